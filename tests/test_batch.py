"""Batched solver (ccd_batch_*): many problems in one persistent launch.

The batched kernel runs each problem's loop with the same tiled passes, in
the same order, as ccd_run with the tiled sweeps, so the bar here is bitwise
equality of W, H and both residual copies with a separate single-problem
run of the same sigma -- which the parity tests in test_gpu_parity.py tie
to the oracle -- plus one direct check against the oracle.  Cases mix
16-bit and 32-bit gathered indices, both orders, empty rows and columns,
k from 3 to 16, 1 to 5 inner iterations, one-block and many-block teams,
runs that continue earlier runs, graph replays and the autotuner's plan.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver():
    from cdmf_amd import solver
    if solver.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need the MI355X")
    return solver


def _say(msg):
    print(msg, flush=True)


def _holey(ds):
    from cdmf_amd import datadir
    rows = np.repeat(np.arange(ds.m), np.diff(ds.csr_ptr.astype(np.int64)))
    keep = (rows % 3 != 0) & (ds.csr_idx % 5 != 0)
    return datadir.from_triplets(ds.m, ds.n, (rows[keep], ds.csr_idx[keep], ds.csr_val[keep]),
                                 (ds.test_row, ds.test_col, ds.test_val))


def _cases():
    """(name, dataset, k, lambda, outer, inner, order, W0, H0)"""
    from cdmf_amd import data as D
    from cdmf_amd.solver import ORDER_H_FIRST, ORDER_W_FIRST
    out = []
    big = D.gen_powerlaw(m=70000, n=3000, nnz_train=400000, nnz_test=20000, max_row_degree=3000, rating_kind=1,
                         seed=11)  # 70,000 gathered rows: 32-bit indices in the CSC sweeps
    out.append(("powerlaw70k", big, 8, 0.05, 2, 3, ORDER_H_FIRST, D.initial_col(8, big.m), np.zeros((8, big.n), np.float32)))
    small = D.small()
    W0, H0 = D.small_init(small, 10)
    out.append(("small_wfirst", small, 10, 0.05, 3, 5, ORDER_W_FIRST, W0, H0))
    pl = D.gen_powerlaw(m=5000, n=800, nnz_train=100000, nnz_test=5000, max_row_degree=800, rating_kind=0, seed=12)
    out.append(("powerlaw5k", pl, 16, 0.1, 2, 1, ORDER_H_FIRST, D.initial_col(16, pl.m), np.zeros((16, pl.n), np.float32)))
    tiny = D.gen_uniform(300, 200, 3000, 300, seed=13)
    out.append(("tiny", tiny, 4, 0.01, 3, 2, ORDER_H_FIRST, D.initial_col(4, tiny.m), np.zeros((4, tiny.n), np.float32)))
    holey = _holey(small)
    out.append(("holey", holey, 3, 0.05, 2, 2, ORDER_H_FIRST, D.initial_col(3, holey.m), np.zeros((3, holey.n), np.float32)))
    return out


def _single(S, ds, k, lam, outer, inner, order, W0, H0, sigma, runs=1):
    # the batched kernel runs on plain tile layouts: the single run must too
    with S.Solver(0) as s:
        s.set_plan(sigma=sigma, split_rows=0)
        s.upload(ds)
        s.set_factors(W0, H0)
        for _ in range(runs):
            s.run(k, lam, outer, inner, variant=S.VARIANT_TILED, order=order, zero_first=False)
        W, H = s.factors()
        a, b = s.residual()
        return W, H, a, b


def _batch_results(solvers):
    out = []
    for s in solvers:
        W, H = s.factors()
        a, b = s.residual()
        out.append((W, H, a, b))
    return out


def _same(got, want, what):
    for x, y, name in zip(got, want, ("W", "H", "residual CSR", "residual CSC")):
        assert np.array_equal(x, y), "%s: %s differs from the single-problem run (max |diff| %.3g)" % (
            what, name, float(np.max(np.abs(x.astype(np.float64) - y.astype(np.float64)))))


def _build(S, cases, sigma=16, tune=False):
    b = S.Batch(0)
    solvers = []
    for name, ds, k, lam, outer, inner, order, W0, H0 in cases:
        s = S.Solver(0)
        s.upload(ds)
        s.set_factors(W0, H0)
        b.add(s, k, lam, outer, inner, order=order, zero_first=False)
        solvers.append(s)
    if tune:
        ms = b.tune()
        _say("tuned plan %s (trial %.3f ms)" % (b.plan(), ms))
    else:
        b.set_plan(sigma=sigma)
    return b, solvers


@pytest.mark.parametrize("sigma", [16, 8])
def test_batch_bitwise_equals_single_runs(sigma):
    S = _solver()
    cases = _cases()
    _say("batch of %d problems, sigma %d" % (len(cases), sigma))
    b, solvers = _build(S, cases, sigma)
    t = b.run()
    _say("batch run %.3f ms, teams %s" % (1e3 * t, b.teams()))
    got = _batch_results(solvers)
    for (name, ds, k, lam, outer, inner, order, W0, H0), g in zip(cases, got):
        _say("compare %s" % name)
        _same(g, _single(S, ds, k, lam, outer, inner, order, W0, H0, sigma), name)
    b.close()


def test_batch_edge_shapes():
    """One row over three tiles, one column, a 3 x 2 matrix in one padded
    tile and a full 40 x 40 block at k = 1, batched with one problem of the
    usual kind: bitwise equal to their single-problem runs."""
    from cdmf_amd import data as D
    from cdmf_amd.solver import ORDER_H_FIRST
    from tests.helpers import DEGENERATE, degenerate
    S = _solver()
    cases = [_cases()[3]]
    for name in DEGENERATE:
        ds = degenerate(name)
        k = 1 if name == "dense" else 4
        cases.append((name, ds, k, 0.05, 3, 3, ORDER_H_FIRST, D.initial_col(k, ds.m), np.zeros((k, ds.n), np.float32)))
    b, solvers = _build(S, cases, 16)
    b.run()
    _say("teams %s" % b.teams())
    for (name, ds, k, lam, outer, inner, order, W0, H0), g in zip(cases, _batch_results(solvers)):
        _say("compare %s" % name)
        _same(g, _single(S, ds, k, lam, outer, inner, order, W0, H0, 16), name)
    b.close()


@pytest.mark.parametrize("deal", [1, 0xffffffff])
def test_batch_tile_deal_bitwise(deal):
    """ccd_batch_plan.deal only changes which wave runs a tile: with every
    sweep dealt through the blocks' LDS counters (1 tile per wave) and with
    none dealt, every problem's W, H and residual equal the single-problem
    runs bit for bit (a continued run included: its first rank's add-back
    and first sweep meet at a team barrier when the sweep is dealt)."""
    S = _solver()
    cases = _cases()
    b, solvers = _build(S, cases, 16)
    b.set_plan(sigma=16, deal=deal)
    assert b.plan()["deal"] == deal
    b.run()
    b.run(use_graph=True)  # continued: two runs
    got = _batch_results(solvers)
    for (name, ds, k, lam, outer, inner, order, W0, H0), g in zip(cases, got):
        _same(g, _single(S, ds, k, lam, outer, inner, order, W0, H0, 16, runs=2), "%s deal %d" % (name, deal))
    b.close()


def test_batch_small_grid_stages_in_lds():
    """16 blocks for 5 problems (two per XCD): every team streams far more
    elements per block than it would stage, so each pass stages the gathered
    vector's hot prefix (or the pairs) in LDS; results stay bit-identical to
    single-problem runs."""
    S = _solver()
    cases = _cases()
    b, solvers = _build(S, cases, 16)
    b.set_plan(sigma=16, blocks=16)
    t = b.run()
    _say("batch run on 16 blocks %.3f ms, teams %s" % (1e3 * t, b.teams()))
    for (name, ds, k, lam, outer, inner, order, W0, H0), g in zip(cases, _batch_results(solvers)):
        _say("compare %s" % name)
        _same(g, _single(S, ds, k, lam, outer, inner, order, W0, H0, 16), name)
    b.close()


def test_batch_continued_and_graph_replay():
    """Second batch run continues the first (add-back from rank 0); graph
    replays after ccd_set_factors repeat the first run exactly."""
    S = _solver()
    cases = _cases()[1:4]
    b, solvers = _build(S, cases, 16)
    _say("graph run 1")
    b.run(use_graph=True)
    first = _batch_results(solvers)
    _say("graph run 2 (continued: recaptured)")
    b.run(use_graph=True)
    second = _batch_results(solvers)
    for (name, ds, k, lam, outer, inner, order, W0, H0), g1, g2 in zip(cases, first, second):
        _say("compare %s" % name)
        _same(g1, _single(S, ds, k, lam, outer, inner, order, W0, H0, 16), name + " run 1")
        _same(g2, _single(S, ds, k, lam, outer, inner, order, W0, H0, 16, runs=2), name + " run 2")
    for rep in range(3):  # reset, replay the first run's graph, compare
        for s, c in zip(solvers, cases):
            s.set_factors(c[7], c[8])
        _say("graph replay %d" % rep)
        b.run(use_graph=True)
        for (name, *_), g, g1 in zip(cases, _batch_results(solvers), first):
            _same(g, g1, "%s replay %d" % (name, rep))
    b.close()


def test_batch_tuned_plan_and_oracle(oracle_lib):
    S = _solver()
    from cdmf_amd import data as D
    cases = _cases()[1:3]
    b, solvers = _build(S, cases, tune=True)
    plan = b.plan()
    assert plan["sigma"] in (8, 16)
    b.run()
    got = _batch_results(solvers)
    for (name, ds, k, lam, outer, inner, order, W0, H0), g in zip(cases, got):
        _same(g, _single(S, ds, k, lam, outer, inner, order, W0, H0, plan["sigma"]), name + " (tuned)")
    # the small case against the oracle, within the FP32 bar of the parity tests
    name, ds, k, lam, outer, inner, order, W0, H0 = cases[0]
    W, H, rm, _, _ = oracle_lib.ccd(ds, W0, H0, k, float(np.float32(lam)), outer, inner, order=order,
                                    zero_first=False)
    Wg, Hg = got[0][0], got[0][1]
    for x, ref, what in ((Wg, W, "W"), (Hg, H, "H")):
        frob = np.linalg.norm(x.astype(np.float64) - ref) / np.linalg.norm(ref)
        assert frob <= 1e-3, "%s vs oracle: %.3g" % (what, frob)
    assert abs(solvers[0].rmse() - rm[-1]) <= 1e-4
    b.close()


def test_batch_graph_after_layout_rebuild():
    """A captured batch run bakes in each context's buffer addresses.  A
    ccd_run with split rows in between rebuilds the tile layout of a problem
    whose gathered dimension (70,000 rows) exceeds the LDS prefix, freeing
    the plain layout the capture used; the next graph run must recapture.
    The same sequence with eager launches is the reference, bit for bit."""
    S = _solver()
    cases = _cases()[:3]
    out = {}
    for graph in (False, True):
        b, solvers = _build(S, cases, 16)
        _say("graph=%s: batch run 1" % graph)
        b.run(use_graph=graph)
        name, ds, k, lam, outer, inner, order, W0, H0 = cases[0]
        _say("graph=%s: ccd_run with split rows on %s" % (graph, name))
        solvers[0].set_plan(split_rows=1)
        solvers[0].run(k, lam, 1, inner, variant=S.VARIANT_TILED, order=order, zero_first=False)
        _say("graph=%s: batch run 2" % graph)
        b.run(use_graph=graph)
        out[graph] = _batch_results(solvers)
        b.close()
        for s in solvers:
            s.close()
    for (name, *_), e, g in zip(cases, out[False], out[True]):
        assert all(np.all(np.isfinite(x)) for x in e), name
        _same(g, e, name + " (graph after a layout rebuild)")


@pytest.mark.parametrize("solo", [1, 2])
def test_batch_solo_problems_bitwise(solo):
    """ccd_batch_plan.solo: the largest problems are solved alone on the
    whole GPU before the persistent launch takes the rest; every problem's
    result is still bit-identical to its own single-problem run, for the
    eager and the graph-replayed launch."""
    S = _solver()
    cases = _cases()
    for graph in (False, True):
        b, solvers = _build(S, cases, 16)
        b.set_plan(sigma=16, solo=solo)
        assert b.plan()["solo"] == solo
        t = b.run(use_graph=graph)
        teams = b.teams()
        _say("solo=%d graph=%s: %.3f ms, teams %s" % (solo, graph, 1e3 * t, teams))
        assert sum(1 for x in teams if x == 0) == solo
        for (name, ds, k, lam, outer, inner, order, W0, H0), g in zip(cases, _batch_results(solvers)):
            _same(g, _single(S, ds, k, lam, outer, inner, order, W0, H0, 16), "%s (solo=%d, graph=%s)" % (
                name, solo, graph))
        b.close()
