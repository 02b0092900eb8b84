"""What the learned warm start gains over plain (zero-start) CCD++ fold-in,
measured end to end on the GPU at a reduced scale of the held-out-users
study (30,000 users x 3,000 items, 2.1M ratings, k = 16): factorize the
kept 90 %, train the MLP on their (features, w_i), fold in the rest.

  * planted rank-8 ratings: the warm start leaves a smaller fold-in
    objective gap (to the value after 30 sweeps) than the zero start after
    every number of sweeps from 1 to 6, and reaches a 2e-4 gap in at most as
    many sweeps;
  * independent ratings (nothing to learn): the warm start's test RMSE
    with no sweep at all is already at or below the zero start's after 6
    sweeps (the zero start fits the rating noise), i.e. <= 1/2 the sweeps.
The full-size numbers are produced by bench.py --foldin.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(kind):
    from cdmf_amd import data as D
    from cdmf_amd import foldin as F
    from cdmf_amd import solver
    if solver.device_count() < 1:
        pytest.fail("no HIP device visible: GPU tests need the MI355X")
    ds = D.gen_powerlaw(m=30000, n=3000, nnz_train=2100000, nnz_test=60000, max_row_degree=3000, rating_kind=kind,
                        seed=21)
    return F.experiment(ds, k=16, lam=0.05, outer=6, inner=3, sweeps=6, epochs=15, converge=30, obj_rtol=2e-4,
                        log=lambda s: print(s, flush=True))


def test_warm_start_shrinks_the_objective_gap():
    r = _run(3)
    gp = np.array(r.plain_obj) / r.obj_star - 1
    gw = np.array(r.warm_obj) / r.obj_star - 1
    print("gaps plain", gp, "warm", gw, flush=True)
    assert np.all(gw[1:] < gp[1:])
    assert r.warm_sweeps_obj >= 0 and (r.plain_sweeps_obj < 0 or r.warm_sweeps_obj <= r.plain_sweeps_obj)


def test_warm_start_on_independent_ratings():
    r = _run(1)
    print("plain", r.plain_rmse, "warm", r.warm_rmse, flush=True)
    assert 0 <= r.warm_sweeps <= r.plain_sweeps // 2
