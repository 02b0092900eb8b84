"""Kernel classification of the rocprofv3 summaries (host-only)."""
from cdmf_amd.profiling import kernel_class


def test_kernel_classes():
    t = "void ccd::(anonymous namespace)::tiled_kernel<%s, 16, float, float, %s, %s, %s, %s>(ccd::TiledArgs)"
    assert kernel_class(t % ("ccd::Idx24", "false", "false", "true", "false")) == "sweep_f"
    assert kernel_class(t % ("unsigned int", "false", "false", "true", "false")) == "sweep_f"
    assert kernel_class(t % ("unsigned short", "false", "false", "true", "false")) == "sweep_s"
    assert kernel_class(t % ("ccd::Idx24", "true", "true", "true", "true")) == "fused_f"
    assert kernel_class(t % ("unsigned short", "true", "false", "true", "false")) == "fused_s"
    assert kernel_class(t % ("ccd::Idx24", "true", "false", "false", "false")) == "residual"
    assert kernel_class("void ccd::(anonymous namespace)::tiled_span_kernel<float, float>(ccd::SpanArgs)") == "finalize"
    assert kernel_class("void ccd::(anonymous namespace)::split_finish_kernel<float, float>(ccd::SplitArgs)") == "finalize"
    assert kernel_class("void ccd::(anonymous namespace)::commit_kernel<float>(ccd::CommitArgs)") == "copy"
    assert kernel_class("__amd_rocclr_copyBuffer") is None
