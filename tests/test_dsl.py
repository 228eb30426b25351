"""The Python kernel DSL prints HIP source that compiles against the device library."""
import pytest

import luisacompute_amd as lc
from luisacompute_amd import dsl
import dsl_kernels


@pytest.mark.parametrize("k", [dsl_kernels.trace_kernel, dsl_kernels.control_kernel, dsl_kernels.shade_kernel,
                               dsl_kernels.query_kernel, dsl_kernels.path_tracer,
                               dsl_kernels.volume_kernel])
def test_dsl_kernels_compile(k):
    assert "extern \"C\" __global__" in k.source
    assert lc.compile_shader_source(k.source) > 1000


def test_dsl_errors():
    with pytest.raises(lc.LCError, match="annotation"):
        @dsl.kernel
        def bad(x):
            pass
    with pytest.raises(lc.LCError, match="unknown function"):
        @dsl.kernel
        def bad2(out: dsl.Buffer("float")):
            out.write(0, frobnicate(1.0))


def test_dsl_true_division():
    src = dsl_kernels.division_kernel.source
    assert "true_div(" in src
    assert lc.compile_shader_source(src) > 1000
