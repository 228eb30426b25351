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
    """'/' converts both integer operands to float; '//' divides in the operands' type."""
    from luisacompute_amd import lcast as A
    k = dsl_kernels.division_kernel
    writes = []
    A.walk_expressions(k.function.body, lambda e: writes.append(e) if isinstance(e, A.CallExpr) and
                       e.op == A.CallOp.BUFFER_WRITE else None)
    true_div, floor_div = (w.arguments[2] for w in writes)
    assert isinstance(true_div, A.BinaryExpr) and true_div.op == A.BinaryOp.DIV and true_div.type is A.FLOAT
    assert all(isinstance(x, A.CastExpr) and x.type is A.FLOAT for x in (true_div.lhs, true_div.rhs))
    assert isinstance(floor_div, A.BinaryExpr) and floor_div.op == A.BinaryOp.DIV and floor_div.type is A.UINT
    assert lc.compile_shader_source(k.source) > 1000


@pytest.mark.parametrize("name", __import__("luisacompute_amd.programs", fromlist=["x"]).WavefrontPathTracer.KERNELS)
def test_wavefront_kernels_compile(name):
    from luisacompute_amd import programs
    k = getattr(programs, name)
    assert "extern \"C\" __global__" in k.source
    assert lc.compile_shader_source(k.source) > 1000


def test_wavefront_shade_argument_usage():
    """The shade kernel's queues and path state carry the usage its AST gives them."""
    from luisacompute_amd import lcast as A, programs
    k = programs.wf_shade
    names = list(k.params)
    use = lambda n: k.argument_usage(names.index(n))
    assert use("rays_in") == A.Usage.READ and use("hits") == A.Usage.READ
    assert use("rays_out") == A.Usage.WRITE and use("shadow_rays") == A.Usage.WRITE
    assert use("counters") == A.Usage.READ_WRITE and use("throughput") == A.Usage.READ_WRITE
    assert use("rng") == A.Usage.READ_WRITE and use("radiance") == A.Usage.READ_WRITE
