"""The kernel AST (luisacompute_amd/lcast.py) and its Python frontend: enumerations numbered as the
reference's op.h / type.h, type layouts, inference and its errors, callables, ray-query statements
and per-argument usage."""
import pytest

import luisacompute_amd as lc
from luisacompute_amd import dsl
from luisacompute_amd import lcast as A
import dsl_kernels


def test_enumerations_follow_the_reference_numbering():
    # positions in include/luisa/ast/op.h's CallOp, type.h's Type::Tag, expression.h, statement.h
    assert (A.CallOp.CUSTOM, A.CallOp.EXTERNAL, A.CallOp.ATOMIC_FETCH_ADD, A.CallOp.BUFFER_WRITE) == (0, 1, 67, 76)
    assert (A.CallOp.MAKE_FLOAT3, A.CallOp.RAY_TRACING_TRACE_CLOSEST, A.CallOp.RAY_TRACING_TRACE_ANY) == (125, 174, 175)
    assert (A.CallOp.RAY_QUERY_COMMITTED_HIT, A.CallOp.SHADER_EXECUTION_REORDER) == (189, 228)
    assert len(A.CallOp) == 229
    assert (A.Type.Tag.VECTOR, A.Type.Tag.STRUCTURE, A.Type.Tag.ACCEL, A.Type.Tag.CUSTOM) == (12, 15, 19, 20)
    assert (A.Expression.Tag.CALL, A.Statement.Tag.RAY_QUERY, A.Variable.Tag.DISPATCH_ID) == (7, 13, 9)
    assert (A.BinaryOp.SHR, A.BinaryOp.NOT_EQUAL, A.UnaryOp.BIT_NOT) == (9, 17, 3)


def test_type_layouts():
    assert A.Type.of("vector<float,3>") is A.FLOAT3 and (A.FLOAT3.size, A.FLOAT3.alignment) == (16, 16)
    assert (A.FLOAT2.size, A.UINT4.size) == (8, 16)
    assert A.Type.of("array<float,3>").size == 12
    assert A.Type.of("buffer<vector<uint,3>>").element is A.UINT3
    # the ray-tracing structures have the reference's sizes (ray.h 32 B, hit.h 24 B)
    assert (A.RAY.size, A.RAY.alignment, A.SURFACE_HIT.size, A.COMMITTED_HIT.size) == (32, 16, 24, 24)
    assert A.SURFACE_HIT.description == "struct<8,uint,uint,vector<float,2>,float>"
    assert A.SURFACE_CANDIDATE is not A.SURFACE_HIT  # one layout, two named structures
    with pytest.raises(TypeError):
        A.Type.of("vector<float,5>")


def test_argument_usage_of_the_trace_kernel():
    k = dsl_kernels.trace_kernel
    R, W = A.Usage.READ, A.Usage.WRITE
    assert [k.argument_usage(i) for i in range(5)] == [R, W, W, R, R]
    k = dsl_kernels.volume_kernel
    assert [k.argument_usage(i) for i in range(3)] == [R, W, R]


def test_path_tracer_callables_share_the_rng_state():
    f = dsl_kernels.path_tracer.function
    names = sorted(c.name for c in f.custom_callables)
    assert names == ["pcg_float", "pcg_next", "to_world"]
    nxt = next(c for c in f.custom_callables if c.name == "pcg_float").custom_callables[0]
    assert nxt.name == "pcg_next" and nxt.return_type is A.UINT
    refs = [v for v in nxt.arguments if v.tag == A.Variable.Tag.REFERENCE]
    assert sorted(v.name for v in refs) == ["inc", "state"] and all(v.type is A.ULONG for v in refs)


def test_query_kernel_lowers_to_a_ray_query_statement():
    f = dsl_kernels.query_kernel.function
    found = []

    def visit(s):
        if isinstance(s, A.RayQueryStmt):
            found.append(s)
        for c in A._stmt_children(s):
            if isinstance(c, A.Statement):
                visit(c)
    visit(f.body)
    assert len(found) == 2  # query_any and query_all branches
    for q in found:
        assert q.query.type in (A.RAY_QUERY_ALL, A.RAY_QUERY_ANY)
        assert q.on_triangle_candidate.statements and q.on_procedural_candidate.statements


def test_inference_and_errors():
    @dsl.kernel
    def mixed(out: dsl.Buffer("float"), n: dsl.uint):
        i = dispatch_id().x
        a = i * 2            # uint (the literal adapts)
        b = a + 0.5          # float
        c = float3(1.0, 2.0, 3.0) * a   # float3 (uint broadcast as float)
        out.write(i, b + c.y + float(n // 3))
    f = mixed.function
    types = {v.name: v.type for v in f.local_variables}
    assert (types["a"], types["b"], types["c"]) == (A.UINT, A.FLOAT, A.FLOAT3)
    assert lc.compile_shader_source(mixed.source) > 1000
    with pytest.raises(lc.LCError, match="line .*assigning vector<float,3>"):
        @dsl.kernel
        def bad(out: dsl.Buffer("float")):
            x = 1.0
            x = float3(1.0, 2.0, 3.0)
    with pytest.raises(lc.LCError, match="unknown name zz"):
        @dsl.kernel
        def bad2(out: dsl.Buffer("float")):
            out.write(0, zz)
    with pytest.raises(lc.LCError, match="cannot assign to parameter"):
        @dsl.kernel
        def bad3(n: dsl.uint):
            n = 3
    with pytest.raises(lc.LCError, match="no lane q"):
        @dsl.kernel
        def bad4(out: dsl.Buffer("float")):
            out.write(0, float3(1.0, 2.0, 3.0).q)


def test_conditional_expression_runs_one_branch():
    @dsl.kernel
    def pick(xs: dsl.Buffer("float"), out: dsl.Buffer("float")):
        i = dispatch_id().x
        out.write(i, xs.read(i + 1) if i + 1 < xs.size() else 0.0)
    f = pick.function
    ifs = [s for s in f.body.statements if isinstance(s, A.IfStmt)]
    assert len(ifs) == 1
    reads = []
    A.walk_expressions(ifs[0].true_branch, lambda e: reads.append(e) if isinstance(e, A.CallExpr) and
                       e.op == A.CallOp.BUFFER_READ else None)
    assert len(reads) == 1  # the read sits in the taken branch only
    assert lc.compile_shader_source(pick.source) > 1000


def test_atomics_and_warp_ops_compile():
    @dsl.kernel(block=(64, 1, 1))
    def append(flags: dsl.Buffer("uint"), queue: dsl.Buffer("uint"), counter: dsl.Buffer("uint")):
        i = dispatch_id().x
        keep = flags.read(i) != 0
        n = warp_active_count_bits(keep)
        base = 0
        if warp_is_first_active_lane():
            base = counter.atomic_fetch_add(0, n)
        base = warp_read_first_active_lane(base)
        if keep:
            queue.write(base + warp_prefix_count_bits(keep), i)
    assert append.argument_usage(2) == A.Usage.READ_WRITE
    assert lc.compile_shader_source(append.source) > 1000
