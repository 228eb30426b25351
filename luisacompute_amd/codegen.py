"""AST -> HIP source: prints a kernel of the lcast AST against the device library of run-time
shaders (lc_hip_device.h), for hiprtc. The counterpart of the reference's CUDA codegen
(src/backends/cuda/cuda_codegen_ast.cpp:1291-1490 emits the kernel, its argument struct and the
builtins; :802-1115 lowers the CallExprs); the output is compiled by Device.create_shader.

Layout of the printed source:
  #include "lc_hip_device.h" + the lcdsl helpers (vector math, swizzles, casts, ray-query glue)
  forward declarations and definitions of the callables, deepest first
  struct Params_<kernel> { alignas(16) <argument>...; alignas(16) uint4 ls_kid; }
  extern "C" __global__ void __launch_bounds__(<block>) <kernel>(const Params_<kernel> lcp_)
Every local is declared, zero-initialised, at the top of its function (the AST keeps locals per
function); kernel arguments live in lcp_, callable arguments are a_<name> (REFERENCE arguments by
reference), locals v_<name>. Expressions print fully parenthesised in AST order, so the compiled
arithmetic is the AST's (no reassociation: shaders compile with -ffp-contract=off).
"""
import numpy as np

from . import lcast as A
from .lcast import CallOp, Type

PRELUDE = r'''
#include "lc_hip_device.h"
namespace lcdsl {
using namespace lchip;
template<class T> __device__ __forceinline__ T lc_min(T a, T b) { return b < a ? b : a; }
template<class T> __device__ __forceinline__ T lc_max(T a, T b) { return a < b ? b : a; }
__device__ __forceinline__ float lc_min(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ float lc_max(float a, float b) { return fmaxf(a, b); }
template<class T> __device__ __forceinline__ T lc_abs(T a) { return a < T(0) ? -a : a; }
__device__ __forceinline__ float lc_abs(float a) { return fabsf(a); }
__device__ __forceinline__ uint32_t lc_abs(uint32_t a) { return a; }
__device__ __forceinline__ uint64_t lc_abs(uint64_t a) { return a; }
template<class T> __device__ __forceinline__ T lc_clamp(T x, T lo, T hi) { return lc_min(lc_max(x, lo), hi); }
__device__ __forceinline__ float lc_lerp(float a, float b, float t) { return a + t * (b - a); }
__device__ __forceinline__ float lc_saturate(float x) { return lc_clamp(x, 0.0f, 1.0f); }
__device__ __forceinline__ float lc_step(float e, float x) { return x < e ? 0.0f : 1.0f; }
__device__ __forceinline__ float lc_fract(float x) { return x - floorf(x); }
__device__ __forceinline__ float lc_rsqrt(float x) { return 1.0f / sqrtf(x); }
__device__ __forceinline__ float lc_exp10(float x) { return powf(10.0f, x); }
__device__ __forceinline__ float3 v3f(rtx::V3 v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ rtx::V3 fv3(float3 v) { return rtx::v3(v.x, v.y, v.z); }
__device__ __forceinline__ float dot(float3 a, float3 b) { return rtx::dot(fv3(a), fv3(b)); }
__device__ __forceinline__ float dot(float2 a, float2 b) { return a.x * b.x + a.y * b.y; }
__device__ __forceinline__ float3 cross(float3 a, float3 b) { return v3f(rtx::cross(fv3(a), fv3(b))); }
__device__ __forceinline__ float length(float3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ float length(float2 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ float length_squared(float3 a) { return dot(a, a); }
__device__ __forceinline__ float3 normalize(float3 a) { return v3f(rtx::normalize(fv3(a))); }
__device__ __forceinline__ float3 lc_min(float3 a, float3 b) { return make_float3(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
__device__ __forceinline__ float3 lc_max(float3 a, float3 b) { return make_float3(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
__device__ __forceinline__ float3 lc_abs(float3 a) { return make_float3(fabsf(a.x), fabsf(a.y), fabsf(a.z)); }
__device__ __forceinline__ float reduce_sum(float3 a) { return a.x + a.y + a.z; }
__device__ __forceinline__ float reduce_max(float3 a) { return fmaxf(fmaxf(a.x, a.y), a.z); }
__device__ __forceinline__ float reduce_min(float3 a) { return fminf(fminf(a.x, a.y), a.z); }
__device__ __forceinline__ float3 offset_ray_origin(float3 p, float3 n) { return v3f(rtx::offset_ray_origin(fv3(p), fv3(n))); }
__device__ __forceinline__ float3 triangle_interpolate(float2 b, float3 x, float3 y, float3 z) {
    return v3f(rtx::triangle_interpolate(b.x, b.y, fv3(x), fv3(y), fv3(z)));
}
__device__ __forceinline__ dev::Ray make_ray(float3 o, float3 d, float tmin, float tmax) {
    return dev::make_ray(fv3(o), fv3(d), tmin, tmax);
}
__device__ __forceinline__ float3 ray_origin(const dev::Ray &r) { return make_float3(r.origin[0], r.origin[1], r.origin[2]); }
__device__ __forceinline__ float3 ray_direction(const dev::Ray &r) {
    return make_float3(r.direction[0], r.direction[1], r.direction[2]);
}
__device__ __forceinline__ float2 hit_bary(const dev::SurfaceHit &h) { return make_float2(h.bary[0], h.bary[1]); }
__device__ __forceinline__ bool hit_miss(const dev::SurfaceHit &h) { return h.inst == 0xFFFFFFFFu; }
// lane i of a vector, swizzles
template<class V> __device__ __forceinline__ auto lane(const V &v, uint32_t i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }
template<class V> __device__ __forceinline__ auto lane3(const V &v, uint32_t i) { return i == 0 ? v.x : (i == 1 ? v.y : v.z); }
template<class V> __device__ __forceinline__ auto lane2(const V &v, uint32_t i) { return i == 0 ? v.x : v.y; }
template<class V> __device__ __forceinline__ auto &lane_ref(V &v, uint32_t i) { return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w)); }
// warp (wave64) operations
__device__ __forceinline__ uint32_t warp_lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t warp_active_count_bits(bool b) { return (uint32_t)__popcll(__ballot(b)); }
__device__ __forceinline__ uint32_t warp_prefix_count_bits(bool b) {
    return (uint32_t)__popcll(__ballot(b) & ((1ull << __lane_id()) - 1ull));
}
__device__ __forceinline__ bool warp_is_first_active_lane() { return __lane_id() == (uint32_t)(__ffsll((long long)__ballot(1)) - 1); }
template<class T> __device__ __forceinline__ T warp_read_first_active_lane(T x) {
    return __shfl(x, __ffsll((long long)__ballot(1)) - 1, 64);
}
template<class T> __device__ __forceinline__ T warp_read_lane(T x, uint32_t i) { return __shfl(x, (int)i, 64); }
// ray queries: the query object, the candidates its callbacks see
struct RayQuery { dev::Accel accel; dev::Ray ray; uint32_t mask; bool any; rq::CommittedHit hit; };
struct SurfaceCandidate { uint32_t inst, prim; float2 bary; float t; };
struct ProceduralCandidate { uint32_t inst, prim; float3 origin, direction, aabb_min, aabb_max; float t_max; };
} // namespace lcdsl
'''

_SCALAR_NAMES = {"bool": "bool", "byte": "int8_t", "ubyte": "uint8_t", "short": "int16_t", "ushort": "uint16_t",
                 "int": "int32_t", "uint": "uint32_t", "long": "int64_t", "ulong": "uint64_t", "float": "float",
                 "double": "double"}
_VECTOR_PREFIX = {"int": "int", "uint": "uint", "float": "float", "long": "longlong", "ulong": "ulonglong",
                  "double": "double"}
_STRUCT_NAMES = {"Ray": "lchip::dev::Ray", "SurfaceHit": "lchip::dev::SurfaceHit",
                 "CommittedHit": "lchip::rq::CommittedHit", "SurfaceCandidate": "lcdsl::SurfaceCandidate",
                 "ProceduralCandidate": "lcdsl::ProceduralCandidate"}

# CallOps printed as a function call on their arguments
_FUNCS = {
    CallOp.ABS: "lcdsl::lc_abs", CallOp.MIN: "lcdsl::lc_min", CallOp.MAX: "lcdsl::lc_max",
    CallOp.CLAMP: "lcdsl::lc_clamp", CallOp.SATURATE: "lcdsl::lc_saturate", CallOp.LERP: "lcdsl::lc_lerp",
    CallOp.STEP: "lcdsl::lc_step", CallOp.CLZ: "__clz", CallOp.CTZ: "__builtin_ctz", CallOp.POPCOUNT: "__popc",
    CallOp.REVERSE: "__brev", CallOp.ISINF: "isinf", CallOp.ISNAN: "isnan", CallOp.ACOS: "acosf",
    CallOp.ACOSH: "acoshf", CallOp.ASIN: "asinf", CallOp.ASINH: "asinhf", CallOp.ATAN: "atanf",
    CallOp.ATAN2: "atan2f", CallOp.ATANH: "atanhf", CallOp.COS: "cosf", CallOp.COSH: "coshf", CallOp.SIN: "sinf",
    CallOp.SINH: "sinhf", CallOp.TAN: "tanf", CallOp.TANH: "tanhf", CallOp.EXP: "expf", CallOp.EXP2: "exp2f",
    CallOp.EXP10: "lcdsl::lc_exp10", CallOp.LOG: "logf", CallOp.LOG2: "log2f", CallOp.LOG10: "log10f",
    CallOp.POW: "powf", CallOp.SQRT: "sqrtf", CallOp.RSQRT: "lcdsl::lc_rsqrt", CallOp.CEIL: "ceilf",
    CallOp.FLOOR: "floorf", CallOp.FRACT: "lcdsl::lc_fract", CallOp.TRUNC: "truncf", CallOp.ROUND: "roundf",
    CallOp.FMA: "fmaf", CallOp.COPYSIGN: "copysignf", CallOp.CROSS: "lcdsl::cross", CallOp.DOT: "lcdsl::dot",
    CallOp.LENGTH: "lcdsl::length", CallOp.LENGTH_SQUARED: "lcdsl::length_squared",
    CallOp.NORMALIZE: "lcdsl::normalize", CallOp.REDUCE_SUM: "lcdsl::reduce_sum",
    CallOp.REDUCE_MIN: "lcdsl::reduce_min", CallOp.REDUCE_MAX: "lcdsl::reduce_max",
    CallOp.WARP_ACTIVE_COUNT_BITS: "lcdsl::warp_active_count_bits",
    CallOp.WARP_PREFIX_COUNT_BITS: "lcdsl::warp_prefix_count_bits",
    CallOp.WARP_IS_FIRST_ACTIVE_LANE: "lcdsl::warp_is_first_active_lane",
    CallOp.WARP_READ_FIRST_ACTIVE_LANE: "lcdsl::warp_read_first_active_lane",
    CallOp.WARP_READ_LANE: "lcdsl::warp_read_lane",
    CallOp.RAY_TRACING_TRACE_CLOSEST: "lchip::dev::trace_closest", CallOp.RAY_TRACING_TRACE_ANY: "lchip::dev::trace_any",
    CallOp.RAY_TRACING_INSTANCE_USER_ID: "lchip::dev::instance_user_id",
    CallOp.RAY_TRACING_INSTANCE_VISIBILITY_MASK: "lchip::dev::instance_visibility_mask",
    CallOp.RAY_TRACING_SET_INSTANCE_USER_ID: "lchip::dev::set_instance_user_id",
    CallOp.RAY_TRACING_SET_INSTANCE_VISIBILITY: "lchip::dev::set_instance_visibility_mask",
    CallOp.BINDLESS_TEXTURE2D_SIZE: "lchip::dev::bindless_tex2d_size",
    CallOp.BINDLESS_BUFFER_SIZE: "lchip::dev::bindless_buffer_size",
}
_ATOMICS = {CallOp.ATOMIC_EXCHANGE: "atomicExch", CallOp.ATOMIC_FETCH_ADD: "atomicAdd",
            CallOp.ATOMIC_FETCH_SUB: "atomicSub", CallOp.ATOMIC_FETCH_AND: "atomicAnd",
            CallOp.ATOMIC_FETCH_OR: "atomicOr", CallOp.ATOMIC_FETCH_XOR: "atomicXor",
            CallOp.ATOMIC_FETCH_MIN: "atomicMin", CallOp.ATOMIC_FETCH_MAX: "atomicMax"}
_UNARY = {A.UnaryOp.PLUS: "+", A.UnaryOp.MINUS: "-", A.UnaryOp.NOT: "!", A.UnaryOp.BIT_NOT: "~"}
_BINARY = {A.BinaryOp.ADD: "+", A.BinaryOp.SUB: "-", A.BinaryOp.MUL: "*", A.BinaryOp.DIV: "/", A.BinaryOp.MOD: "%",
           A.BinaryOp.BIT_AND: "&", A.BinaryOp.BIT_OR: "|", A.BinaryOp.BIT_XOR: "^", A.BinaryOp.SHL: "<<",
           A.BinaryOp.SHR: ">>", A.BinaryOp.AND: "&&", A.BinaryOp.OR: "||", A.BinaryOp.LESS: "<",
           A.BinaryOp.GREATER: ">", A.BinaryOp.LESS_EQUAL: "<=", A.BinaryOp.GREATER_EQUAL: ">=",
           A.BinaryOp.EQUAL: "==", A.BinaryOp.NOT_EQUAL: "!="}
_LANES = "xyzw"


class CodegenError(Exception):
    pass


def type_name(t):
    """C++ spelling of an AST type in the device library."""
    if t is None:
        return "void"
    if t.is_scalar:
        return _SCALAR_NAMES[t.description]
    if t.is_vector:
        p = _VECTOR_PREFIX.get(t.element.description)
        if p is None:
            raise CodegenError(f"no device type for {t.description}")
        return f"{p}{t.dimension}"
    if t.tag == Type.Tag.ARRAY:
        return f"lchip::dev::Array<{type_name(t.element)}, {t.dimension}>"
    if t.is_structure:
        name = getattr(t, "struct_name", None)
        if name not in _STRUCT_NAMES:
            raise CodegenError(f"no device type for {t.description}")
        return _STRUCT_NAMES[name]
    if t.tag == Type.Tag.BUFFER:
        return f"lchip::dev::Buffer<{type_name(t.element)}>"
    if t.tag == Type.Tag.TEXTURE:
        return "lchip::dev::Image" if t.dimension == 2 else "lchip::dev::Volume"
    if t.tag == Type.Tag.BINDLESS_ARRAY:
        return "lchip::dev::BindlessArray"
    if t.tag == Type.Tag.ACCEL:
        return "lchip::dev::Accel"
    if t is A.RAY_QUERY_ALL or t is A.RAY_QUERY_ANY:
        return "lcdsl::RayQuery"
    raise CodegenError(f"no device type for {t.description}")


def literal(t, v):
    if t.is_bool:
        return "true" if v else "false"
    if t.is_float:
        f = np.float32(v)
        if np.isnan(f):
            return "__builtin_nanf(\"\")"
        if np.isinf(f):
            return "__builtin_inff()" if f > 0 else "(-__builtin_inff())"
        s = repr(float(f))
        if "e" not in s and "." not in s and "inf" not in s:
            s += ".0"
        return f"{s}f" if f >= 0 else f"({s}f)"
    v = int(v)
    suffix = {"uint": "u", "long": "ll", "ulong": "ull"}.get(t.description, "")
    if t.description == "int" and v == -2 ** 31:
        return "(-2147483647 - 1)"
    return f"{v}{suffix}" if v >= 0 else f"({v}{suffix})"


class _Printer:
    def __init__(self, fn, callable_names, kernel):
        self.fn, self.callable_names, self.kernel = fn, callable_names, kernel
        self.lines = []
        self.depth = 1
        self.names = {}
        self.query_depth = 0
        used = set()
        for v in fn.arguments + fn.local_variables + fn.shared_variables:
            base = (v.name or f"t{v.uid}").replace(".", "_")
            prefix = "a_" if v in fn.arguments and not kernel else "v_"
            n = prefix + base
            while n in used:
                n = f"{prefix}{base}_{v.uid}"
            used.add(n)
            self.names[v.uid] = n

    def line(self, s):
        self.lines.append("    " * self.depth + s)

    # -- variables
    def var(self, v):
        if v.tag == A.Variable.Tag.DISPATCH_ID:
            return "lchip::dev::dispatch_id()"
        if v.tag == A.Variable.Tag.DISPATCH_SIZE:
            return "make_uint3(lcp_.ls_kid.x, lcp_.ls_kid.y, lcp_.ls_kid.z)"
        if v.tag == A.Variable.Tag.THREAD_ID:
            return "make_uint3(threadIdx.x, threadIdx.y, threadIdx.z)"
        if v.tag == A.Variable.Tag.BLOCK_ID:
            return "make_uint3(blockIdx.x, blockIdx.y, blockIdx.z)"
        if v.tag == A.Variable.Tag.KERNEL_ID:
            return "lchip::dev::kernel_id(lcp_.ls_kid)"
        if v.tag == A.Variable.Tag.WARP_LANE_ID:
            return "lcdsl::warp_lane_id()"
        if v.tag == A.Variable.Tag.WARP_LANE_COUNT:
            return "64u"
        if self.kernel and v in self.fn.arguments:
            return f"lcp_.{v.name}"
        return self.names[v.uid]

    # -- expressions
    def expr(self, e):
        if isinstance(e, A.LiteralExpr):
            return literal(e.type, e.value)
        if isinstance(e, A.RefExpr):
            return self.var(e.variable)
        if isinstance(e, A.UnaryExpr):
            return f"({_UNARY[e.op]}{self.expr(e.operand)})"
        if isinstance(e, A.BinaryExpr):
            return f"({self.expr(e.lhs)} {_BINARY[e.op]} {self.expr(e.rhs)})"
        if isinstance(e, A.MemberExpr):
            s = self.expr(e.self)
            if e.is_swizzle:
                if len(e.swizzle) == 1:
                    return f"{s}.{_LANES[e.swizzle[0]]}"
                return f"[&](const auto &lcs_) {{ return make_{type_name(e.type)}(" + \
                       ", ".join(f"lcs_.{_LANES[i]}" for i in e.swizzle) + f"); }}({s})"
            st = e.self.type
            name = st.member_names[e.member]
            if getattr(st, "struct_name", None) == "SurfaceHit" and name == "bary":
                return f"lcdsl::hit_bary({s})"
            return f"{s}.{name}"
        if isinstance(e, A.AccessExpr):
            r = self.expr(e.range)
            if e.range.type.is_vector:
                return f"lcdsl::lane({r}, {self.expr(e.index)})"
            return f"{r}[{self.expr(e.index)}]"
        if isinstance(e, A.CastExpr):
            inner = self.expr(e.expression)
            if e.op == A.CastOp.BITWISE:
                return f"__builtin_bit_cast({type_name(e.type)}, {inner})"
            if e.type.is_vector:
                src = e.expression.type
                lanes = ", ".join(f"static_cast<{type_name(e.type.element)}>(lcc_v_.{_LANES[i]})"
                                  for i in range(src.dimension))
                return f"[&](const auto &lcc_v_) {{ return make_{type_name(e.type)}({lanes}); }}({inner})"
            return f"static_cast<{type_name(e.type)}>({inner})"
        if isinstance(e, A.CallExpr):
            return self.call(e)
        raise CodegenError(f"cannot print {type(e).__name__}")

    def call(self, e):
        a = [self.expr(x) for x in e.arguments]
        op = e.op
        if op == CallOp.CUSTOM:
            return f"{self.callable_names[id(e.custom)]}({', '.join(a)})"
        if op == CallOp.EXTERNAL:
            return f"{e.external.name}({', '.join(a)})"
        if op in _FUNCS:
            return f"{_FUNCS[op]}({', '.join(a)})"
        if op == CallOp.SELECT:  # select(false_value, true_value, condition)
            return f"({a[2]} ? {a[1]} : {a[0]})"
        if CallOp.MAKE_BOOL2 <= op <= CallOp.MAKE_UBYTE4:
            return f"make_{type_name(e.type)}({', '.join(a)})"
        if op == CallOp.SYNCHRONIZE_BLOCK:
            return "__syncthreads()"
        if op == CallOp.BUFFER_READ:
            return f"{a[0]}[{a[1]}]"
        if op == CallOp.BUFFER_WRITE:
            return f"({a[0]}[{a[1]}] = {a[2]})"
        if op == CallOp.BUFFER_SIZE:
            return f"{a[0]}.size()"
        if op in _ATOMICS:
            return f"{_ATOMICS[op]}(&{a[0]}[{a[1]}], {a[2]})"
        if op == CallOp.ATOMIC_COMPARE_EXCHANGE:
            return f"atomicCAS(&{a[0]}[{a[1]}], {a[2]}, {a[3]})"
        if op in (CallOp.TEXTURE_READ, CallOp.TEXTURE_WRITE):
            tex = e.arguments[0].type
            kind = {"float": "float", "uint": "uint", "int": "int"}[tex.element.description]
            xyz = "xy" if tex.dimension == 2 else "xyz"
            name = "image" if tex.dimension == 2 else "volume"
            coords = ", ".join(f"lct_.{c}" for c in xyz)
            if op == CallOp.TEXTURE_READ:
                return f"[&](const auto &lct_) {{ return lchip::dev::{name}_read_{kind}({a[0]}, {coords}); }}({a[1]})"
            return (f"[&](const auto &lct_) {{ lchip::dev::{name}_write_{kind}({a[0]}, {coords}, {a[2]}); }}({a[1]})")
        if op == CallOp.BINDLESS_TEXTURE2D_SAMPLE:
            return f"[&](const float2 lcu_) {{ return lchip::dev::bindless_tex2d_sample({a[0]}, {a[1]}, lcu_.x, lcu_.y); }}({a[2]})"
        if op == CallOp.BINDLESS_TEXTURE2D_READ:
            return f"[&](const uint2 lcu_) {{ return lchip::dev::bindless_tex2d_read({a[0]}, {a[1]}, lcu_.x, lcu_.y); }}({a[2]})"
        if op == CallOp.BINDLESS_BUFFER_READ:
            return f"lchip::dev::bindless_buffer_read<{type_name(e.type)}>({a[0]}, {a[1]}, {a[2]})"
        if op == CallOp.BINDLESS_BUFFER_WRITE:
            return f"lchip::dev::bindless_buffer_write({a[0]}, {a[1]}, {a[2]}, {a[3]})"
        if op in (CallOp.RAY_TRACING_QUERY_ALL, CallOp.RAY_TRACING_QUERY_ANY):
            any_ = "true" if op == CallOp.RAY_TRACING_QUERY_ANY else "false"
            return f"lcdsl::RayQuery{{{a[0]}, {a[1]}, {a[2]}, {any_}, {{}}}}"
        if op in (CallOp.RAY_QUERY_TRIANGLE_CANDIDATE_HIT, CallOp.RAY_QUERY_PROCEDURAL_CANDIDATE_HIT):
            if not self.query_depth:
                raise CodegenError("a ray-query candidate outside its RayQueryStmt")
            return "lcq_cand"
        if op == CallOp.RAY_QUERY_COMMIT_TRIANGLE:
            return "(lcq_commit = true)"
        if op == CallOp.RAY_QUERY_COMMIT_PROCEDURAL:
            return f"(lcq_t = {a[1]})"
        if op == CallOp.RAY_QUERY_COMMITTED_HIT:
            return f"{a[0]}.hit"
        raise CodegenError(f"CallOp {op.name} is not supported by this backend's codegen")

    # -- statements
    def scope(self, s):
        for st in s.statements:
            self.stmt(st)

    def block(self, head, scope):
        self.line(head + " {")
        self.depth += 1
        self.scope(scope)
        self.depth -= 1
        self.line("}")

    def stmt(self, s):
        if isinstance(s, A.BreakStmt):
            self.line("break;")
        elif isinstance(s, A.ContinueStmt):
            self.line("continue;")
        elif isinstance(s, A.ReturnStmt):
            self.line("return;" if s.expression is None else f"return {self.expr(s.expression)};")
        elif isinstance(s, A.ScopeStmt):
            self.block("", s)
        elif isinstance(s, A.IfStmt):
            self.line(f"if ({self.expr(s.condition)}) {{")
            self.depth += 1
            self.scope(s.true_branch)
            self.depth -= 1
            if s.false_branch.statements:
                self.line("} else {")
                self.depth += 1
                self.scope(s.false_branch)
                self.depth -= 1
            self.line("}")
        elif isinstance(s, A.LoopStmt):
            self.block("for (;;)", s.body)
        elif isinstance(s, A.ForStmt):
            v = self.expr(s.variable)
            self.block(f"for (; {self.expr(s.condition)}; {v} += {self.expr(s.step)})", s.body)
        elif isinstance(s, A.ExprStmt):
            self.line(f"{self.expr(s.expression)};")
        elif isinstance(s, A.AssignStmt):
            self.line(f"{self.expr(s.lhs)} = {self.expr(s.rhs)};")
        elif isinstance(s, A.CommentStmt):
            self.line("// " + s.comment.replace("\n", " "))
        elif isinstance(s, A.RayQueryStmt):
            self.ray_query(s)
        elif isinstance(s, A.SwitchStmt):
            self.block(f"switch ({self.expr(s.expression)})", s.body)
        elif isinstance(s, A.SwitchCaseStmt):
            self.block(f"case {self.expr(s.expression)}:", s.body)
            self.line("break;")
        elif isinstance(s, A.SwitchDefaultStmt):
            self.block("default:", s.body)
            self.line("break;")
        else:
            raise CodegenError(f"cannot print {type(s).__name__}")

    def ray_query(self, s):
        """query.hit = the device library's walk with the two candidate scopes as its callbacks."""
        q = self.expr(s.query)
        self.query_depth += 1
        self.line(f"{q}.hit = lchip::dev::query({q}.accel, {q}.ray, {q}.mask, {q}.any,")
        self.depth += 1
        self.line("[&](uint32_t lcq_inst, uint32_t lcq_prim, float lcq_u, float lcq_v, float lcq_tt) -> bool {")
        self.depth += 1
        self.line("const lcdsl::SurfaceCandidate lcq_cand{lcq_inst, lcq_prim, make_float2(lcq_u, lcq_v), lcq_tt};")
        self.line("bool lcq_commit = false;")
        self.line("{")
        self.depth += 1
        self.scope(s.on_triangle_candidate)
        self.depth -= 1
        self.line("}")
        self.line("return lcq_commit;")
        self.depth -= 1
        self.line("},")
        self.line("[&](uint32_t lcq_inst, uint32_t lcq_prim, const float *lcq_o, const float *lcq_d, const float *,")
        self.line("    const float *, float4 lcq_lo, float4 lcq_hi, float lcq_tmax, float &lcq_out) -> bool {")
        self.depth += 1
        self.line("const lcdsl::ProceduralCandidate lcq_cand{lcq_inst, lcq_prim, make_float3(lcq_o[0], lcq_o[1], lcq_o[2]),")
        self.line("    make_float3(lcq_d[0], lcq_d[1], lcq_d[2]), make_float3(lcq_lo.x, lcq_lo.y, lcq_lo.z),")
        self.line("    make_float3(lcq_hi.x, lcq_hi.y, lcq_hi.z), lcq_tmax};")
        # a distance outside [t_min, t_max) commits nothing (ray_query.h)
        self.line("float lcq_t = -1.0f;")
        self.line("{")
        self.depth += 1
        self.scope(s.on_procedural_candidate)
        self.depth -= 1
        self.line("}")
        self.line("lcq_out = lcq_t;")
        self.line("return true;")
        self.depth -= 1
        self.line("});")
        self.depth -= 1
        self.query_depth -= 1

    def declarations(self):
        for v in self.fn.shared_variables:  # not initialised (a shared variable has no initialiser)
            self.line(f"__shared__ {type_name(v.type)} {self.var(v)};")
        for v in self.fn.local_variables:
            self.line(f"{type_name(v.type)} {self.var(v)}{{}};")


def _callables(fn, out, seen):
    """Callables reachable from fn, each after the callables it calls."""
    for c in fn.custom_callables:
        if id(c) in seen:
            continue
        seen.add(id(c))
        _callables(c, out, seen)
        out.append(c)


def emit(kernel):
    """HIP source of a kernel Function (its callables included)."""
    if kernel.tag != A.Function.Tag.KERNEL:
        raise CodegenError("emit() takes a kernel")
    callables = []
    _callables(kernel, callables, set())
    names = {}
    for i, c in enumerate(callables):
        names[id(c)] = f"lcc_{c.name}_{i}"
    out = [PRELUDE]

    def signature(c):
        args = []
        for v in c.arguments:
            t = type_name(v.type)
            n = "a_" + (v.name or f"t{v.uid}")
            args.append(f"{t} &{n}" if v.tag == A.Variable.Tag.REFERENCE else f"{t} {n}")
        return f"__device__ {type_name(c.return_type)} {names[id(c)]}({', '.join(args)})"

    for c in callables:
        out.append(signature(c) + ";")
    for c in callables:
        p = _Printer(c, names, kernel=False)
        p.declarations()
        p.scope(c.body)
        out.append(signature(c) + " {\n" + "\n".join(p.lines) + "\n}")
    fields = "\n".join(f"    alignas(16) {type_name(v.type)} {v.name};" for v in kernel.arguments)
    bx, by, bz = kernel.block_size
    p = _Printer(kernel, names, kernel=True)
    p.line("using namespace lcdsl;")
    p.line("if (!lchip::dev::in_dispatch(lcp_.ls_kid)) return;")
    p.declarations()
    p.scope(kernel.body)
    out.append(f"struct Params_{kernel.name} {{\n{fields}\n    alignas(16) uint4 ls_kid;\n}};")
    out.append(f'extern "C" __global__ void __launch_bounds__({bx * by * bz}) {kernel.name}('
               f"const Params_{kernel.name} lcp_) {{\n" + "\n".join(p.lines) + "\n}\n")
    return "\n".join(out)
