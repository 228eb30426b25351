"""A kernel DSL on Python syntax: the AST -> HIP source half of the codegen path.

The reference's Python frontend turns a decorated Python function into LuisaCompute AST
(src/py/luisa/astbuilder.py) that a backend's codegen prints as device source
(src/backends/cuda/cuda_codegen_ast.cpp:1291-1490). Here the Python AST of the function is
printed as HIP source directly, against the device library of run-time shaders
(lc_hip_device.h), and compiled with hiprtc by Device.create_shader:

    @dsl.kernel
    def shade(rays: dsl.Buffer("Ray"), hits: dsl.Buffer("SurfaceHit"), accel: dsl.Accel, mask: dsl.uint):
        i = dispatch_id().x
        hits.write(i, accel.trace_closest(rays.read(i), mask))

    shader = shade.compile(device)
    stream << shader(ray_buffer, hit_buffer, accel, 0xFF).dispatch(n)

Supported: parameters annotated with Buffer(element), Accel, Image, Volume, BindlessArray or a scalar type
(uint, int_, float_, ulong, float3_: uniforms); nested functions (helpers and ray-query callbacks,
printed as lambdas capturing the kernel's variables by reference); assignments (the first one in a block declares with `auto`),
augmented assignments, if / elif / else, for over range(...), while, break, continue, return,
conditional expressions, arithmetic, comparisons, `and` / `or` / `not`, calls of the builtins below
and attribute access (hit.inst, v.x, ...). C++ infers the variable types.

Builtins: dispatch_id(), dispatch_size(), float2/3/4, int2/3/4, uint2/3/4, make_ray(o, d[, tmin,
tmax]), offset_ray_origin(p, n), triangle_interpolate(bary, a, b, c), dot, cross, normalize, length,
sqrt, abs, min, max, floor, ceil, sin, cos, exp, log, pow, select(c, a, b), ray_origin(r),
ray_direction(r), hit_miss(h); buffer.read(i) / buffer.write(i, v) / buffer.size();
accel.trace_closest(ray, mask) / trace_any(ray, mask) / instance_user_id(i) /
instance_visibility_mask(i); accel.query_all / query_any(ray, mask, on_surface, on_procedural)
(accel.traverse / traverse_any) with callbacks defined as nested functions of one parameter, the
candidate: on_surface(c) gets c.inst, c.prim, c.bary, c.t and returns whether to commit;
on_procedural(c) gets c.inst, c.prim, c.origin, c.direction (object space), c.aabb_min,
c.aabb_max, c.t_max and returns the distance to commit (outside [t_min, t_max): none); either may
be None; the result has inst, prim, u, v, type (0 miss, 1 surface, 2 procedural), t. image.read(x, y) / image.write(x, y, v) (float images);
heap.buffer_read(type, slot, i) with type a string naming the element type; heap.tex2d_sample(slot,
uv) / tex2d_read(slot, x, y) / tex2d_size(slot).
"""
import ast
import inspect
import textwrap

import numpy as np

from . import LCError

# ---- parameter annotations


class Buffer:
    def __init__(self, element):
        self.element = _ELEMENT_TYPES.get(element, element)

    def decl(self):
        return f"lchip::dev::Buffer<{self.element}>"


class _Resource:
    def __init__(self, decl):
        self._decl = decl

    def decl(self):
        return self._decl


class _Scalar:
    def __init__(self, decl, dtype):
        self._decl, self.dtype = decl, dtype

    def decl(self):
        return self._decl


Accel = _Resource("lchip::dev::Accel")
Image = _Resource("lchip::dev::Image")
Volume = _Resource("lchip::dev::Volume")
BindlessArray = _Resource("lchip::dev::BindlessArray")
uint = _Scalar("uint32_t", np.uint32)
int_ = _Scalar("int32_t", np.int32)
float_ = _Scalar("float", np.float32)
ulong = _Scalar("uint64_t", np.uint64)
float3_ = _Scalar("float3", np.float32)  # a 3-float uniform

_ELEMENT_TYPES = {"Ray": "lchip::dev::Ray", "SurfaceHit": "lchip::dev::SurfaceHit",
                  "CommittedHit": "lchip::rq::CommittedHit", "uint": "uint32_t", "int": "int32_t",
                  "float": "float"}

# ---- device prelude: float3 helpers over HIP's vector types

_PRELUDE = r'''
#include "lc_hip_device.h"
namespace lcdsl {
using namespace lchip;
// Python '/' converts both operands to float first (src/py/luisa/builtin.py:191-194); '//' is the
// C division of the operand types.
__device__ __forceinline__ float as_real(int32_t x) { return (float)x; }
__device__ __forceinline__ float as_real(uint32_t x) { return (float)x; }
__device__ __forceinline__ float as_real(int64_t x) { return (float)x; }
__device__ __forceinline__ float as_real(uint64_t x) { return (float)x; }
__device__ __forceinline__ float as_real(float x) { return x; }
__device__ __forceinline__ float2 as_real(float2 x) { return x; }
__device__ __forceinline__ float3 as_real(float3 x) { return x; }
__device__ __forceinline__ float4 as_real(float4 x) { return x; }
__device__ __forceinline__ float2 as_real(int2 x) { return make_float2((float)x.x, (float)x.y); }
__device__ __forceinline__ float3 as_real(int3 x) { return make_float3((float)x.x, (float)x.y, (float)x.z); }
__device__ __forceinline__ float2 as_real(uint2 x) { return make_float2((float)x.x, (float)x.y); }
__device__ __forceinline__ float3 as_real(uint3 x) { return make_float3((float)x.x, (float)x.y, (float)x.z); }
template<class A, class B> __device__ __forceinline__ auto true_div(A a, B b) { return as_real(a) / as_real(b); }
__device__ __forceinline__ float3 v3f(rtx::V3 v) { return make_float3(v.x, v.y, v.z); }
__device__ __forceinline__ rtx::V3 fv3(float3 v) { return rtx::v3(v.x, v.y, v.z); }
__device__ __forceinline__ float dot(float3 a, float3 b) { return rtx::dot(fv3(a), fv3(b)); }
__device__ __forceinline__ float3 cross(float3 a, float3 b) { return v3f(rtx::cross(fv3(a), fv3(b))); }
__device__ __forceinline__ float length(float3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ float3 normalize(float3 a) { return v3f(rtx::normalize(fv3(a))); }
__device__ __forceinline__ float3 offset_ray_origin(float3 p, float3 n) { return v3f(rtx::offset_ray_origin(fv3(p), fv3(n))); }
__device__ __forceinline__ float3 triangle_interpolate(float2 b, float3 x, float3 y, float3 z) {
    return v3f(rtx::triangle_interpolate(b.x, b.y, fv3(x), fv3(y), fv3(z)));
}
__device__ __forceinline__ dev::Ray make_ray(float3 o, float3 d, float tmin = 0.0f, float tmax = FLT_MAX) {
    return dev::make_ray(fv3(o), fv3(d), tmin, tmax);
}
__device__ __forceinline__ float3 ray_origin(const dev::Ray &r) { return make_float3(r.origin[0], r.origin[1], r.origin[2]); }
__device__ __forceinline__ float3 ray_direction(const dev::Ray &r) {
    return make_float3(r.direction[0], r.direction[1], r.direction[2]);
}
__device__ __forceinline__ float2 hit_bary(const dev::SurfaceHit &h) { return make_float2(h.bary[0], h.bary[1]); }
__device__ __forceinline__ bool hit_miss(const dev::SurfaceHit &h) { return h.inst == 0xFFFFFFFFu; }
template<class T> __device__ __forceinline__ T select(bool c, T a, T b) { return c ? a : b; }
// ray-query candidates as the DSL callbacks see them
struct SurfaceCandidate { uint32_t inst, prim; float2 bary; float t; };
struct ProceduralCandidate { uint32_t inst, prim; float3 origin, direction, aabb_min, aabb_max; float t_max; };
template<class F> __device__ __forceinline__ auto surface_cb(F f) {
    return [f](uint32_t inst, uint32_t prim, float u, float v, float t) {
        return (bool)f(SurfaceCandidate{inst, prim, make_float2(u, v), t});
    };
}
template<class F> __device__ __forceinline__ auto procedural_cb(F f) {
    return [f](uint32_t inst, uint32_t prim, const float *o, const float *d, const float *, const float *, float4 lo,
               float4 hi, float tmax, float &t) {
        t = (float)f(ProceduralCandidate{inst, prim, make_float3(o[0], o[1], o[2]), make_float3(d[0], d[1], d[2]),
                                         make_float3(lo.x, lo.y, lo.z), make_float3(hi.x, hi.y, hi.z), tmax});
        return true; // a distance outside [t_min, t_max) commits nothing
    };
}
__device__ __forceinline__ auto no_surface() { return [](uint32_t, uint32_t, float, float, float) { return false; }; }
__device__ __forceinline__ auto no_procedural() {
    return [](uint32_t, uint32_t, const float *, const float *, const float *, const float *, float4, float4, float,
              float &) { return false; };
}
} // namespace lcdsl
'''

_FUNCS = {"float2": "make_float2", "float3": "make_float3", "float4": "make_float4", "int2": "make_int2",
          "int3": "make_int3", "int4": "make_int4", "uint2": "make_uint2", "uint3": "make_uint3",
          "uint4": "make_uint4", "sqrt": "sqrtf", "abs": "fabsf", "min": "fminf", "max": "fmaxf",
          "floor": "floorf", "ceil": "ceilf", "sin": "sinf", "cos": "cosf", "exp": "expf", "log": "logf",
          "pow": "powf", "dot": "lcdsl::dot", "cross": "lcdsl::cross", "normalize": "lcdsl::normalize",
          "length": "lcdsl::length", "offset_ray_origin": "lcdsl::offset_ray_origin",
          "triangle_interpolate": "lcdsl::triangle_interpolate", "make_ray": "lcdsl::make_ray",
          "ray_origin": "lcdsl::ray_origin", "ray_direction": "lcdsl::ray_direction", "hit_miss": "lcdsl::hit_miss",
          "hit_bary": "lcdsl::hit_bary", "select": "lcdsl::select", "uint": "(uint32_t)", "int": "(int32_t)",
          "float": "(float)", "ulong": "(uint64_t)", "isnan": "isnan"}

_BINOPS = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.FloorDiv: "/", ast.Mod: "%", ast.BitAnd: "&", ast.BitOr: "|",
           ast.BitXor: "^", ast.LShift: "<<", ast.RShift: ">>"}
_CMPOPS = {ast.Eq: "==", ast.NotEq: "!=", ast.Lt: "<", ast.LtE: "<=", ast.Gt: ">", ast.GtE: ">="}


class _Codegen(ast.NodeVisitor):
    def __init__(self, params, constants=None):
        self.params = params  # name -> annotation
        self.constants = constants or {}  # numbers the kernel names from its module
        self.lines = []
        self.indent = 1
        self.scopes = [set(params)]
        self.in_callback = 0

    def emit(self, line):
        self.lines.append("    " * self.indent + line)

    def error(self, node, msg):
        raise LCError(f"dsl: line {getattr(node, 'lineno', '?')}: {msg}")

    # -- statements
    def block(self, body):
        self.scopes.append(set())
        self.indent += 1
        for st in body:
            self.visit(st)
        self.indent -= 1
        self.scopes.pop()

    def declared(self, name):
        return any(name in s for s in self.scopes)

    def visit_Assign(self, node):
        if len(node.targets) != 1:
            self.error(node, "one assignment target at a time")
        self.assign(node.targets[0], self.expr(node.value), node)

    def assign(self, target, value, node):
        if isinstance(target, ast.Name):
            if target.id in self.params:
                self.error(node, f"cannot assign to parameter {target.id}")
            if self.declared(target.id):
                self.emit(f"{target.id} = {value};")
            else:
                self.scopes[-1].add(target.id)
                self.emit(f"auto {target.id} = {value};")
        elif isinstance(target, (ast.Attribute, ast.Subscript)):
            self.emit(f"{self.expr(target)} = {value};")
        else:
            self.error(node, "unsupported assignment target")

    def visit_AugAssign(self, node):
        if isinstance(node.op, ast.Div):
            t = self.expr(node.target)
            self.emit(f"{t} = true_div({t}, {self.expr(node.value)});")
            return
        op = _BINOPS.get(type(node.op))
        if op is None:
            self.error(node, "unsupported operator")
        self.emit(f"{self.expr(node.target)} {op}= {self.expr(node.value)};")

    def visit_Expr(self, node):
        if isinstance(node.value, ast.Constant) and isinstance(node.value.value, str):
            return  # docstring
        self.emit(f"{self.expr(node.value)};")

    def visit_If(self, node):
        self.emit(f"if ({self.expr(node.test)}) {{")
        self.block(node.body)
        if node.orelse:
            self.emit("} else {")
            self.block(node.orelse)
        self.emit("}")

    def visit_While(self, node):
        self.emit(f"while ({self.expr(node.test)}) {{")
        self.block(node.body)
        self.emit("}")

    def visit_For(self, node):
        it = node.iter
        if not (isinstance(it, ast.Call) and isinstance(it.func, ast.Name) and it.func.id == "range"
                and isinstance(node.target, ast.Name) and not node.orelse):
            self.error(node, "for loops run over range(...)")
        args = [self.expr(a) for a in it.args]
        lo, hi, step = ("0", args[0], "1") if len(args) == 1 else (args[0], args[1], args[2] if len(args) > 2 else "1")
        v = node.target.id
        self.emit(f"for (int32_t {v} = {lo}; ({step}) > 0 ? {v} < ({hi}) : {v} > ({hi}); {v} += {step}) {{")
        self.scopes.append({v})
        self.block(node.body)
        self.scopes.pop()
        self.emit("}")

    def visit_Break(self, node):
        self.emit("break;")

    def visit_Continue(self, node):
        self.emit("continue;")

    def visit_Return(self, node):
        if self.in_callback:
            self.emit("return;" if node.value is None else f"return {self.expr(node.value)};")
            return
        if node.value is not None:
            self.error(node, "kernels return nothing")
        self.emit("return;")

    def visit_FunctionDef(self, node):
        # a nested function (a helper, or a ray-query callback of one parameter, the candidate):
        # a lambda capturing the kernel's variables by reference, parameters by value
        if node.decorator_list or node.args.vararg or node.args.kwarg or node.args.defaults:
            self.error(node, "nested functions take plain parameters")
        names = [a.arg for a in node.args.args]
        self.scopes[-1].add(node.name)
        self.emit(f"auto {node.name} = [&](" + ", ".join(f"auto {n}" for n in names) + ") {")
        self.in_callback += 1
        self.scopes.append(set(names))
        self.block(node.body)
        self.scopes.pop()
        self.in_callback -= 1
        self.emit("};")

    def visit_Pass(self, node):
        pass

    def generic_visit(self, node):
        self.error(node, f"unsupported statement {type(node).__name__}")

    # -- expressions
    def expr(self, node):
        if isinstance(node, ast.Constant) and node.value is None:
            return "nullptr"
        if isinstance(node, ast.Constant) and isinstance(node.value, str):
            return '"' + node.value + '"'  # only as the element type of a bindless read
        if isinstance(node, ast.Constant):
            v = node.value
            if isinstance(v, bool):
                return "true" if v else "false"
            if isinstance(v, int):
                return f"{v}ull" if v > 0xFFFFFFFF else (f"{v}u" if v > 0x7FFFFFFF else str(v))
            if isinstance(v, float):
                return f"{np.float32(v)!r}".replace("np.float32(", "").rstrip(")") + "f" if np.isfinite(v) else \
                    ("FLT_MAX" if v > 0 else "-FLT_MAX")
            self.error(node, "unsupported constant")
        if isinstance(node, ast.Name):
            if node.id in self.params and isinstance(self.params[node.id], (Buffer, _Resource, _Scalar)):
                return f"lcp_.{node.id}"
            if not self.declared(node.id) and isinstance(self.constants.get(node.id), (int, float)) and \
                    not isinstance(self.constants.get(node.id), bool):
                return self.expr(ast.Constant(self.constants[node.id]))
            return node.id
        if isinstance(node, ast.BinOp):
            if isinstance(node.op, ast.Div):
                return f"true_div({self.expr(node.left)}, {self.expr(node.right)})"
            op = _BINOPS.get(type(node.op))
            if op is None:
                self.error(node, "unsupported operator")
            return f"({self.expr(node.left)} {op} {self.expr(node.right)})"
        if isinstance(node, ast.UnaryOp):
            op = {ast.USub: "-", ast.UAdd: "+", ast.Not: "!", ast.Invert: "~"}[type(node.op)]
            return f"({op}{self.expr(node.operand)})"
        if isinstance(node, ast.BoolOp):
            op = " && " if isinstance(node.op, ast.And) else " || "
            return "(" + op.join(self.expr(v) for v in node.values) + ")"
        if isinstance(node, ast.Compare):
            parts, left = [], self.expr(node.left)
            for o, c in zip(node.ops, node.comparators):
                right = self.expr(c)
                parts.append(f"({left} {_CMPOPS[type(o)]} {right})")
                left = right
            return "(" + " && ".join(parts) + ")"
        if isinstance(node, ast.IfExp):
            return f"({self.expr(node.test)} ? {self.expr(node.body)} : {self.expr(node.orelse)})"
        if isinstance(node, ast.Attribute):
            return f"{self.expr(node.value)}.{node.attr}"
        if isinstance(node, ast.Subscript):
            return f"{self.expr(node.value)}[{self.expr(node.slice)}]"
        if isinstance(node, ast.Call):
            return self.call(node)
        self.error(node, f"unsupported expression {type(node).__name__}")

    def call(self, node):
        args = [self.expr(a) for a in node.args]
        f = node.func
        if isinstance(f, ast.Name):
            if f.id == "dispatch_id":
                return "lchip::dev::dispatch_id()"
            if f.id == "dispatch_size":
                return "make_uint3(lcp_.ls_kid.x, lcp_.ls_kid.y, lcp_.ls_kid.z)"
            if self.declared(f.id) and f.id not in self.params:  # a nested function
                return f"{f.id}({', '.join(args)})"
            name = _FUNCS.get(f.id)
            if name is None:
                self.error(node, f"unknown function {f.id}")
            return f"{name}({', '.join(args)})"
        if isinstance(f, ast.Attribute) and isinstance(f.value, ast.Name) and f.value.id in self.params:
            obj, kind, m = f"lcp_.{f.value.id}", self.params[f.value.id], f.attr
            if isinstance(kind, Buffer):
                if m == "read":
                    return f"{obj}[{args[0]}]"
                if m == "write":
                    return f"({obj}[{args[0]}] = {args[1]})"
                if m == "size":
                    return f"{obj}.size()"
            if kind is Accel:
                if m in ("trace_closest", "trace_any", "instance_user_id", "instance_visibility_mask"):
                    return f"lchip::dev::{m}({obj}, {', '.join(args)})"
                if m in ("query_all", "query_any"):
                    # accel.traverse / traverse_any (ray_query.h): (ray, mask, on_surface, on_procedural);
                    # a callback may be None
                    if len(node.args) != 4:
                        self.error(node, f"{m}(ray, mask, on_surface, on_procedural)")
                    none = lambda a: isinstance(a, ast.Constant) and a.value is None
                    sfc = "lcdsl::no_surface()" if none(node.args[2]) else f"lcdsl::surface_cb({args[2]})"
                    prc = "lcdsl::no_procedural()" if none(node.args[3]) else f"lcdsl::procedural_cb({args[3]})"
                    first = "true" if m == "query_any" else "false"
                    return f"lchip::dev::query({obj}, {args[0]}, {args[1]}, {first}, {sfc}, {prc})"
            if kind is Image:
                if m == "read":
                    return f"lchip::dev::image_read_float({obj}, {args[0]}, {args[1]})"
                if m == "write":
                    return f"lchip::dev::image_write_float({obj}, {args[0]}, {args[1]}, {args[2]})"
            if kind is Volume:
                if m == "read":
                    return f"lchip::dev::volume_read_float({obj}, {args[0]}, {args[1]}, {args[2]})"
                if m == "write":
                    return f"lchip::dev::volume_write_float({obj}, {args[0]}, {args[1]}, {args[2]}, {args[3]})"
            if kind is BindlessArray and m == "tex2d_sample":
                return f"lchip::dev::bindless_tex2d_sample({obj}, {args[0]}, {args[1]}.x, {args[1]}.y)"
            if kind is BindlessArray and m == "tex2d_read":
                return f"lchip::dev::bindless_tex2d_read({obj}, {args[0]}, {args[1]}, {args[2]})"
            if kind is BindlessArray and m == "tex2d_size":
                return f"lchip::dev::bindless_tex2d_size({obj}, {args[0]})"
            if kind is BindlessArray and m == "buffer_read":
                el = node.args[0]
                if not (isinstance(el, ast.Constant) and isinstance(el.value, str)):
                    self.error(node, "buffer_read takes the element type as a string")
                t = _ELEMENT_TYPES.get(el.value, el.value)
                return f"lchip::dev::bindless_buffer_read<{t}>({obj}, {args[1]}, {args[2]})"
            self.error(node, f"unknown method {m} of parameter {f.value.id}")
        self.error(node, "unsupported call")


class Kernel:
    """A DSL kernel: `source` is its HIP text; compile(device) returns a Shader whose call packs
    the arguments in parameter order (scalars as uniforms of the annotated type)."""

    def __init__(self, fn, block=(128, 1, 1)):
        self.fn, self.block, self.name = fn, tuple(block), fn.__name__
        tree = ast.parse(textwrap.dedent(inspect.getsource(fn)))
        fdef = tree.body[0]
        hints = fn.__annotations__
        self.params = {}
        for a in fdef.args.args:
            ann = hints.get(a.arg)
            if not isinstance(ann, (Buffer, _Resource, _Scalar)):
                raise LCError(f"dsl: parameter {a.arg} needs a Buffer / Accel / Image / Volume / BindlessArray / scalar annotation")
            self.params[a.arg] = ann
        gen = _Codegen(self.params, {k: v for k, v in fn.__globals__.items() if isinstance(v, (int, float))})
        for st in fdef.body:
            gen.visit(st)
        fields = "\n".join(f"    alignas(16) {ann.decl()} {n};" for n, ann in self.params.items())
        bx, by, bz = self.block
        self.source = (_PRELUDE + f"struct Params_{self.name} {{\n{fields}\n    alignas(16) uint4 ls_kid;\n}};\n"
                       f"extern \"C\" __global__ void __launch_bounds__({bx * by * bz}) {self.name}("
                       f"const Params_{self.name} lcp_) {{\n    using namespace lcdsl;\n"
                       f"    if (!lchip::dev::in_dispatch(lcp_.ls_kid)) return;\n" + "\n".join(gen.lines) + "\n}\n")

    def compile(self, device):
        shader = device.create_shader(self.source, self.name, self.block)
        params = list(self.params.values())

        def bind(*args):
            if len(args) != len(params):
                raise LCError(f"{self.name} takes {len(params)} arguments")
            packed = [np.asarray(a, p.dtype) if isinstance(p, _Scalar) else a for a, p in zip(args, params)]
            return shader(*packed)
        bind.shader = shader
        return bind


def kernel(fn=None, block=(128, 1, 1)):
    """Decorator: @kernel or @kernel(block=(bx, by, bz))."""
    if fn is None:
        return lambda f: Kernel(f, block)
    return Kernel(fn, block)
