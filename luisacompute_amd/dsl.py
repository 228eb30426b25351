"""A kernel DSL on Python syntax: the frontend of the AST -> HIP codegen path.

The reference's Python frontend turns a decorated Python function into LuisaCompute AST
(src/py/luisa/astbuilder.py, types and builtins in src/py/luisa/types.py and builtin.py) that a
backend's codegen prints as device source (src/backends/cuda/cuda_codegen_ast.cpp:1291-1490).
Here likewise: the Python AST of a `@dsl.kernel` function is built into a typed lcast.Function
(types inferred, builtins lowered to CallOps, nested functions to callables) and codegen.emit
prints it as HIP source for hiprtc (Device.create_shader, code objects cached by source hash):

    @dsl.kernel
    def shade(rays: dsl.Buffer("Ray"), hits: dsl.Buffer("SurfaceHit"), accel: dsl.Accel, mask: dsl.uint):
        i = dispatch_id().x
        hits.write(i, accel.trace_closest(rays.read(i), mask))

    shader = shade.compile(device)
    stream << shader(ray_buffer, hit_buffer, accel, 0xFF).dispatch(n)

Parameters are annotated with Buffer(element), Accel, Image, Volume, BindlessArray or a uniform
type (uint, int_, float_, ulong, float3_). Types: bool, int, uint, long, ulong, float, their 2-4
vectors, the structures Ray, SurfaceHit, CommittedHit and the ray-query candidates. Python int
literals are int (or take the type of the other operand), float literals float; arithmetic
converts both operands to a common type (float over integers, 64 over 32 bits, unsigned over
signed; a scalar broadcasts to a vector); '/' is float division (src/py/luisa/builtin.py:191-194),
'//' integer division (floor for floats). A variable keeps the type of its first assignment
(numeric assignments convert). Statements: assignments (to names, members, lanes), augmented
assignments, if / elif / else, for over range(...), while, break, continue, return, expression
statements; conditional expressions evaluate only the branch taken.

Nested functions are callables: typed per argument types at their calls, returning the type of
their first return. A name a nested function reads or assigns that belongs to an enclosing
function is shared with it (passed by reference; resources and uniforms by value).

Builtins: dispatch_id(), dispatch_size(), thread_id(), block_id(), warp_lane_id(); float2/3/4,
int2/3/4, uint2/3/4, float(), int(), uint(), ulong(), bool(); sqrt, rsqrt, abs, min, max, clamp,
lerp, saturate, step, fma, floor, ceil, fract, trunc, round, sin, cos, tan, asin, acos, atan,
atan2, exp, exp2, log, log2, log10, pow (and **), copysign, isnan, isinf, dot, cross, normalize,
length, length_squared, reduce_sum/min/max, select(c, a, b) (= c ? a : b), synchronize_block() (or
sync_block()); name = shared_array("type", n) declares a workgroup-shared array (kernels only);
warp_active_count_bits(b), warp_prefix_count_bits(b), warp_is_first_active_lane(),
warp_read_first_active_lane(x), warp_read_lane(x, i); make_ray(o, d[, t_min, t_max]),
offset_ray_origin(p, n), triangle_interpolate(bary, a, b, c), ray_origin(r), ray_direction(r),
hit_miss(h), hit_bary(h); constants FLT_MAX and the module's int / float globals.
Methods: buffer.read(i) / write(i, v) / size() / atomic_fetch_add|sub|and|or|xor|min|max(i, v) /
atomic_exchange(i, v) / atomic_compare_exchange(i, expected, desired); accel.trace_closest(ray,
mask=255) / trace_any / instance_user_id(i) / instance_visibility_mask(i) / set_instance_user_id(i,
v) / set_instance_visibility(i, m) / query_all(ray, mask, on_surface, on_procedural) / query_any
(accel.traverse / traverse_any: a RayQueryStmt whose candidate scopes call the callbacks; either
may be None; on_surface(c) gets c.inst, c.prim, c.bary, c.t and returns whether to commit;
on_procedural(c) gets c.inst, c.prim, c.origin, c.direction (object space), c.aabb_min,
c.aabb_max, c.t_max and returns the distance to commit (outside [t_min, t_max): none); the result
has inst, prim, u, v, type (0 miss, 1 surface, 2 procedural), t); image.read(x, y) /
write(x, y, v); volume.read(x, y, z) / write(x, y, z, v); heap.buffer_read(type, slot, i) with
type a string naming the element type, heap.tex2d_sample(slot, uv) / tex2d_read(slot, x, y) /
tex2d_size(slot).

Errors (unknown names, type mismatches, unsupported syntax) raise LCError with the kernel's line.
"""
import ast
import inspect
import textwrap

import numpy as np

from . import LCError
from . import codegen
from . import lcast as A
from .lcast import CallOp, Type

# ---- parameter annotations ---------------------------------------------------------------------

_ELEMENT_TYPES = {"Ray": A.RAY, "SurfaceHit": A.SURFACE_HIT, "CommittedHit": A.COMMITTED_HIT, "uint": A.UINT,
                  "int": A.INT, "float": A.FLOAT, "ulong": A.ULONG, "long": A.LONG, "bool": A.BOOL,
                  "float2": A.FLOAT2, "float3": A.FLOAT3, "float4": A.FLOAT4, "uint2": A.UINT2, "uint3": A.UINT3,
                  "uint4": A.UINT4, "int2": A.INT2, "int3": A.INT3, "int4": A.INT4}


def _element(name):
    if isinstance(name, Type):
        return name
    t = _ELEMENT_TYPES.get(name)
    if t is None:
        t = Type.of(name)
    return t


class Buffer:
    def __init__(self, element):
        self.type = Type.buffer(_element(element))


class _Annotation:
    def __init__(self, t, dtype=None):
        self.type, self.dtype = t, dtype


Accel = _Annotation(Type.accel())
Image = _Annotation(Type.texture(2, A.FLOAT))
Volume = _Annotation(Type.texture(3, A.FLOAT))
BindlessArray = _Annotation(Type.bindless_array())
uint = _Annotation(A.UINT, np.uint32)
int_ = _Annotation(A.INT, np.int32)
float_ = _Annotation(A.FLOAT, np.float32)
ulong = _Annotation(A.ULONG, np.uint64)
float3_ = _Annotation(A.FLOAT3, np.float32)  # a 3-float uniform

FLT_MAX = 3.4028234663852886e38

_BINOPS = {ast.Add: A.BinaryOp.ADD, ast.Sub: A.BinaryOp.SUB, ast.Mult: A.BinaryOp.MUL, ast.Mod: A.BinaryOp.MOD,
           ast.BitAnd: A.BinaryOp.BIT_AND, ast.BitOr: A.BinaryOp.BIT_OR, ast.BitXor: A.BinaryOp.BIT_XOR,
           ast.LShift: A.BinaryOp.SHL, ast.RShift: A.BinaryOp.SHR}
_CMPOPS = {ast.Eq: A.BinaryOp.EQUAL, ast.NotEq: A.BinaryOp.NOT_EQUAL, ast.Lt: A.BinaryOp.LESS,
           ast.LtE: A.BinaryOp.LESS_EQUAL, ast.Gt: A.BinaryOp.GREATER, ast.GtE: A.BinaryOp.GREATER_EQUAL}
_VEC_MAKERS = {"float2": (A.FLOAT, 2, CallOp.MAKE_FLOAT2), "float3": (A.FLOAT, 3, CallOp.MAKE_FLOAT3),
               "float4": (A.FLOAT, 4, CallOp.MAKE_FLOAT4), "int2": (A.INT, 2, CallOp.MAKE_INT2),
               "int3": (A.INT, 3, CallOp.MAKE_INT3), "int4": (A.INT, 4, CallOp.MAKE_INT4),
               "uint2": (A.UINT, 2, CallOp.MAKE_UINT2), "uint3": (A.UINT, 3, CallOp.MAKE_UINT3),
               "uint4": (A.UINT, 4, CallOp.MAKE_UINT4)}
_CASTS = {"float": A.FLOAT, "int": A.INT, "uint": A.UINT, "ulong": A.ULONG, "long": A.LONG, "bool": A.BOOL}
_FLOAT_UNARY = {"sqrt": CallOp.SQRT, "rsqrt": CallOp.RSQRT, "floor": CallOp.FLOOR, "ceil": CallOp.CEIL,
                "fract": CallOp.FRACT, "trunc": CallOp.TRUNC, "round": CallOp.ROUND, "sin": CallOp.SIN,
                "cos": CallOp.COS, "tan": CallOp.TAN, "asin": CallOp.ASIN, "acos": CallOp.ACOS, "atan": CallOp.ATAN,
                "exp": CallOp.EXP, "exp2": CallOp.EXP2, "log": CallOp.LOG, "log2": CallOp.LOG2,
                "log10": CallOp.LOG10, "saturate": CallOp.SATURATE}
_FLOAT_BINARY = {"pow": CallOp.POW, "atan2": CallOp.ATAN2, "copysign": CallOp.COPYSIGN, "step": CallOp.STEP}
_ATOMIC_METHODS = {"atomic_fetch_add": CallOp.ATOMIC_FETCH_ADD, "atomic_fetch_sub": CallOp.ATOMIC_FETCH_SUB,
                   "atomic_fetch_and": CallOp.ATOMIC_FETCH_AND, "atomic_fetch_or": CallOp.ATOMIC_FETCH_OR,
                   "atomic_fetch_xor": CallOp.ATOMIC_FETCH_XOR, "atomic_fetch_min": CallOp.ATOMIC_FETCH_MIN,
                   "atomic_fetch_max": CallOp.ATOMIC_FETCH_MAX, "atomic_exchange": CallOp.ATOMIC_EXCHANGE}
_EXTERNAL = {
    "offset_ray_origin": ("lcdsl::offset_ray_origin", A.FLOAT3, (A.FLOAT3, A.FLOAT3)),
    "triangle_interpolate": ("lcdsl::triangle_interpolate", A.FLOAT3, (A.FLOAT2, A.FLOAT3, A.FLOAT3, A.FLOAT3)),
    "ray_origin": ("lcdsl::ray_origin", A.FLOAT3, (A.RAY,)),
    "ray_direction": ("lcdsl::ray_direction", A.FLOAT3, (A.RAY,)),
    "hit_miss": ("lcdsl::hit_miss", A.BOOL, (A.SURFACE_HIT,)),
    "hit_bary": ("lcdsl::hit_bary", A.FLOAT2, (A.SURFACE_HIT,)),
}
_SWIZZLE = {"x": 0, "y": 1, "z": 2, "w": 3, "r": 0, "g": 1, "b": 2, "a": 3}


def _rank(t):
    """Scalar conversion rank: float > 64-bit > 32-bit > smaller; unsigned over signed."""
    if t.is_float:
        return 100 + t.size
    if t.is_bool:
        return 0
    return t.size * 10 + (1 if t.is_unsigned else 0)


def _lit(t, v, adaptive=False):
    """A literal; `adaptive` ones (written as Python numbers) take the type of the other operand
    of an arithmetic operation, explicitly typed ones (uint(3), ulong(64)) keep theirs."""
    e = A.LiteralExpr(t, v)
    e.adaptive = adaptive
    return e


def _adaptive(e):
    return isinstance(e, A.LiteralExpr) and getattr(e, "adaptive", False)


class _CallableDef:
    """A nested function: its Python definition and the builder it was defined in."""

    def __init__(self, node, owner):
        self.node, self.owner = node, owner
        self.instances = {}  # argument types -> (Function, captures)


class _Builder:
    """Builds one lcast.Function from Python statements."""

    def __init__(self, fn, kernel, globals_, parent=None, fname=""):
        self.fn, self.kernel, self.globals, self.parent, self.fname = fn, kernel, globals_, parent, fname
        self.names = {}          # name -> Variable | _CallableDef
        self.captures = []       # (outer variable, own argument) for callables
        self.scope_stack = [fn.body]
        self.loop_depth = 0
        self.line = 0

    # -- errors and statements
    def error(self, node, msg):
        raise LCError(f"dsl: {self.fname} line {getattr(node, 'lineno', self.line)}: {msg}")

    @property
    def scope(self):
        return self.scope_stack[-1]

    def push(self, s):
        return self.scope.append(s)

    def in_scope(self, scope, body):
        self.scope_stack.append(scope)
        try:
            for st in body:
                self.stmt(st)
        finally:
            self.scope_stack.pop()

    # -- names
    def lookup(self, name, node):
        """Variable or callable bound to `name` here, capturing from enclosing functions."""
        if name in self.names:
            return self.names[name]
        if self.parent is not None:
            outer = self.parent.lookup_soft(name)
            if isinstance(outer, _CallableDef):
                return outer
            if isinstance(outer, A.Variable):
                return self.capture(outer, name)
        return None

    def lookup_soft(self, name):
        return self.lookup(name, None)

    def capture(self, outer, name):
        """An argument standing for the enclosing function's variable `outer`."""
        for o, mine in self.captures:
            if o is outer:
                return mine
        ref = outer.tag in (A.Variable.Tag.LOCAL, A.Variable.Tag.REFERENCE) and not outer.is_builtin and \
            not (self.parent.kernel and outer in self.parent.fn.arguments)
        if outer.is_builtin:
            v = self.fn.argument(outer.type, name)
        else:
            v = self.fn.argument(outer.type, name, reference=ref)
        self.captures.append((outer, v))
        self.names[name] = v
        return v

    def outer_expr(self, outer):
        """This builder's expression for a variable of some enclosing function (or its own)."""
        if outer in self.fn.arguments or outer in self.fn.local_variables or outer in self.fn.builtin_variables:
            return A.RefExpr(outer)
        if self.parent is None:
            raise LCError(f"dsl: {self.fname}: variable {outer.name} is not visible here")
        name = outer.name
        for o, mine in self.captures:
            if o is outer:
                return A.RefExpr(mine)
        # walk up: the parent must see it first
        self.parent.outer_expr(outer)
        return A.RefExpr(self.capture(outer, f"{name}_c{outer.uid}"))

    def ref(self, v, usage=A.Usage.READ):
        self.fn.mark(v, usage)
        return A.RefExpr(v)

    # -- conversions
    def cast(self, e, t, node=None):
        if e.type is t:
            return e
        if isinstance(e, A.LiteralExpr) and t.is_scalar:
            v = e.value
            if t.is_bool:
                return _lit(t, bool(v))
            if t.is_float:
                return _lit(t, float(v))
            if isinstance(v, float) and not float(v).is_integer():
                return A.CastExpr(t, A.CastOp.STATIC, e)
            return _lit(t, int(v))
        if e.type is not None and e.type.is_scalar and t.is_scalar:
            return A.CastExpr(t, A.CastOp.STATIC, e)
        if e.type is not None and e.type.is_scalar and t.is_vector:
            s = self.cast(e, t.element, node)
            return A.CallExpr(t, _maker(t), [s] * t.dimension)
        if e.type is not None and e.type.is_vector and t.is_vector and e.type.dimension == t.dimension:
            return A.CastExpr(t, A.CastOp.STATIC, e)
        self.error(node, f"cannot convert {e.type.description if e.type else 'void'} to {t.description}")

    def to_bool(self, e, node):
        if e.type is A.BOOL:
            return e
        if e.type is not None and e.type.is_scalar:
            return A.BinaryExpr(A.BOOL, A.BinaryOp.NOT_EQUAL, e, self.cast(_lit(A.INT, 0), e.type))
        self.error(node, f"condition of type {e.type.description if e.type else 'void'}")

    def common(self, a, b, node):
        """Common type of two operands (literals adapt to the other side)."""
        ta, tb = a.type, b.type
        if ta is None or tb is None:
            self.error(node, "void operand")
        if ta is tb:
            return ta
        la, lb = _adaptive(a), _adaptive(b)
        sa, sb = ta.scalar_type, tb.scalar_type
        if ta.is_vector and tb.is_vector:
            if ta.dimension != tb.dimension:
                self.error(node, f"{ta.description} and {tb.description}")
            el = self._scalar_common(sa, sb, False, False)
            return Type.vector(el, ta.dimension)
        if ta.is_vector or tb.is_vector:
            v = ta if ta.is_vector else tb
            s = sb if ta.is_vector else sa
            lit = lb if ta.is_vector else la
            if not s.is_scalar:
                self.error(node, f"{ta.description} and {tb.description}")
            el = v.element if (lit or _rank(v.element) >= _rank(s)) else s
            return Type.vector(el, v.dimension)
        if not (ta.is_scalar and tb.is_scalar):
            self.error(node, f"{ta.description} and {tb.description}")
        return self._scalar_common(ta, tb, la, lb)

    @staticmethod
    def _scalar_common(ta, tb, la, lb):
        if la and not lb:
            if tb.is_float or not ta.is_float:
                return tb
            return ta
        if lb and not la:
            if ta.is_float or not tb.is_float:
                return ta
            return tb
        return ta if _rank(ta) >= _rank(tb) else tb

    # -- statements
    def stmt(self, node):
        self.line = getattr(node, "lineno", self.line)
        m = getattr(self, "s_" + type(node).__name__, None)
        if m is None:
            self.error(node, f"unsupported statement {type(node).__name__}")
        m(node)

    def s_Pass(self, node):
        pass

    def s_Expr(self, node):
        if isinstance(node.value, ast.Constant) and isinstance(node.value.value, str):
            return  # docstring
        e = self.expr(node.value)
        if e is not None:
            self.push(A.ExprStmt(e))

    def s_Assign(self, node):
        if len(node.targets) != 1:
            self.error(node, "one assignment target at a time")
        v = node.value
        if isinstance(v, ast.Call) and isinstance(v.func, ast.Name) and v.func.id == "shared_array":
            self.shared_array(node.targets[0], v, node)
            return
        self.assign(node.targets[0], self.expr(node.value), node)

    def shared_array(self, target, call, node):
        """name = shared_array("elem", n): an array of n elements shared by the workgroup's threads
        (SharedArrayType(n, dtype)() of the reference Python frontend, src/py/luisa/array.py:73-95);
        not initialised; indexed like an array, with atomic_* methods for integer elements."""
        if not self.kernel:
            self.error(node, "shared arrays belong to kernels")
        a = call.args
        if not (isinstance(target, ast.Name) and len(a) == 2 and isinstance(a[0], ast.Constant) and
                isinstance(a[0].value, str) and isinstance(a[1], ast.Constant) and isinstance(a[1].value, int)
                and a[1].value > 0):
            self.error(node, 'name = shared_array("element type", positive size)')
        if self.lookup_soft(target.id) is not None:
            self.error(node, f"{target.id} is already defined")
        el = Type.of(a[0].value)
        if not el.is_scalar and not el.is_vector:
            self.error(node, "shared arrays hold scalars or vectors")
        self.names[target.id] = self.fn.shared(Type.array(el, a[1].value), target.id)

    def assign(self, target, value, node):
        if value is None or value.type is None:
            self.error(node, "assigning a value of type void")
        if isinstance(target, ast.Name):
            v = self.lookup(target.id, node)
            if isinstance(v, _CallableDef):
                self.error(node, f"{target.id} names a function")
            if v is None:
                if value.type.is_resource:
                    self.error(node, "resources cannot be copied into variables")
                v = self.fn.local(value.type, target.id)
                self.names[target.id] = v
            elif v.is_resource or (self.kernel and v in self.fn.arguments):
                self.error(node, f"cannot assign to parameter {target.id}")
            self.fn.mark(v, A.Usage.WRITE)
            self.push(A.AssignStmt(A.RefExpr(v), self.convert_assign(value, v.type, node)))
            return
        if isinstance(target, (ast.Attribute, ast.Subscript)):
            lhs = self.expr(target)
            self._mark_write(lhs)
            self.push(A.AssignStmt(lhs, self.convert_assign(value, lhs.type, node)))
            return
        self.error(node, "unsupported assignment target")

    def _mark_write(self, e):
        while isinstance(e, (A.MemberExpr, A.AccessExpr)):
            e = e.self if isinstance(e, A.MemberExpr) else e.range
        if isinstance(e, A.RefExpr):
            if self.kernel and e.variable in self.fn.arguments:
                self.error(None, f"cannot assign to parameter {e.variable.name}")
            self.fn.mark(e.variable, A.Usage.WRITE)

    def convert_assign(self, value, t, node):
        if value.type is t:
            return value
        if value.type.is_scalar and t.is_scalar or (value.type.is_vector and t.is_vector and
                                                    value.type.dimension == t.dimension):
            return self.cast(value, t, node)
        self.error(node, f"assigning {value.type.description} to a variable of type {t.description}")

    def s_AugAssign(self, node):
        cur = self.expr(node.target)
        rhs = self.expr(node.value)
        val = self.binop(node.op, cur, rhs, node)
        if isinstance(node.target, ast.Name):
            self.assign(node.target, val, node)
        else:
            self._mark_write(cur)
            self.push(A.AssignStmt(cur, self.convert_assign(val, cur.type, node)))

    def s_If(self, node):
        s = A.IfStmt(self.to_bool(self.expr(node.test), node))
        self.push(s)
        self.in_scope(s.true_branch, node.body)
        if node.orelse:
            self.in_scope(s.false_branch, node.orelse)

    def s_While(self, node):
        if node.orelse:
            self.error(node, "while ... else")
        loop = A.LoopStmt()
        self.push(loop)
        self.scope_stack.append(loop.body)
        try:
            cond = self.to_bool(self.expr(node.test), node)
            brk = A.IfStmt(A.UnaryExpr(A.BOOL, A.UnaryOp.NOT, cond))
            brk.true_branch.append(A.BreakStmt())
            self.push(brk)
        finally:
            self.scope_stack.pop()
        self.loop_depth += 1
        self.in_scope(loop.body, node.body)
        self.loop_depth -= 1

    def s_For(self, node):
        it = node.iter
        if not (isinstance(it, ast.Call) and isinstance(it.func, ast.Name) and it.func.id == "range"
                and isinstance(node.target, ast.Name) and not node.orelse and 1 <= len(it.args) <= 3):
            self.error(node, "for loops run over range(...)")
        args = [self.expr(a) for a in it.args]
        for a in args:
            if a.type is None or not a.type.is_int:
                self.error(node, "range() takes integers")
        start, stop, step = (_lit(A.INT, 0, True), args[0], _lit(A.INT, 1, True)) if len(args) == 1 else \
            (args[0], args[1], args[2] if len(args) > 2 else _lit(A.INT, 1, True))
        # the loop variable takes the bounds' type (literals adapt)
        t = start.type
        for a in (stop, step):
            if not _adaptive(a):
                t = a.type if _adaptive(start) and t is start.type else self._scalar_common(t, a.type, False, False)
        name = node.target.id
        v = self.lookup(name, node)
        if v is None:
            v = self.fn.local(t, name)
            self.names[name] = v
        elif not isinstance(v, A.Variable) or v.type is not t:
            self.error(node, f"loop variable {name} has another type")
        self.fn.mark(v, A.Usage.READ_WRITE)
        self.push(A.AssignStmt(A.RefExpr(v), self.cast(start, t, node)))
        stop_t, step_t = self.cast(stop, t, node), self.cast(step, t, node)
        ref = A.RefExpr(v)
        if isinstance(step, A.LiteralExpr):
            cond = A.BinaryExpr(A.BOOL, A.BinaryOp.LESS if step.value > 0 else A.BinaryOp.GREATER, ref, stop_t)
        else:
            zero = self.cast(_lit(A.INT, 0), t)
            up = A.BinaryExpr(A.BOOL, A.BinaryOp.GREATER, step_t, zero)
            cond = A.CallExpr(A.BOOL, CallOp.SELECT, [A.BinaryExpr(A.BOOL, A.BinaryOp.GREATER, ref, stop_t),
                                                      A.BinaryExpr(A.BOOL, A.BinaryOp.LESS, ref, stop_t), up])
        f = A.ForStmt(A.RefExpr(v), cond, step_t)
        self.push(f)
        self.loop_depth += 1
        self.in_scope(f.body, node.body)
        self.loop_depth -= 1

    def s_Break(self, node):
        if not self.loop_depth:
            self.error(node, "break outside a loop")
        self.push(A.BreakStmt())

    def s_Continue(self, node):
        if not self.loop_depth:
            self.error(node, "continue outside a loop")
        self.push(A.ContinueStmt())

    def s_Return(self, node):
        if self.kernel:
            if node.value is not None:
                self.error(node, "kernels return nothing")
            self.push(A.ReturnStmt())
            return
        if node.value is None:
            if self.fn.return_type is not None:
                self.error(node, "return without a value in a function returning one")
            self.push(A.ReturnStmt())
            self._returned_void = True
            return
        e = self.expr(node.value)
        if e is None or e.type is None:
            self.error(node, "returning void")
        if self.fn.return_type is None:
            if getattr(self, "_returned_void", False):
                self.error(node, "return with and without a value")
            self.fn.return_type = e.type
        else:
            e = self.convert_assign(e, self.fn.return_type, node)
        self.push(A.ReturnStmt(e))

    def s_FunctionDef(self, node):
        if node.decorator_list or node.args.vararg or node.args.kwarg or node.args.defaults or node.args.kwonlyargs:
            self.error(node, "nested functions take plain parameters")
        self.names[node.name] = _CallableDef(node, self)

    # -- expressions
    def expr(self, node):
        self.line = getattr(node, "lineno", self.line)
        m = getattr(self, "e_" + type(node).__name__, None)
        if m is None:
            self.error(node, f"unsupported expression {type(node).__name__}")
        return m(node)

    def e_Constant(self, node):
        v = node.value
        if isinstance(v, bool):
            return _lit(A.BOOL, v)
        if isinstance(v, int):
            if -2 ** 31 <= v < 2 ** 31:
                return _lit(A.INT, v, True)
            if 0 <= v < 2 ** 32:
                return _lit(A.UINT, v, True)
            if -2 ** 63 <= v < 2 ** 63:
                return _lit(A.LONG, v, True)
            if 0 <= v < 2 ** 64:
                return _lit(A.ULONG, v, True)
            self.error(node, "integer literal out of range")
        if isinstance(v, float):
            return _lit(A.FLOAT, v, True)
        self.error(node, f"unsupported constant {v!r}")

    def e_Name(self, node):
        v = self.lookup(node.id, node)
        if isinstance(v, A.Variable):
            return self.ref(v)
        if isinstance(v, _CallableDef):
            self.error(node, f"{node.id} is a function (call it)")
        if node.id == "FLT_MAX":
            return _lit(A.FLOAT, FLT_MAX, True)
        c = self.globals.get(node.id)
        if isinstance(c, (int, float)) and not isinstance(c, bool):
            return self.e_Constant(ast.Constant(c))
        self.error(node, f"unknown name {node.id}")

    def binop(self, op, a, b, node):
        if isinstance(op, ast.Div):  # true division: float operands
            ta = a.type if a.type.scalar_type.is_float else (Type.vector(A.FLOAT, a.type.dimension) if a.type.is_vector else A.FLOAT)
            tb = b.type if b.type.scalar_type.is_float else (Type.vector(A.FLOAT, b.type.dimension) if b.type.is_vector else A.FLOAT)
            a, b = self.cast(a, ta, node), self.cast(b, tb, node)
            t = self.common(a, b, node)
            return A.BinaryExpr(t, A.BinaryOp.DIV, self.cast(a, t, node), self.cast(b, t, node))
        if isinstance(op, ast.FloorDiv):
            t = self.common(a, b, node)
            q = A.BinaryExpr(t, A.BinaryOp.DIV, self.cast(a, t, node), self.cast(b, t, node))
            return A.CallExpr(t, CallOp.FLOOR, [q]) if t.scalar_type.is_float else q
        if isinstance(op, ast.Pow):
            t = self.common(a, b, node)
            if not t.scalar_type.is_float:
                t = A.FLOAT
            return A.CallExpr(t, CallOp.POW, [self.cast(a, t, node), self.cast(b, t, node)])
        bop = _BINOPS.get(type(op))
        if bop is None:
            self.error(node, f"unsupported operator {type(op).__name__}")
        if bop in (A.BinaryOp.SHL, A.BinaryOp.SHR):  # the shifted operand keeps its type
            if not a.type.scalar_type.is_int or not b.type.scalar_type.is_int:
                self.error(node, "shifts take integers")
            t = a.type
            if _adaptive(a) and not _adaptive(b):
                t = b.type
                a = self.cast(a, t, node)
            return A.BinaryExpr(t, bop, a, self.cast(b, t, node) if b.type.is_vector == t.is_vector else b)
        t = self.common(a, b, node)
        if bop in (A.BinaryOp.BIT_AND, A.BinaryOp.BIT_OR, A.BinaryOp.BIT_XOR, A.BinaryOp.MOD) and \
                not (t.scalar_type.is_int or t.scalar_type.is_bool):
            self.error(node, f"{type(op).__name__} takes integers")
        return A.BinaryExpr(t, bop, self.cast(a, t, node), self.cast(b, t, node))

    def e_BinOp(self, node):
        return self.binop(node.op, self.expr(node.left), self.expr(node.right), node)

    def e_UnaryOp(self, node):
        e = self.expr(node.operand)
        if isinstance(node.op, ast.Not):
            return A.UnaryExpr(A.BOOL, A.UnaryOp.NOT, self.to_bool(e, node))
        if isinstance(node.op, ast.USub):
            if isinstance(e, A.LiteralExpr) and not e.type.is_bool and not e.type.is_unsigned:
                return _lit(e.type, -e.value, _adaptive(e))
            return A.UnaryExpr(e.type, A.UnaryOp.MINUS, e)
        if isinstance(node.op, ast.UAdd):
            return e
        if isinstance(node.op, ast.Invert):
            return A.UnaryExpr(e.type, A.UnaryOp.BIT_NOT, e)
        self.error(node, "unsupported unary operator")

    def e_BoolOp(self, node):
        op = A.BinaryOp.AND if isinstance(node.op, ast.And) else A.BinaryOp.OR
        vals = [self.to_bool(self.expr(v), node) for v in node.values]
        e = vals[0]
        for v in vals[1:]:
            e = A.BinaryExpr(A.BOOL, op, e, v)
        return e

    def e_Compare(self, node):
        left = self.expr(node.left)
        parts = []
        for o, c in zip(node.ops, node.comparators):
            right = self.expr(c)
            bop = _CMPOPS.get(type(o))
            if bop is None:
                self.error(node, "unsupported comparison")
            t = self.common(left, right, node)
            if not t.is_scalar:
                self.error(node, "comparisons take scalars")
            parts.append(A.BinaryExpr(A.BOOL, bop, self.cast(left, t, node), self.cast(right, t, node)))
            left = right
        e = parts[0]
        for p in parts[1:]:
            e = A.BinaryExpr(A.BOOL, A.BinaryOp.AND, e, p)
        return e

    def e_IfExp(self, node):
        """c ? a : b as a temporary assigned in an if statement (only the branch taken runs)."""
        cond = self.to_bool(self.expr(node.test), node)
        s = A.IfStmt(cond)
        self.push(s)
        self.scope_stack.append(s.true_branch)
        try:
            a = self.expr(node.body)
        finally:
            self.scope_stack.pop()
        self.scope_stack.append(s.false_branch)
        try:
            b = self.expr(node.orelse)
        finally:
            self.scope_stack.pop()
        t = self.common(a, b, node)
        tmp = self.fn.local(t, f"sel{len(self.fn.local_variables)}")
        self.fn.mark(tmp, A.Usage.READ_WRITE)
        s.true_branch.append(A.AssignStmt(A.RefExpr(tmp), self.cast(a, t, node)))
        s.false_branch.append(A.AssignStmt(A.RefExpr(tmp), self.cast(b, t, node)))
        return A.RefExpr(tmp)

    def e_Attribute(self, node):
        base = self.expr(node.value)
        t = base.type
        if t is None:
            self.error(node, "member of void")
        if t.is_vector:
            lanes = [_SWIZZLE.get(ch) for ch in node.attr]
            if None in lanes or not 1 <= len(lanes) <= 4 or max(lanes) >= t.dimension:
                self.error(node, f"no lane {node.attr} in {t.description}")
            rt = t.element if len(lanes) == 1 else Type.vector(t.element, len(lanes))
            return A.MemberExpr(rt, base, swizzle=tuple(lanes))
        if t.is_structure:
            try:
                i = t.member_index(node.attr)
            except KeyError:
                self.error(node, f"no member {node.attr} in {getattr(t, 'struct_name', t.description)}")
            return A.MemberExpr(t.members[i], base, member=i)
        self.error(node, f"{t.description} has no members")

    def e_Subscript(self, node):
        base = self.expr(node.value)
        idx = self.expr(node.slice)
        t = base.type
        if t is None or not (t.is_vector or t.tag == Type.Tag.ARRAY):
            self.error(node, "indexing takes a vector or an array")
        if not idx.type.is_int:
            self.error(node, "index must be an integer")
        return A.AccessExpr(t.element, base, self.cast(idx, A.UINT, node))

    # -- calls
    def e_Call(self, node):
        if node.keywords:
            self.error(node, "keyword arguments are not supported")
        f = node.func
        if isinstance(f, ast.Attribute) and isinstance(f.value, ast.Name):
            v = self.lookup(f.value.id, node)
            if isinstance(v, A.Variable) and (v.type.is_resource or v.tag == A.Variable.Tag.SHARED):
                return self.method(v, f.attr, node)
        if isinstance(f, ast.Name):
            v = self.lookup(f.id, node)
            if isinstance(v, _CallableDef):
                return self.call_callable(v, [self.expr(a) for a in node.args], node)
            return self.builtin(f.id, node)
        self.error(node, "unsupported call")

    def builtin(self, name, node):
        if name in ("dispatch_id", "dispatch_size", "thread_id", "block_id", "warp_lane_id"):
            tag = {"dispatch_id": A.Variable.Tag.DISPATCH_ID, "dispatch_size": A.Variable.Tag.DISPATCH_SIZE,
                   "thread_id": A.Variable.Tag.THREAD_ID, "block_id": A.Variable.Tag.BLOCK_ID,
                   "warp_lane_id": A.Variable.Tag.WARP_LANE_ID}[name]
            return self.builtin_var(tag)
        if name in ("select", "synchronize_block", "warp_is_first_active_lane") or name in _CASTS or \
                name in _VEC_MAKERS or True:
            args = [self.expr(a) for a in node.args]
        if name in _VEC_MAKERS:
            el, n, op = _VEC_MAKERS[name]
            t = Type.vector(el, n)
            if len(args) == 1 and args[0].type.is_scalar:
                return A.CallExpr(t, op, [self.cast(args[0], el, node)] * n)
            flat = []
            for a in args:
                if a.type.is_vector:
                    flat += [A.MemberExpr(a.type.element, a, swizzle=(i,)) for i in range(a.type.dimension)]
                else:
                    flat.append(a)
            if len(flat) != n:
                self.error(node, f"{name} takes {n} lanes")
            return A.CallExpr(t, op, [self.cast(a, el, node) for a in flat])
        if name in _CASTS:
            if len(args) != 1:
                self.error(node, f"{name}() takes one value")
            t = _CASTS[name]
            if t.is_bool:
                return self.to_bool(args[0], node)
            return self.cast(args[0], t, node)
        if name in _FLOAT_UNARY:
            a = self._float(args, 1, name, node)
            return A.CallExpr(a[0].type, _FLOAT_UNARY[name], a)
        if name in _FLOAT_BINARY:
            a = self._float(args, 2, name, node)
            return A.CallExpr(a[0].type, _FLOAT_BINARY[name], a)
        if name in ("min", "max"):
            if len(args) != 2:
                self.error(node, f"{name} takes two values")
            t = self.common(args[0], args[1], node)
            return A.CallExpr(t, CallOp.MIN if name == "min" else CallOp.MAX, [self.cast(x, t, node) for x in args])
        if name == "abs":
            return A.CallExpr(args[0].type, CallOp.ABS, args)
        if name == "clamp":
            t = self.common(self.common(args[0], args[1], node) and args[0], args[2], node)
            t = self.common(A.LiteralExpr(t, 0) if False else args[0], args[1], node) if t is None else t
            return A.CallExpr(t, CallOp.CLAMP, [self.cast(x, t, node) for x in args])
        if name in ("lerp", "fma"):
            a = self._float(args, 3, name, node)
            return A.CallExpr(a[0].type, CallOp.LERP if name == "lerp" else CallOp.FMA, a)
        if name in ("isnan", "isinf"):
            a = self._float(args, 1, name, node)
            return A.CallExpr(A.BOOL, CallOp.ISNAN if name == "isnan" else CallOp.ISINF, a)
        if name in ("dot", "cross", "normalize", "length", "length_squared", "reduce_sum", "reduce_min",
                    "reduce_max"):
            n = {"dot": 2, "cross": 2}.get(name, 1)
            if len(args) != n or not all(a.type.is_vector and a.type.element is A.FLOAT for a in args):
                self.error(node, f"{name} takes {n} float vectors")
            op = {"dot": CallOp.DOT, "cross": CallOp.CROSS, "normalize": CallOp.NORMALIZE, "length": CallOp.LENGTH,
                  "length_squared": CallOp.LENGTH_SQUARED, "reduce_sum": CallOp.REDUCE_SUM,
                  "reduce_min": CallOp.REDUCE_MIN, "reduce_max": CallOp.REDUCE_MAX}[name]
            rt = args[0].type if name in ("cross", "normalize") else A.FLOAT
            return A.CallExpr(rt, op, args)
        if name == "select":  # select(c, a, b) = c ? a : b; the AST's argument order is (f, t, c)
            if len(args) != 3:
                self.error(node, "select(c, a, b)")
            t = self.common(args[1], args[2], node)
            return A.CallExpr(t, CallOp.SELECT, [self.cast(args[2], t, node), self.cast(args[1], t, node),
                                                 self.to_bool(args[0], node)])
        if name in ("synchronize_block", "sync_block"):  # sync_block: src/py/luisa/builtin.py:371
            return A.CallExpr(None, CallOp.SYNCHRONIZE_BLOCK, [])
        if name in ("warp_active_count_bits", "warp_prefix_count_bits"):
            op = CallOp.WARP_ACTIVE_COUNT_BITS if name == "warp_active_count_bits" else CallOp.WARP_PREFIX_COUNT_BITS
            return A.CallExpr(A.UINT, op, [self.to_bool(args[0], node)])
        if name == "warp_is_first_active_lane":
            return A.CallExpr(A.BOOL, CallOp.WARP_IS_FIRST_ACTIVE_LANE, [])
        if name == "warp_read_first_active_lane":
            return A.CallExpr(args[0].type, CallOp.WARP_READ_FIRST_ACTIVE_LANE, args)
        if name == "warp_read_lane":
            return A.CallExpr(args[0].type, CallOp.WARP_READ_LANE, [args[0], self.cast(args[1], A.UINT, node)])
        if name == "make_ray":
            if not 2 <= len(args) <= 4:
                self.error(node, "make_ray(origin, direction[, t_min, t_max])")
            full = args + [_lit(A.FLOAT, 0.0), _lit(A.FLOAT, FLT_MAX)][len(args) - 2:]
            ext = A.ExternalFunction("lcdsl::make_ray", A.RAY, (A.FLOAT3, A.FLOAT3, A.FLOAT, A.FLOAT))
            return A.CallExpr(A.RAY, CallOp.EXTERNAL, [self.cast(a, t, node) for a, t in zip(full, ext.argument_types)],
                              external=ext)
        if name in _EXTERNAL:
            cname, rt, ats = _EXTERNAL[name]
            if len(args) != len(ats):
                self.error(node, f"{name} takes {len(ats)} arguments")
            return A.CallExpr(rt, CallOp.EXTERNAL, [self.cast(a, t, node) for a, t in zip(args, ats)],
                              external=A.ExternalFunction(cname, rt, ats))
        self.error(node, f"unknown function {name}")

    def _float(self, args, n, name, node):
        if len(args) != n:
            self.error(node, f"{name} takes {n} arguments")
        t = args[0].type
        for a in args[1:]:
            t = self.common(A.RefExpr(A.Variable(t, A.Variable.Tag.LOCAL, -1)), a, node)
        if not t.scalar_type.is_float:
            t = A.FLOAT if t.is_scalar else Type.vector(A.FLOAT, t.dimension)
        return [self.cast(a, t, node) for a in args]

    def builtin_var(self, tag):
        if self.kernel:
            v = self.fn.builtin(tag)
            self.fn.mark(v, A.Usage.READ)
            return A.RefExpr(v)
        # a callable reads the kernel's builtin through an argument (the dispatch size lives in the
        # kernel's parameters)
        top = self
        while top.parent is not None:
            top = top.parent
        kv = top.fn.builtin(tag)
        return self.outer_expr(kv)

    # -- resource methods
    def method(self, v, m, node):
        t = v.type
        if t.tag == Type.Tag.ACCEL and m in ("query_all", "query_any"):
            self.fn.mark(v, A.Usage.READ)
            return self.ray_query(A.RefExpr(v), m == "query_any", node)
        # string arguments name types (bindless buffer_read)
        args = [None if isinstance(a, ast.Constant) and isinstance(a.value, str) else self.expr(a) for a in node.args]
        res = A.RefExpr(v)
        if v.tag == A.Variable.Tag.SHARED:
            el = t.element
            if m in _ATOMIC_METHODS and el.is_int and el.is_scalar:
                if len(args) != 2:
                    self.error(node, f"{m}(index, value)")
                return A.CallExpr(el, _ATOMIC_METHODS[m], [res, self._index(args, node), self.cast(args[1], el, node)])
            self.error(node, f"shared arrays of {el.description} have no method {m}")
        if t.tag == Type.Tag.BUFFER:
            el = t.element
            if m == "read":
                self.fn.mark(v, A.Usage.READ)
                return A.CallExpr(el, CallOp.BUFFER_READ, [res, self._index(args, node)])
            if m == "write":
                self.fn.mark(v, A.Usage.WRITE)
                if len(args) != 2:
                    self.error(node, "write(index, value)")
                return A.CallExpr(None, CallOp.BUFFER_WRITE, [res, self._index(args, node),
                                                              self.convert_assign(args[1], el, node)])
            if m == "size":
                return A.CallExpr(A.ULONG, CallOp.BUFFER_SIZE, [res])
            if m in _ATOMIC_METHODS:
                self.fn.mark(v, A.Usage.READ_WRITE)
                if len(args) != 2:
                    self.error(node, f"{m}(index, value)")
                return A.CallExpr(el, _ATOMIC_METHODS[m], [res, self._index(args, node), self.cast(args[1], el, node)])
            if m == "atomic_compare_exchange":
                self.fn.mark(v, A.Usage.READ_WRITE)
                return A.CallExpr(el, CallOp.ATOMIC_COMPARE_EXCHANGE,
                                  [res, self._index(args, node), self.cast(args[1], el, node), self.cast(args[2], el, node)])
        elif t.tag == Type.Tag.ACCEL:
            self.fn.mark(v, A.Usage.READ)
            if m in ("trace_closest", "trace_any"):
                if not 1 <= len(args) <= 2 or args[0].type is not A.RAY:
                    self.error(node, f"{m}(ray[, mask])")
                mask = self.cast(args[1], A.UINT, node) if len(args) > 1 else _lit(A.UINT, 255)
                op = CallOp.RAY_TRACING_TRACE_CLOSEST if m == "trace_closest" else CallOp.RAY_TRACING_TRACE_ANY
                return A.CallExpr(A.SURFACE_HIT if m == "trace_closest" else A.BOOL, op, [res, args[0], mask])
            if m in ("instance_user_id", "instance_visibility_mask"):
                op = CallOp.RAY_TRACING_INSTANCE_USER_ID if m == "instance_user_id" else \
                    CallOp.RAY_TRACING_INSTANCE_VISIBILITY_MASK
                return A.CallExpr(A.UINT, op, [res, self.cast(args[0], A.UINT, node)])
            if m in ("set_instance_user_id", "set_instance_visibility"):
                self.fn.mark(v, A.Usage.WRITE)
                op = CallOp.RAY_TRACING_SET_INSTANCE_USER_ID if m == "set_instance_user_id" else \
                    CallOp.RAY_TRACING_SET_INSTANCE_VISIBILITY
                return A.CallExpr(None, op, [res, self.cast(args[0], A.UINT, node), self.cast(args[1], A.UINT, node)])
        elif t.tag == Type.Tag.TEXTURE:
            dim = t.dimension
            coord_t = Type.vector(A.UINT, dim)
            if m == "read":
                self.fn.mark(v, A.Usage.READ)
                if len(args) != dim:
                    self.error(node, f"read takes {dim} coordinates")
                return A.CallExpr(A.FLOAT4, CallOp.TEXTURE_READ, [res, self._coords(args, coord_t, node)])
            if m == "write":
                self.fn.mark(v, A.Usage.WRITE)
                if len(args) != dim + 1:
                    self.error(node, f"write takes {dim} coordinates and a value")
                return A.CallExpr(None, CallOp.TEXTURE_WRITE, [res, self._coords(args[:dim], coord_t, node),
                                                               self.cast(args[dim], A.FLOAT4, node)])
        elif t.tag == Type.Tag.BINDLESS_ARRAY:
            self.fn.mark(v, A.Usage.READ)
            if m == "tex2d_sample":
                return A.CallExpr(A.FLOAT4, CallOp.BINDLESS_TEXTURE2D_SAMPLE,
                                  [res, self.cast(args[0], A.UINT, node), self.cast(args[1], A.FLOAT2, node)])
            if m == "tex2d_read":
                return A.CallExpr(A.FLOAT4, CallOp.BINDLESS_TEXTURE2D_READ,
                                  [res, self.cast(args[0], A.UINT, node), self._coords(args[1:3], A.UINT2, node)])
            if m == "tex2d_size":
                return A.CallExpr(A.UINT2, CallOp.BINDLESS_TEXTURE2D_SIZE, [res, self.cast(args[0], A.UINT, node)])
            if m == "buffer_read":
                el = node.args[0]
                if not (isinstance(el, ast.Constant) and isinstance(el.value, str)):
                    self.error(node, "buffer_read takes the element type as a string")
                et = _element(el.value)
                return A.CallExpr(et, CallOp.BINDLESS_BUFFER_READ,
                                  [res, self.cast(args[1], A.UINT, node), self.cast(args[2], A.ULONG, node)])
        self.error(node, f"unknown method {m} of parameter {v.name}")

    def _index(self, args, node):
        if not args or args[0].type is None or not args[0].type.is_int:
            self.error(node, "index must be an integer")
        return self.cast(args[0], A.ULONG if args[0].type.size == 8 else A.UINT, node)

    def _coords(self, args, t, node):
        return A.CallExpr(t, _maker(t), [self.cast(a, A.UINT, node) for a in args])

    def ray_query(self, accel, any_hit, node):
        """accel.traverse(...): a query local, its RayQueryStmt (callbacks as callables called in
        the candidate scopes), then the committed hit."""
        if len(node.args) != 4:
            self.error(node, "query_all(ray, mask, on_surface, on_procedural)")
        ray = self.expr(node.args[0])
        if ray.type is not A.RAY:
            self.error(node, "query_all takes a Ray")
        mask = self.cast(self.expr(node.args[1]), A.UINT, node)
        qt = A.RAY_QUERY_ANY if any_hit else A.RAY_QUERY_ALL
        q = self.fn.local(qt, f"query{len(self.fn.local_variables)}")
        self.fn.mark(q, A.Usage.READ_WRITE)
        op = CallOp.RAY_TRACING_QUERY_ANY if any_hit else CallOp.RAY_TRACING_QUERY_ALL
        self.push(A.AssignStmt(A.RefExpr(q), A.CallExpr(qt, op, [accel, ray, mask])))
        rq = A.RayQueryStmt(A.RefExpr(q))
        callbacks = []
        for arg in node.args[2:]:
            if isinstance(arg, ast.Constant) and arg.value is None:
                callbacks.append(None)
                continue
            cb = self.lookup(arg.id, node) if isinstance(arg, ast.Name) else None
            if not isinstance(cb, _CallableDef):
                self.error(node, "query callbacks are nested functions of one parameter (or None)")
            callbacks.append(cb)
        self.push(rq)
        if callbacks[0] is not None:
            self.scope_stack.append(rq.on_triangle_candidate)
            try:
                cand = A.CallExpr(A.SURFACE_CANDIDATE, CallOp.RAY_QUERY_TRIANGLE_CANDIDATE_HIT, [A.RefExpr(q)])
                ok = self.to_bool(self.call_callable(callbacks[0], [cand], node), node)
                s = A.IfStmt(ok)
                s.true_branch.append(A.ExprStmt(A.CallExpr(None, CallOp.RAY_QUERY_COMMIT_TRIANGLE, [A.RefExpr(q)])))
                self.push(s)
            finally:
                self.scope_stack.pop()
        if callbacks[1] is not None:
            self.scope_stack.append(rq.on_procedural_candidate)
            try:
                cand = A.CallExpr(A.PROCEDURAL_CANDIDATE, CallOp.RAY_QUERY_PROCEDURAL_CANDIDATE_HIT, [A.RefExpr(q)])
                t = self.cast(self.call_callable(callbacks[1], [cand], node), A.FLOAT, node)
                self.push(A.ExprStmt(A.CallExpr(None, CallOp.RAY_QUERY_COMMIT_PROCEDURAL, [A.RefExpr(q), t])))
            finally:
                self.scope_stack.pop()
        return A.CallExpr(A.COMMITTED_HIT, CallOp.RAY_QUERY_COMMITTED_HIT, [A.RefExpr(q)])

    # -- callables
    def call_callable(self, cd, args, node):
        """Compiles the nested function for these argument types (once) and calls it; variables it
        shares with enclosing functions follow the explicit arguments."""
        if len(args) != len(cd.node.args.args):
            self.error(node, f"{cd.node.name} takes {len(cd.node.args.args)} arguments")
        for a in args:
            if a.type is None:
                self.error(node, "void argument")
        key = tuple(a.type.description for a in args)
        inst = cd.instances.get(key)
        if inst is None:
            fn = A.Function(A.Function.Tag.CALLABLE, cd.node.name)
            b = _Builder(fn, False, self.globals, parent=cd.owner, fname=f"{self.fname}.{cd.node.name}")
            for p, a in zip(cd.node.args.args, args):
                b.names[p.arg] = fn.argument(a.type, p.arg)
            cd.instances[key] = (fn, b)  # (recursion is rejected below)
            b._compiling = True
            for st in cd.node.body:
                b.stmt(st)
            b._compiling = False
            inst = (fn, b)
        fn, b = inst
        if getattr(b, "_compiling", False):
            self.error(node, f"recursive call of {cd.node.name}")
        own = self.fn
        if fn not in own.custom_callables:
            own.custom_callables.append(fn)
        call_args = list(args)
        for outer, _ in b.captures:
            e = self.outer_expr(outer) if outer not in self.fn.builtin_variables else A.RefExpr(outer)
            if isinstance(e, A.RefExpr):
                self.fn.mark(e.variable, A.Usage.READ_WRITE if _ is not None and _.tag == A.Variable.Tag.REFERENCE
                             else A.Usage.READ)
            call_args.append(e)
        return A.CallExpr(fn.return_type, CallOp.CUSTOM, call_args, custom=fn)


def _maker(t):
    return {A.FLOAT2: CallOp.MAKE_FLOAT2, A.FLOAT3: CallOp.MAKE_FLOAT3, A.FLOAT4: CallOp.MAKE_FLOAT4,
            A.UINT2: CallOp.MAKE_UINT2, A.UINT3: CallOp.MAKE_UINT3, A.UINT4: CallOp.MAKE_UINT4,
            A.INT2: CallOp.MAKE_INT2, A.INT3: CallOp.MAKE_INT3, A.INT4: CallOp.MAKE_INT4}[t]


# ---- kernels ------------------------------------------------------------------------------------

class Kernel:
    """A DSL kernel: `function` is its typed AST (lcast.Function), `source` the HIP text the
    codegen prints; compile(device) returns a callable that packs the arguments in parameter order
    (uniforms as the annotated type) into a shader dispatch command."""

    def __init__(self, fn, block=(128, 1, 1)):
        self.fn, self.block, self.name = fn, tuple(block), fn.__name__
        tree = ast.parse(textwrap.dedent(inspect.getsource(fn)))
        fdef = tree.body[0]
        hints = fn.__annotations__
        self.function = A.Function(A.Function.Tag.KERNEL, self.name, self.block)
        self.params = {}
        b = _Builder(self.function, True, fn.__globals__, fname=self.name)
        for a in fdef.args.args:
            ann = hints.get(a.arg)
            if not isinstance(ann, (Buffer, _Annotation)):
                raise LCError(f"dsl: parameter {a.arg} needs a Buffer / Accel / Image / Volume / BindlessArray / "
                              f"uniform annotation")
            self.params[a.arg] = ann
            b.names[a.arg] = self.function.argument(ann.type, a.arg)
        for st in fdef.body:
            b.stmt(st)
        self.source = codegen.emit(self.function)

    def argument_usage(self, index):
        """Usage (read / write) of argument `index` (DeviceInterface::shader_argument_usage)."""
        return self.function.argument_usage(index)

    def compile(self, device):
        shader = device.create_shader(self.source, self.name, self.block)
        usages = [int(self.argument_usage(i)) for i in range(len(self.params))]
        if hasattr(shader, "set_argument_usage"):
            shader.set_argument_usage(usages)
        params = list(self.params.values())

        def bind(*args):
            if len(args) != len(params):
                raise LCError(f"{self.name} takes {len(params)} arguments")
            packed = [np.asarray(a, p.dtype) if isinstance(p, _Annotation) and p.dtype is not None else a
                      for a, p in zip(args, params)]
            return shader(*packed)
        bind.shader = shader
        return bind


def kernel(fn=None, block=(128, 1, 1)):
    """Decorator: @kernel or @kernel(block=(bx, by, bz))."""
    if fn is None:
        return lambda f: Kernel(f, block)
    return Kernel(fn, block)
