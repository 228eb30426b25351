"""The kernel AST: types, variables, expressions, statements and functions shaped like the
reference's C++ AST (include/luisa/ast/{type,variable,expression,statement,function,op}.h), so a
program is a typed tree that a backend prints (codegen.py) instead of text assembled on the fly.

What is kept alike, and why: the tag enumerations (Type.Tag, Variable.Tag, Usage, Expression.Tag,
Statement.Tag, UnaryOp, BinaryOp and CallOp in the order of op.h) are the AST's interface -- a
program or a tool that reads CallOp values sees the same numbers. Type descriptions use the
reference's spelling ("float", "vector<float,3>", "array<float,3>", "struct<16,uint,uint,...>",
"buffer<...>", "accel"). Everything else (how the tree is built and checked) is this package's.

Each Function records, per argument, whether the body reads or writes it (Usage, as
Function::variable_usage does), which is what shader_argument_usage reports.
"""
import enum
import hashlib


# ---- enumerations (values in the reference's order) --------------------------------------------

class Usage(enum.IntFlag):
    """include/luisa/ast/usage.h"""
    NONE = 0
    READ = 1
    WRITE = 2
    READ_WRITE = 3


class UnaryOp(enum.IntEnum):
    PLUS = 0
    MINUS = 1
    NOT = 2
    BIT_NOT = 3


class BinaryOp(enum.IntEnum):
    ADD = 0
    SUB = 1
    MUL = 2
    DIV = 3
    MOD = 4
    BIT_AND = 5
    BIT_OR = 6
    BIT_XOR = 7
    SHL = 8
    SHR = 9
    AND = 10
    OR = 11
    LESS = 12
    GREATER = 13
    LESS_EQUAL = 14
    GREATER_EQUAL = 15
    EQUAL = 16
    NOT_EQUAL = 17


_CALL_OPS = """
CUSTOM EXTERNAL ALL ANY SELECT CLAMP SATURATE LERP SMOOTHSTEP STEP ABS MIN MAX CLZ CTZ POPCOUNT REVERSE
ISINF ISNAN ACOS ACOSH ASIN ASINH ATAN ATAN2 ATANH COS COSH SIN SINH TAN TANH EXP EXP2 EXP10 LOG LOG2 LOG10
POW SQRT RSQRT CEIL FLOOR FRACT TRUNC ROUND FMA COPYSIGN CROSS DOT LENGTH LENGTH_SQUARED NORMALIZE
FACEFORWARD REFLECT REDUCE_SUM REDUCE_PRODUCT REDUCE_MIN REDUCE_MAX OUTER_PRODUCT
MATRIX_COMPONENT_WISE_MULTIPLICATION DETERMINANT TRANSPOSE INVERSE SYNCHRONIZE_BLOCK ATOMIC_EXCHANGE
ATOMIC_COMPARE_EXCHANGE ATOMIC_FETCH_ADD ATOMIC_FETCH_SUB ATOMIC_FETCH_AND ATOMIC_FETCH_OR ATOMIC_FETCH_XOR
ATOMIC_FETCH_MIN ATOMIC_FETCH_MAX ADDRESS_OF BUFFER_READ BUFFER_WRITE BUFFER_SIZE BUFFER_ADDRESS
BYTE_BUFFER_READ BYTE_BUFFER_WRITE BYTE_BUFFER_SIZE TEXTURE_READ TEXTURE_WRITE TEXTURE_SIZE
BINDLESS_TEXTURE2D_SAMPLE BINDLESS_TEXTURE2D_SAMPLE_LEVEL BINDLESS_TEXTURE2D_SAMPLE_GRAD
BINDLESS_TEXTURE2D_SAMPLE_GRAD_LEVEL BINDLESS_TEXTURE3D_SAMPLE BINDLESS_TEXTURE3D_SAMPLE_LEVEL
BINDLESS_TEXTURE3D_SAMPLE_GRAD BINDLESS_TEXTURE3D_SAMPLE_GRAD_LEVEL BINDLESS_TEXTURE2D_SAMPLE_SAMPLER
BINDLESS_TEXTURE2D_SAMPLE_LEVEL_SAMPLER BINDLESS_TEXTURE2D_SAMPLE_GRAD_SAMPLER
BINDLESS_TEXTURE2D_SAMPLE_GRAD_LEVEL_SAMPLER BINDLESS_TEXTURE3D_SAMPLE_SAMPLER
BINDLESS_TEXTURE3D_SAMPLE_LEVEL_SAMPLER BINDLESS_TEXTURE3D_SAMPLE_GRAD_SAMPLER
BINDLESS_TEXTURE3D_SAMPLE_GRAD_LEVEL_SAMPLER BINDLESS_TEXTURE2D_READ BINDLESS_TEXTURE3D_READ
BINDLESS_TEXTURE2D_READ_LEVEL BINDLESS_TEXTURE3D_READ_LEVEL BINDLESS_TEXTURE2D_SIZE BINDLESS_TEXTURE3D_SIZE
BINDLESS_TEXTURE2D_SIZE_LEVEL BINDLESS_TEXTURE3D_SIZE_LEVEL BINDLESS_BUFFER_READ BINDLESS_BUFFER_WRITE
BINDLESS_BYTE_BUFFER_READ BINDLESS_BUFFER_SIZE BINDLESS_BUFFER_TYPE BINDLESS_BUFFER_ADDRESS
MAKE_BOOL2 MAKE_BOOL3 MAKE_BOOL4 MAKE_INT2 MAKE_INT3 MAKE_INT4 MAKE_UINT2 MAKE_UINT3 MAKE_UINT4
MAKE_FLOAT2 MAKE_FLOAT3 MAKE_FLOAT4 MAKE_SHORT2 MAKE_SHORT3 MAKE_SHORT4 MAKE_USHORT2 MAKE_USHORT3 MAKE_USHORT4
MAKE_LONG2 MAKE_LONG3 MAKE_LONG4 MAKE_ULONG2 MAKE_ULONG3 MAKE_ULONG4 MAKE_HALF2 MAKE_HALF3 MAKE_HALF4
MAKE_DOUBLE2 MAKE_DOUBLE3 MAKE_DOUBLE4 MAKE_BYTE2 MAKE_BYTE3 MAKE_BYTE4 MAKE_UBYTE2 MAKE_UBYTE3 MAKE_UBYTE4
MAKE_FLOAT2X2 MAKE_FLOAT3X3 MAKE_FLOAT4X4 ASSERT ASSUME UNREACHABLE ZERO ONE PACK UNPACK REQUIRES_GRADIENT
GRADIENT GRADIENT_MARKER ACCUMULATE_GRADIENT BACKWARD DETACH RAY_TRACING_INSTANCE_TRANSFORM
RAY_TRACING_INSTANCE_USER_ID RAY_TRACING_INSTANCE_VISIBILITY_MASK RAY_TRACING_SET_INSTANCE_TRANSFORM
RAY_TRACING_SET_INSTANCE_VISIBILITY RAY_TRACING_SET_INSTANCE_OPACITY RAY_TRACING_SET_INSTANCE_USER_ID
RAY_TRACING_TRACE_CLOSEST RAY_TRACING_TRACE_ANY RAY_TRACING_QUERY_ALL RAY_TRACING_QUERY_ANY
RAY_TRACING_INSTANCE_MOTION_MATRIX RAY_TRACING_INSTANCE_MOTION_SRT RAY_TRACING_SET_INSTANCE_MOTION_MATRIX
RAY_TRACING_SET_INSTANCE_MOTION_SRT RAY_TRACING_TRACE_CLOSEST_MOTION_BLUR RAY_TRACING_TRACE_ANY_MOTION_BLUR
RAY_TRACING_QUERY_ALL_MOTION_BLUR RAY_TRACING_QUERY_ANY_MOTION_BLUR RAY_QUERY_WORLD_SPACE_RAY
RAY_QUERY_PROCEDURAL_CANDIDATE_HIT RAY_QUERY_TRIANGLE_CANDIDATE_HIT RAY_QUERY_COMMITTED_HIT
RAY_QUERY_COMMIT_TRIANGLE RAY_QUERY_COMMIT_PROCEDURAL RAY_QUERY_TERMINATE RAY_QUERY_PROCEED
RAY_QUERY_IS_TRIANGLE_CANDIDATE RAY_QUERY_IS_PROCEDURAL_CANDIDATE RASTER_DISCARD DDX DDY
WARP_IS_FIRST_ACTIVE_LANE WARP_FIRST_ACTIVE_LANE WARP_ACTIVE_ALL_EQUAL WARP_ACTIVE_BIT_AND
WARP_ACTIVE_BIT_OR WARP_ACTIVE_BIT_XOR WARP_ACTIVE_COUNT_BITS WARP_ACTIVE_MAX WARP_ACTIVE_MIN
WARP_ACTIVE_PRODUCT WARP_ACTIVE_SUM WARP_ACTIVE_ALL WARP_ACTIVE_ANY WARP_ACTIVE_BIT_MASK
WARP_PREFIX_COUNT_BITS WARP_PREFIX_SUM WARP_PREFIX_PRODUCT WARP_READ_LANE WARP_READ_FIRST_ACTIVE_LANE
INDIRECT_SET_DISPATCH_KERNEL INDIRECT_SET_DISPATCH_COUNT TEXTURE2D_SAMPLE TEXTURE2D_SAMPLE_LEVEL
TEXTURE2D_SAMPLE_GRAD TEXTURE2D_SAMPLE_GRAD_LEVEL TEXTURE3D_SAMPLE TEXTURE3D_SAMPLE_LEVEL
TEXTURE3D_SAMPLE_GRAD TEXTURE3D_SAMPLE_GRAD_LEVEL SHADER_EXECUTION_REORDER
""".split()

# CallOp: the builtin functions and intrinsics of op.h, numbered in its order
CallOp = enum.IntEnum("CallOp", [(n, i) for i, n in enumerate(_CALL_OPS)])


class CastOp(enum.IntEnum):
    STATIC = 0
    BITWISE = 1


# ---- types --------------------------------------------------------------------------------------

class Type:
    """A type of the AST (include/luisa/ast/type.h). Interned: Type.of(description) returns the
    same object for the same description, so types compare with `is` and `==`."""

    class Tag(enum.IntEnum):
        BOOL = 0
        INT8 = 1
        UINT8 = 2
        INT16 = 3
        UINT16 = 4
        INT32 = 5
        UINT32 = 6
        INT64 = 7
        UINT64 = 8
        FLOAT16 = 9
        FLOAT32 = 10
        FLOAT64 = 11
        VECTOR = 12
        MATRIX = 13
        ARRAY = 14
        STRUCTURE = 15
        BUFFER = 16
        TEXTURE = 17
        BINDLESS_ARRAY = 18
        ACCEL = 19
        CUSTOM = 20

    _SCALARS = {"bool": (Tag.BOOL, 1), "byte": (Tag.INT8, 1), "ubyte": (Tag.UINT8, 1), "short": (Tag.INT16, 2),
                "ushort": (Tag.UINT16, 2), "int": (Tag.INT32, 4), "uint": (Tag.UINT32, 4), "long": (Tag.INT64, 8),
                "ulong": (Tag.UINT64, 8), "half": (Tag.FLOAT16, 2), "float": (Tag.FLOAT32, 4),
                "double": (Tag.FLOAT64, 8)}
    _interned = {}

    def __init__(self, tag, description, size, alignment, element=None, dimension=1, members=(), names=()):
        self.tag, self.description, self.size, self.alignment = tag, description, size, alignment
        self.element, self.dimension, self.members, self.member_names = element, dimension, tuple(members), tuple(names)

    def __repr__(self):
        return f"Type({self.description})"

    # construction
    @classmethod
    def _intern(cls, desc, make):
        t = cls._interned.get(desc)
        if t is None:
            t = make()
            cls._interned[desc] = t
        return t

    @classmethod
    def scalar(cls, name):
        tag, size = cls._SCALARS[name]
        return cls._intern(name, lambda: cls(tag, name, size, size))

    @classmethod
    def vector(cls, element, n):
        if not element.is_scalar or not 2 <= n <= 4:
            raise TypeError(f"vector<{element.description},{n}>")
        desc = f"vector<{element.description},{n}>"
        lanes = 4 if n == 3 else n
        return cls._intern(desc, lambda: cls(cls.Tag.VECTOR, desc, element.size * lanes, element.size * lanes,
                                             element, n))

    @classmethod
    def matrix(cls, n):
        desc = f"matrix<{n}>"
        col = cls.vector(cls.scalar("float"), n)
        return cls._intern(desc, lambda: cls(cls.Tag.MATRIX, desc, col.size * n, col.alignment, col, n))

    @classmethod
    def array(cls, element, n):
        desc = f"array<{element.description},{n}>"
        return cls._intern(desc, lambda: cls(cls.Tag.ARRAY, desc, element.size * n, element.alignment, element, n))

    @classmethod
    def structure(cls, alignment, members, names=()):
        """struct<alignment,member,...>; member names are this package's (the description, like the
        reference's, has none)."""
        desc = f"struct<{alignment}," + ",".join(m.description for m in members) + ">"
        size = 0
        for m in members:
            size = (size + m.alignment - 1) // m.alignment * m.alignment + m.size
        size = (size + alignment - 1) // alignment * alignment
        t = cls._intern(desc, lambda: cls(cls.Tag.STRUCTURE, desc, size, alignment, None, len(members), members,
                                          names))
        return t

    @classmethod
    def buffer(cls, element):
        desc = f"buffer<{element.description}>"
        return cls._intern(desc, lambda: cls(cls.Tag.BUFFER, desc, 8, 8, element))

    @classmethod
    def texture(cls, dimension, element):
        desc = f"texture<{dimension},{element.description}>"
        return cls._intern(desc, lambda: cls(cls.Tag.TEXTURE, desc, 8, 8, element, dimension))

    @classmethod
    def bindless_array(cls):
        return cls._intern("bindless_array", lambda: cls(cls.Tag.BINDLESS_ARRAY, "bindless_array", 8, 8))

    @classmethod
    def accel(cls):
        return cls._intern("accel", lambda: cls(cls.Tag.ACCEL, "accel", 8, 8))

    @classmethod
    def custom(cls, name):
        return cls._intern(name, lambda: cls(cls.Tag.CUSTOM, name, 0, 16))

    @classmethod
    def of(cls, desc):
        """Parses a type description ("float", "vector<uint,3>", "buffer<float>", ...)."""
        desc = desc.replace(" ", "")
        if desc in cls._interned:
            return cls._interned[desc]
        if desc in cls._SCALARS:
            return cls.scalar(desc)
        if desc == "accel":
            return cls.accel()
        if desc == "bindless_array":
            return cls.bindless_array()
        head, _, rest = desc.partition("<")
        if not rest.endswith(">"):
            raise TypeError(f"bad type description {desc!r}")
        args = _split_args(rest[:-1])
        if head == "vector":
            return cls.vector(cls.of(args[0]), int(args[1]))
        if head == "matrix":
            return cls.matrix(int(args[0]))
        if head == "array":
            return cls.array(cls.of(args[0]), int(args[1]))
        if head == "struct":
            return cls.structure(int(args[0]), [cls.of(a) for a in args[1:]])
        if head == "buffer":
            return cls.buffer(cls.of(args[0]))
        if head == "texture":
            return cls.texture(int(args[0]), cls.of(args[1]))
        raise TypeError(f"bad type description {desc!r}")

    # queries
    @property
    def is_scalar(self):
        return self.tag <= Type.Tag.FLOAT64

    @property
    def is_bool(self):
        return self.tag == Type.Tag.BOOL

    @property
    def is_float(self):
        return self.tag in (Type.Tag.FLOAT16, Type.Tag.FLOAT32, Type.Tag.FLOAT64)

    @property
    def is_int(self):
        return Type.Tag.INT8 <= self.tag <= Type.Tag.UINT64

    @property
    def is_unsigned(self):
        return self.tag in (Type.Tag.UINT8, Type.Tag.UINT16, Type.Tag.UINT32, Type.Tag.UINT64)

    @property
    def is_vector(self):
        return self.tag == Type.Tag.VECTOR

    @property
    def is_structure(self):
        return self.tag == Type.Tag.STRUCTURE

    @property
    def is_resource(self):
        return self.tag in (Type.Tag.BUFFER, Type.Tag.TEXTURE, Type.Tag.BINDLESS_ARRAY, Type.Tag.ACCEL)

    @property
    def scalar_type(self):
        """The lane type of a scalar or vector."""
        return self.element if self.is_vector else self

    def member_index(self, name):
        if name not in self.member_names:
            raise KeyError(name)
        return self.member_names.index(name)


def _split_args(s):
    out, depth, cur = [], 0, ""
    for ch in s:
        if ch == "," and depth == 0:
            out.append(cur)
            cur = ""
            continue
        depth += ch == "<"
        depth -= ch == ">"
        cur += ch
    out.append(cur)
    return out


def named_struct(name, alignment, fields):
    """A structure type with member names (fields: [(name, Type)]), registered under `name`. Its
    description is the layout's, as in the reference; the object is its own (two structures of
    one layout stay distinct, so the codegen can name each)."""
    layout = Type.structure(alignment, [f[1] for f in fields])
    t = Type(Type.Tag.STRUCTURE, layout.description, layout.size, alignment, None, len(fields),
             [f[1] for f in fields], [f[0] for f in fields])
    t.struct_name = name
    NAMED_STRUCTS[name] = t
    return t


NAMED_STRUCTS = {}

BOOL, INT, UINT, LONG, ULONG, FLOAT = (Type.scalar(n) for n in ("bool", "int", "uint", "long", "ulong", "float"))
FLOAT2, FLOAT3, FLOAT4 = (Type.vector(FLOAT, n) for n in (2, 3, 4))
UINT2, UINT3, UINT4 = (Type.vector(UINT, n) for n in (2, 3, 4))
INT2, INT3, INT4 = (Type.vector(INT, n) for n in (2, 3, 4))

# structures of the ray-tracing library (include/luisa/runtime/rtx/ray.h, hit.h): the layouts of
# the device structs the codegen names
RAY = named_struct("Ray", 16, [("origin", Type.array(FLOAT, 3)), ("t_min", FLOAT),
                               ("direction", Type.array(FLOAT, 3)), ("t_max", FLOAT)])
SURFACE_HIT = named_struct("SurfaceHit", 8, [("inst", UINT), ("prim", UINT), ("bary", FLOAT2),
                                             ("committed_ray_t", FLOAT)])
COMMITTED_HIT = named_struct("CommittedHit", 8, [("inst", UINT), ("prim", UINT), ("u", FLOAT), ("v", FLOAT),
                                                 ("type", UINT), ("t", FLOAT)])
# a ray query's candidates as its callbacks see them
SURFACE_CANDIDATE = named_struct("SurfaceCandidate", 8, [("inst", UINT), ("prim", UINT), ("bary", FLOAT2),
                                                         ("t", FLOAT)])
PROCEDURAL_CANDIDATE = named_struct("ProceduralCandidate", 16, [
    ("inst", UINT), ("prim", UINT), ("origin", FLOAT3), ("direction", FLOAT3), ("aabb_min", FLOAT3),
    ("aabb_max", FLOAT3), ("t_max", FLOAT)])
RAY_QUERY_ALL = Type.custom("LC_RayQueryAll")
RAY_QUERY_ANY = Type.custom("LC_RayQueryAny")


# ---- variables ----------------------------------------------------------------------------------

class Variable:
    """include/luisa/ast/variable.h"""

    class Tag(enum.IntEnum):
        LOCAL = 0
        SHARED = 1
        REFERENCE = 2
        BUFFER = 3
        TEXTURE = 4
        BINDLESS_ARRAY = 5
        ACCEL = 6
        THREAD_ID = 7
        BLOCK_ID = 8
        DISPATCH_ID = 9
        DISPATCH_SIZE = 10
        KERNEL_ID = 11
        WARP_LANE_COUNT = 12
        WARP_LANE_ID = 13
        OBJECT_ID = 14

    def __init__(self, type_, tag, uid, name=None):
        self.type, self.tag, self.uid, self.name = type_, tag, uid, name

    def __repr__(self):
        return f"Variable({self.name or self.uid}: {self.type.description}, {self.tag.name})"

    @property
    def is_resource(self):
        return self.tag in (Variable.Tag.BUFFER, Variable.Tag.TEXTURE, Variable.Tag.BINDLESS_ARRAY, Variable.Tag.ACCEL)

    @property
    def is_builtin(self):
        return self.tag >= Variable.Tag.THREAD_ID


# ---- expressions --------------------------------------------------------------------------------

class Expression:
    """include/luisa/ast/expression.h; every expression has a type (None for void calls)."""

    class Tag(enum.IntEnum):
        UNARY = 0
        BINARY = 1
        MEMBER = 2
        ACCESS = 3
        LITERAL = 4
        REF = 5
        CONSTANT = 6
        CALL = 7
        CAST = 8
        TYPE_ID = 9
        STRING_ID = 10
        FUNC_REF = 11
        CPUCUSTOM = 12
        GPUCUSTOM = 13

    tag = None

    def __init__(self, type_):
        self.type = type_


class UnaryExpr(Expression):
    tag = Expression.Tag.UNARY

    def __init__(self, type_, op, operand):
        super().__init__(type_)
        self.op, self.operand = op, operand


class BinaryExpr(Expression):
    tag = Expression.Tag.BINARY

    def __init__(self, type_, op, lhs, rhs):
        super().__init__(type_)
        self.op, self.lhs, self.rhs = op, lhs, rhs


class MemberExpr(Expression):
    """A structure member (member index) or a vector swizzle (swizzle: lane indices)."""
    tag = Expression.Tag.MEMBER

    def __init__(self, type_, self_, member=None, swizzle=None):
        super().__init__(type_)
        self.self, self.member, self.swizzle = self_, member, swizzle

    @property
    def is_swizzle(self):
        return self.swizzle is not None


class AccessExpr(Expression):
    tag = Expression.Tag.ACCESS

    def __init__(self, type_, range_, index):
        super().__init__(type_)
        self.range, self.index = range_, index


class LiteralExpr(Expression):
    tag = Expression.Tag.LITERAL

    def __init__(self, type_, value):
        super().__init__(type_)
        self.value = value


class RefExpr(Expression):
    tag = Expression.Tag.REF

    def __init__(self, variable):
        super().__init__(variable.type)
        self.variable = variable


class ExternalFunction:
    """A function the device library provides (CallOp::EXTERNAL): name and signature."""

    def __init__(self, name, return_type, argument_types):
        self.name, self.return_type, self.argument_types = name, return_type, tuple(argument_types)


class CallExpr(Expression):
    """A builtin (op), a callable (op CUSTOM, custom = the Function) or a device-library function
    (op EXTERNAL, external = the ExternalFunction)."""
    tag = Expression.Tag.CALL

    def __init__(self, type_, op, arguments, custom=None, external=None):
        super().__init__(type_)
        self.op, self.arguments, self.custom, self.external = op, list(arguments), custom, external


class CastExpr(Expression):
    tag = Expression.Tag.CAST

    def __init__(self, type_, op, expression):
        super().__init__(type_)
        self.op, self.expression = op, expression


# ---- statements ---------------------------------------------------------------------------------

class Statement:
    """include/luisa/ast/statement.h"""

    class Tag(enum.IntEnum):
        BREAK = 0
        CONTINUE = 1
        RETURN = 2
        SCOPE = 3
        IF = 4
        LOOP = 5
        EXPR = 6
        SWITCH = 7
        SWITCH_CASE = 8
        SWITCH_DEFAULT = 9
        ASSIGN = 10
        FOR = 11
        COMMENT = 12
        RAY_QUERY = 13
        AUTO_DIFF = 14
        PRINT = 15

    tag = None


class BreakStmt(Statement):
    tag = Statement.Tag.BREAK


class ContinueStmt(Statement):
    tag = Statement.Tag.CONTINUE


class ReturnStmt(Statement):
    tag = Statement.Tag.RETURN

    def __init__(self, expression=None):
        self.expression = expression


class ScopeStmt(Statement):
    tag = Statement.Tag.SCOPE

    def __init__(self):
        self.statements = []

    def append(self, s):
        self.statements.append(s)
        return s


class IfStmt(Statement):
    tag = Statement.Tag.IF

    def __init__(self, condition):
        self.condition, self.true_branch, self.false_branch = condition, ScopeStmt(), ScopeStmt()


class LoopStmt(Statement):
    tag = Statement.Tag.LOOP

    def __init__(self):
        self.body = ScopeStmt()


class ExprStmt(Statement):
    tag = Statement.Tag.EXPR

    def __init__(self, expression):
        self.expression = expression


class SwitchStmt(Statement):
    tag = Statement.Tag.SWITCH

    def __init__(self, expression):
        self.expression, self.body = expression, ScopeStmt()


class SwitchCaseStmt(Statement):
    tag = Statement.Tag.SWITCH_CASE

    def __init__(self, expression):
        self.expression, self.body = expression, ScopeStmt()


class SwitchDefaultStmt(Statement):
    tag = Statement.Tag.SWITCH_DEFAULT

    def __init__(self):
        self.body = ScopeStmt()


class AssignStmt(Statement):
    tag = Statement.Tag.ASSIGN

    def __init__(self, lhs, rhs):
        self.lhs, self.rhs = lhs, rhs


class ForStmt(Statement):
    """for (; variable <cond>; variable += step) body -- the variable is assigned its start before."""
    tag = Statement.Tag.FOR

    def __init__(self, variable, condition, step):
        self.variable, self.condition, self.step, self.body = variable, condition, step, ScopeStmt()


class CommentStmt(Statement):
    tag = Statement.Tag.COMMENT

    def __init__(self, comment):
        self.comment = comment


class RayQueryStmt(Statement):
    """Runs a ray query (a local of type LC_RayQueryAll / LC_RayQueryAny assigned by
    RAY_TRACING_QUERY_ALL / _ANY): on_triangle_candidate runs for every surface candidate and may
    RAY_QUERY_COMMIT_TRIANGLE; on_procedural_candidate for every procedural one and may
    RAY_QUERY_COMMIT_PROCEDURAL(query, t). Afterwards RAY_QUERY_COMMITTED_HIT reads the result."""
    tag = Statement.Tag.RAY_QUERY

    def __init__(self, query):
        self.query = query
        self.on_triangle_candidate, self.on_procedural_candidate = ScopeStmt(), ScopeStmt()


# ---- functions ----------------------------------------------------------------------------------

class Function:
    """A kernel or a callable (include/luisa/ast/function.h): arguments, builtin and local
    variables, the body, the return type (callables), the callables it calls and the usage of
    every variable. Callables that capture a variable of the function defining them take it as a
    REFERENCE argument (locals) or by value (resources and uniforms)."""

    class Tag(enum.IntEnum):
        KERNEL = 0
        CALLABLE = 1
        RASTER_STAGE = 2

    def __init__(self, tag, name, block_size=(1, 1, 1)):
        self.tag, self.name, self.block_size = tag, name, tuple(block_size)
        self.arguments = []
        self.builtin_variables = []
        self.local_variables = []
        self.shared_variables = []
        self.body = ScopeStmt()
        self.return_type = None
        self.custom_callables = []
        self._usage = {}
        self._uid = 0

    # variables
    def _new(self, type_, tag, name):
        v = Variable(type_, tag, self._uid, name)
        self._uid += 1
        return v

    def argument(self, type_, name, reference=False):
        if type_.tag == Type.Tag.BUFFER:
            tag = Variable.Tag.BUFFER
        elif type_.tag == Type.Tag.TEXTURE:
            tag = Variable.Tag.TEXTURE
        elif type_.tag == Type.Tag.BINDLESS_ARRAY:
            tag = Variable.Tag.BINDLESS_ARRAY
        elif type_.tag == Type.Tag.ACCEL:
            tag = Variable.Tag.ACCEL
        else:
            tag = Variable.Tag.REFERENCE if reference else Variable.Tag.LOCAL
        v = self._new(type_, tag, name)
        self.arguments.append(v)
        return v

    def local(self, type_, name):
        v = self._new(type_, Variable.Tag.LOCAL, name)
        self.local_variables.append(v)
        return v

    def shared(self, type_, name):
        """A workgroup-shared variable (FunctionBuilder::shared; kernels only)."""
        if self.tag != Function.Tag.KERNEL:
            raise ValueError("shared variables belong to kernels")
        v = self._new(type_, Variable.Tag.SHARED, name)
        self.shared_variables.append(v)
        return v

    def builtin(self, tag):
        for v in self.builtin_variables:
            if v.tag == tag:
                return v
        t = UINT3 if tag in (Variable.Tag.THREAD_ID, Variable.Tag.BLOCK_ID, Variable.Tag.DISPATCH_ID,
                             Variable.Tag.DISPATCH_SIZE) else UINT
        v = self._new(t, tag, tag.name.lower())
        self.builtin_variables.append(v)
        return v

    # usage
    def mark(self, variable, usage):
        self._usage[variable.uid] = self._usage.get(variable.uid, Usage.NONE) | usage

    def variable_usage(self, variable):
        return self._usage.get(variable.uid, Usage.NONE)

    def argument_usage(self, index):
        return self.variable_usage(self.arguments[index])

    def hash(self):
        """A digest of the printed function (the codegen's text is the canonical form)."""
        from . import codegen
        return hashlib.sha1(codegen.emit(self).encode()).hexdigest()[:16]


def walk_expressions(node, fn):
    """Calls fn on every expression under a statement or expression (pre-order)."""
    if isinstance(node, Expression):
        fn(node)
        for child in _expr_children(node):
            walk_expressions(child, fn)
        return
    for child in _stmt_children(node):
        walk_expressions(child, fn)


def _expr_children(e):
    if isinstance(e, UnaryExpr):
        return [e.operand]
    if isinstance(e, BinaryExpr):
        return [e.lhs, e.rhs]
    if isinstance(e, MemberExpr):
        return [e.self]
    if isinstance(e, AccessExpr):
        return [e.range, e.index]
    if isinstance(e, CallExpr):
        return list(e.arguments)
    if isinstance(e, CastExpr):
        return [e.expression]
    return []


def _stmt_children(s):
    if isinstance(s, ScopeStmt):
        return list(s.statements)
    if isinstance(s, IfStmt):
        return [s.condition, s.true_branch, s.false_branch]
    if isinstance(s, LoopStmt):
        return [s.body]
    if isinstance(s, ExprStmt):
        return [s.expression]
    if isinstance(s, AssignStmt):
        return [s.lhs, s.rhs]
    if isinstance(s, ReturnStmt):
        return [s.expression] if s.expression is not None else []
    if isinstance(s, ForStmt):
        return [s.condition, s.step, s.body]
    if isinstance(s, (SwitchStmt, SwitchCaseStmt)):
        return [s.expression, s.body]
    if isinstance(s, SwitchDefaultStmt):
        return [s.body]
    if isinstance(s, RayQueryStmt):
        return [s.on_triangle_candidate, s.on_procedural_candidate]
    return []
