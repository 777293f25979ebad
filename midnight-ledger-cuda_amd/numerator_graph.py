"""Expression graphs for the quotient numerator, in the GPU encoding.

The prover describes every custom gate, lookup and trash argument as an
expression over fixed / advice / instance columns at rotations.  The
reference builds a straight-line "graph" from those expressions
(vendor/midnight-proofs-0.8.1/src/plonk/evaluation.rs:2723-2820
Evaluator::new; :3638-3786 add_rotation / add_constant / add_calculation /
add_expression), then lowers it for the GPU with intermediate-slot reuse
(evaluation.rs:392-560 intermediate_slot_plan, :582-803 compile_source /
compile_resident_graph).  This module restates both steps for Python
callers; `RawGraphBuffers.as_raw()` yields the `RawGraph` struct that
`midnight_sppark_bls12_381_eval_numerator` and the numerator context take
(layout bls12_381_g1_msm.cu:351-376).

Expressions are tuples:
    ("constant", v)            v an int (reduced mod r)
    ("fixed", col, rot)        ("advice", col, rot)    ("instance", col, rot)
    ("challenge", i)
    ("neg", e)   ("sum", a, b)   ("prod", a, b)   ("scaled", e, f)
Field values are canonical Python ints modulo the scalar field order.
"""
from __future__ import annotations

import ctypes

import numpy as np

FR_MODULUS = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001

# ValueSource variants, in declaration order (evaluation.rs:49-73); the
# derived ordering compares the variant first, then its fields.
CONST, INTER, FIXED, ADVICE, INSTANCE, CHALLENGE, BETA, GAMMA, THETA, TRASH, Y, PREV = range(12)
# Opcodes of the GPU encoding (midnight-gpu-prover/src/expression.rs:80-98)
OP_ADD, OP_SUB, OP_MUL, OP_SQUARE, OP_DOUBLE, OP_NEGATE, OP_HORNER, OP_STORE = range(8)

_OPCODE = {"add": OP_ADD, "sub": OP_SUB, "mul": OP_MUL, "square": OP_SQUARE,
           "double": OP_DOUBLE, "negate": OP_NEGATE, "horner": OP_HORNER, "store": OP_STORE}


def vs(kind, a=0, b=0):
    """A ValueSource as a (variant, a, b) tuple; tuple order = derived order."""
    return (kind, a, b)


ZERO_CONST = vs(CONST, 0)
ONE_CONST = vs(CONST, 1)
TWO_CONST = vs(CONST, 2)


class GraphEvaluator:
    """evaluation.rs GraphEvaluator: constants start as [0, 1, 2]
    (Default impl, evaluation.rs:3531-3541)."""

    def __init__(self):
        self.constants = [0, 1, 2]
        self.rotations = []
        self.calculations = []      # (calculation tuple, target)
        self.num_intermediates = 0

    def add_rotation(self, rot):
        if rot in self.rotations:
            return self.rotations.index(rot)
        self.rotations.append(rot)
        return len(self.rotations) - 1

    def add_constant(self, c):
        c %= FR_MODULUS
        if c in self.constants:
            return vs(CONST, self.constants.index(c))
        self.constants.append(c)
        return vs(CONST, len(self.constants) - 1)

    def add_calculation(self, calc):
        for existing, target in self.calculations:
            if existing == calc:
                return vs(INTER, target)
        target = self.num_intermediates
        self.calculations.append((calc, target))
        self.num_intermediates += 1
        return vs(INTER, target)

    def add_expression(self, e):
        tag = e[0]
        if tag == "constant":
            return self.add_constant(e[1])
        if tag in ("fixed", "advice", "instance"):
            kind = {"fixed": FIXED, "advice": ADVICE, "instance": INSTANCE}[tag]
            rot = self.add_rotation(e[2])
            return self.add_calculation(("store", vs(kind, e[1], rot)))
        if tag == "challenge":
            return self.add_calculation(("store", vs(CHALLENGE, e[1])))
        if tag == "neg":
            inner = e[1]
            if inner[0] == "constant":
                return self.add_constant(-inner[1])
            ra = self.add_expression(inner)
            if ra == ZERO_CONST:
                return ra
            return self.add_calculation(("negate", ra))
        if tag == "sum":
            a, b = e[1], e[2]
            if b[0] == "neg":           # a - b is stored as a + (-b)
                ra = self.add_expression(a)
                rb = self.add_expression(b[1])
                if ra == ZERO_CONST:
                    return self.add_calculation(("negate", rb))
                if rb == ZERO_CONST:
                    return ra
                return self.add_calculation(("sub", ra, rb))
            ra = self.add_expression(a)
            rb = self.add_expression(b)
            if ra == ZERO_CONST:
                return rb
            if rb == ZERO_CONST:
                return ra
            return self.add_calculation(("add", ra, rb) if ra <= rb else ("add", rb, ra))
        if tag == "prod":
            ra = self.add_expression(e[1])
            rb = self.add_expression(e[2])
            if ra == ZERO_CONST or rb == ZERO_CONST:
                return ZERO_CONST
            if ra == ONE_CONST:
                return rb
            if rb == ONE_CONST:
                return ra
            if ra == TWO_CONST:
                return self.add_calculation(("double", rb))
            if rb == TWO_CONST:
                return self.add_calculation(("double", ra))
            if ra == rb:
                return self.add_calculation(("square", ra))
            return self.add_calculation(("mul", ra, rb) if ra <= rb else ("mul", rb, ra))
        if tag == "scaled":
            f = e[2] % FR_MODULUS
            if f == 0:
                return ZERO_CONST
            if f == 1:
                return self.add_expression(e[1])
            cst = self.add_constant(f)
            ra = self.add_expression(e[1])
            return self.add_calculation(("mul", ra, cst))
        raise ValueError("unknown expression tag %r" % (tag,))


def custom_gates_graph(polys):
    """Custom-gate graph: every gate polynomial, folded in y over the
    running value (Evaluator::new, evaluation.rs:2726-2739)."""
    g = GraphEvaluator()
    parts = [g.add_expression(p) for p in polys]
    g.add_calculation(("horner", vs(PREV), tuple(parts), vs(Y)))
    return g


def lookup_graph(inputs, tables):
    """Per-lookup graph (a'(X)+beta-side compression times table side),
    evaluation.rs:2768-2797."""
    g = GraphEvaluator()

    def lc(exprs):
        parts = tuple(g.add_expression(x) for x in exprs)
        return g.add_calculation(("horner", ZERO_CONST, parts, vs(THETA)))

    ci = lc(inputs)
    ct = lc(tables)
    right_gamma = g.add_calculation(("add", ct, vs(GAMMA)))
    left = g.add_calculation(("add", ci, vs(BETA)))
    g.add_calculation(("mul", left, right_gamma))
    return g


def compression_graph(exprs):
    """A lookup's input or table compression, Horner in theta from constant
    zero (evaluation.rs:2743-2766): the graphs of
    numerator_context_create's lookup_input_graphs / lookup_table_graphs."""
    g = GraphEvaluator()
    parts = tuple(g.add_expression(x) for x in exprs)
    g.add_calculation(("horner", ZERO_CONST, parts, vs(THETA)))
    return g


def trash_graph(constraints):
    """Trash-argument graph: constraints folded in the trash challenge
    (evaluation.rs:2801-2817)."""
    g = GraphEvaluator()
    parts = tuple(g.add_expression(x) for x in constraints)
    g.add_calculation(("horner", ZERO_CONST, parts, vs(TRASH)))
    return g


def _sources(calc):
    """Sources in the order the slot planner visits them
    (visit_calculation_sources, evaluation.rs:398-417)."""
    op = calc[0]
    if op in ("add", "sub", "mul"):
        return [calc[1], calc[2]]
    if op in ("square", "double", "negate", "store"):
        return [calc[1]]
    if op == "horner":
        return [calc[1], calc[3]] + list(calc[2])
    raise ValueError(op)


def slot_plan(g):
    """intermediate_slot_plan (evaluation.rs:419-545): reuse an
    intermediate's slot once its last reader has run."""
    remaining = [0] * g.num_intermediates
    for calc, _ in g.calculations:
        for s in _sources(calc):
            if s[0] == INTER:
                remaining[s[1]] += 1
    slots = [None] * g.num_intermediates
    free = []
    next_slot = 0
    for calc, target in g.calculations:
        if slots[target] is not None:
            raise ValueError("duplicate target %d" % target)
        for s in _sources(calc):
            if s[0] != INTER:
                continue
            if slots[s[1]] is None or remaining[s[1]] == 0:
                raise ValueError("intermediate %d read before write" % s[1])
            remaining[s[1]] -= 1
            if remaining[s[1]] == 0:
                free.append(slots[s[1]])
        if free:
            slot = free.pop()
        else:
            slot = next_slot
            next_slot += 1
        slots[target] = slot
        if remaining[target] == 0:
            free.append(slot)
    return slots, next_slot


# ---- ctypes mirror of the ABI structs (bls12_381_g1_msm.cu:351-381) ----
class RawValueSource(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("a", ctypes.c_uint32), ("b", ctypes.c_uint32)]


class RawCalculation(ctypes.Structure):
    _fields_ = [("op", ctypes.c_uint32), ("target", ctypes.c_uint32),
                ("parts_offset", ctypes.c_uint32), ("parts_len", ctypes.c_uint32),
                ("lhs", RawValueSource), ("rhs", RawValueSource)]


class RawGraph(ctypes.Structure):
    _fields_ = [("constants", ctypes.c_void_p), ("constants_len", ctypes.c_size_t),
                ("rotations", ctypes.c_void_p), ("rotations_len", ctypes.c_size_t),
                ("calculations", ctypes.c_void_p), ("calculations_len", ctypes.c_size_t),
                ("horner_parts", ctypes.c_void_p), ("horner_parts_len", ctypes.c_size_t),
                ("num_intermediates", ctypes.c_uint32)]


class RawColumnRef(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_uint32), ("index", ctypes.c_uint32)]


def to_montgomery_limbs(values):
    """Canonical ints -> (n, 4) uint64 Montgomery limbs (R = 2^256)."""
    out = np.zeros((len(values), 4), dtype=np.uint64)
    for i, v in enumerate(values):
        m = (v % FR_MODULUS) * (1 << 256) % FR_MODULUS
        for j in range(4):
            out[i, j] = (m >> (64 * j)) & 0xFFFFFFFFFFFFFFFF
    return out


def from_montgomery_limbs(arr):
    """(n, 4) uint64 Montgomery limbs -> canonical ints."""
    rinv = pow(1 << 256, FR_MODULUS - 2, FR_MODULUS)
    a = np.asarray(arr, dtype=np.uint64).reshape(-1, 4)
    return [(int(r[0]) | int(r[1]) << 64 | int(r[2]) << 128 | int(r[3]) << 192) * rinv % FR_MODULUS
            for r in a]


class RawGraphBuffers:
    """A lowered graph (compile_resident_graph, evaluation.rs:677-800):
    Montgomery constants, rotations, operations with slot targets and the
    flattened Horner parts.  Keeps the arrays alive for `as_raw()`."""

    def __init__(self, g: GraphEvaluator):
        slots, capacity = slot_plan(g)

        def src(s):
            kind, a, b = s
            if kind == INTER:
                a = slots[a]
            return (kind, a, b)

        ops, parts = [], []
        for calc, target in g.calculations:
            op = calc[0]
            if op == "horner":
                off = len(parts)
                parts.extend(src(p) for p in calc[2])
                ops.append((OP_HORNER, slots[target], off, len(calc[2]), src(calc[1]), src(calc[3])))
            elif op in ("add", "sub", "mul"):
                ops.append((_OPCODE[op], slots[target], 0, 0, src(calc[1]), src(calc[2])))
            else:
                ops.append((_OPCODE[op], slots[target], 0, 0, src(calc[1]), (0, 0, 0)))
        self.constants = to_montgomery_limbs(g.constants)
        self.rotations = np.array(g.rotations, dtype=np.int32)
        self.calculations = (RawCalculation * max(1, len(ops)))()
        for i, (op, t, off, n, lhs, rhs) in enumerate(ops):
            self.calculations[i] = RawCalculation(op, t, off, n, RawValueSource(*lhs), RawValueSource(*rhs))
        self.horner_parts = (RawValueSource * max(1, len(parts)))()
        for i, p in enumerate(parts):
            self.horner_parts[i] = RawValueSource(*p)
        self.ops = ops
        self.parts = parts
        self.num_intermediates = capacity

    def as_raw(self) -> RawGraph:
        return RawGraph(self.constants.ctypes.data, len(self.constants),
                        self.rotations.ctypes.data if len(self.rotations) else None, len(self.rotations),
                        ctypes.addressof(self.calculations), len(self.ops),
                        ctypes.addressof(self.horner_parts), len(self.parts),
                        self.num_intermediates)


def empty_raw_graph() -> RawGraph:
    return RawGraph(None, 0, None, 0, None, 0, None, 0, 0)
