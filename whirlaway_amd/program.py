"""Constraint programs: the build's form of an AIR's symbolic constraints.

The reference evaluates an AIR's constraints through p3-air's AirBuilder
(ConstraintFolder, crates/utils/src/constraints_folder.rs; the count comes from
get_symbolic_constraints, crates/air/src/table.rs:42).  Here an AIR, or any
SumcheckComputation, is written once against ConstraintProgram, which records the
expression DAG as the straight-line SSA code wl_sumcheck_prove and the generic
AIR table take (4 words per instruction: op, a, b, 0).

    prog = ConstraintProgram(width)           # inputs: local row, then next row
    a, b = prog.local(0), prog.next(1)
    prog.assert_zero(a * a - b + 3)           # i-th assert gets batching[i]
    code = prog.code()
"""
import numpy as np

P = 0x7F000001
OP_CONST, OP_ADD, OP_SUB, OP_MUL, OP_NEG, OP_ASSERT = 0, 1, 2, 3, 4, 5


class Expr:
    __slots__ = ("prog", "id")

    def __init__(self, prog, vid):
        self.prog, self.id = prog, vid

    def _wrap(self, o):
        return o if isinstance(o, Expr) else self.prog.const(o)

    def __add__(self, o):
        return self.prog._op(OP_ADD, self, self._wrap(o))

    def __radd__(self, o):
        return self.prog._op(OP_ADD, self._wrap(o), self)

    def __sub__(self, o):
        return self.prog._op(OP_SUB, self, self._wrap(o))

    def __rsub__(self, o):
        return self.prog._op(OP_SUB, self._wrap(o), self)

    def __mul__(self, o):
        return self.prog._op(OP_MUL, self, self._wrap(o))

    def __rmul__(self, o):
        return self.prog._op(OP_MUL, self._wrap(o), self)

    def __neg__(self):
        return self.prog._op(OP_NEG, self, None)

    def __pow__(self, e):
        if not isinstance(e, int) or e < 1:
            raise ValueError("positive integer powers only")
        r, b = None, self
        while e:
            if e & 1:
                r = b if r is None else r * b
            e >>= 1
            if e:
                b = b * b
        return r


class ConstraintProgram:
    """Records constraints over a point of n_inputs values.  For an AIR of width C
    the point is [local row (C), next row (C)] (sc_computation.rs:27: the point is
    split in two rows of point.len() / 2)."""

    def __init__(self, width=None, n_inputs=None):
        if (width is None) == (n_inputs is None):
            raise ValueError("give the AIR width or the number of inputs")
        self.width = width
        self.n_inputs = 2 * width if width is not None else n_inputs
        self._code = []
        self._consts = {}
        self.n_constraints = 0

    def input(self, i):
        if not 0 <= i < self.n_inputs:
            raise IndexError(i)
        return Expr(self, i)

    def local(self, col):
        return self.input(col)

    def next(self, col):
        return self.input(self.width + col)

    def const(self, c):
        c = int(c) % P
        if c not in self._consts:
            self._consts[c] = self._emit(OP_CONST, c, 0)
        return self._consts[c]

    def _emit(self, op, a, b):
        self._code.append((op, a, b))
        return Expr(self, self.n_inputs + len(self._code) - 1)

    def _op(self, op, x, y):
        return self._emit(op, x.id, 0 if y is None else y.id)

    def assert_zero(self, e):
        e = e if isinstance(e, Expr) else self.const(e)
        self._code.append((OP_ASSERT, e.id, 0))
        self.n_constraints += 1

    def code(self):
        out = np.zeros((len(self._code), 4), dtype=np.uint32)
        for j, (op, a, b) in enumerate(self._code):
            out[j, 0], out[j, 1], out[j, 2] = op, a, b
        return out.reshape(-1)

    @property
    def n_instr(self):
        return len(self._code)


def product_program():
    """InnerSumcheckCircuit (crates/air/src/prove.rs:229-237): point[0] * point[1]."""
    p = ConstraintProgram(n_inputs=2)
    p.assert_zero(p.input(0) * p.input(1))
    return p


def a2b_program():
    """point[0]^2 * point[1] (the kernel micro's product sumcheck)."""
    p = ConstraintProgram(n_inputs=2)
    a, b = p.input(0), p.input(1)
    p.assert_zero(a * a * b)
    return p


# Poseidon2-16 KoalaBear internal diagonal (the same as the device tables): -2, 1, 2,
# 1/2, 3, 4, -1/2, -3, -4, 1/2^8, 1/8, 1/2^24, -1/2^8, -1/8, -1/16, -1/2^24
def _inv2k(k):
    return pow(2, P - 1 - k, P)


P2_DIAG16 = [P - 2, 1, 2, _inv2k(1), 3, 4, P - _inv2k(1), P - 3, P - 4, _inv2k(8), _inv2k(3), _inv2k(24),
             P - _inv2k(8), P - _inv2k(3), P - _inv2k(4), P - _inv2k(24)]


def _mds_light16(s):
    """Poseidon2 external layer on 16 elements: M4 on each block of 4, then each
    element plus the sum of the elements in its position across blocks."""
    out = list(s)
    for c in range(0, 16, 4):
        x = out[c:c + 4]
        t01, t23 = x[0] + x[1], x[2] + x[3]
        t0123 = t01 + t23
        t01123, t01233 = t0123 + x[1], t0123 + x[3]
        out[c:c + 4] = [t01123 + t01, t01123 + (x[2] + x[2]), t01233 + t23, t01233 + (x[0] + x[0])]
    sums = [out[i] + out[i + 4] + out[i + 8] + out[i + 12] for i in range(4)]
    return [out[i] + sums[i & 3] for i in range(16)]


def poseidon2_air_program(round_constants):
    """The constraints of Poseidon2Air<16, 3, 0, 4, 20> (the reference's
    src/examples/poseidon2.rs:108-116, width 165) in its column and constraint
    order: export * (export - 1); 16 per beginning full round; one per partial
    round; 16 per ending full round (149 in all).  round_constants: the 148
    canonical RoundConstants (beginning[4][16], partial[20], ending[4][16])."""
    rc = [int(x) for x in np.asarray(round_constants).reshape(-1)]
    if len(rc) != 148:
        raise ValueError("148 round constants")
    beg, part, end = rc[:64], rc[64:84], rc[84:]
    p = ConstraintProgram(width=165)
    x = [p.local(i) for i in range(165)]
    p.assert_zero(x[0] * (x[0] - 1))
    s = _mds_light16(x[1:17])
    col = 17
    for r in range(4):
        s = _mds_light16([(s[i] + beg[16 * r + i]) ** 3 for i in range(16)])
        for i in range(16):
            p.assert_zero(s[i] - x[col + i])
        s = x[col:col + 16]
        col += 16
    for r in range(20):
        p.assert_zero((s[0] + part[r]) ** 3 - x[col])
        s = [x[col]] + s[1:]
        col += 1
        tot = s[0]
        for i in range(1, 16):
            tot = tot + s[i]
        s = [tot + s[i] * P2_DIAG16[i] for i in range(16)]
    for r in range(4):
        s = _mds_light16([(s[i] + end[16 * r + i]) ** 3 for i in range(16)])
        for i in range(16):
            p.assert_zero(s[i] - x[col + i])
        s = x[col:col + 16]
        col += 16
    return p
