"""Example AIRs for the generic-AIR tests, written against ConstraintProgram.

MulFib (width 4: k preprocessed, then a, b, c), constraints of degree 3:
    c = a * b * k          (local row)
    next.a = b
    next.b = a + b + k
The up/down views of the reference make the last row repeat the pair
(N-2, N-1), so transition constraints that hold on consecutive rows hold
everywhere (crates/air/src/utils.rs:140-160)."""
import numpy as np

P = 0x7F000001


def mul_fib_program():
    from whirlaway_amd.program import ConstraintProgram
    p = ConstraintProgram(width=4)
    k, a, b, c = (p.local(i) for i in range(4))
    p.assert_zero(c - a * b * k)
    p.assert_zero(p.next(1) - b)
    p.assert_zero(p.next(2) - a - b - k)
    return p


def mul_fib_columns(log_n, seed):
    """(preprocessed k column (1 x N), witness a, b, c (3 x N))."""
    rng = np.random.default_rng(seed)
    N = 1 << log_n
    k = rng.integers(0, P, N).astype(object)
    a, b = [int(rng.integers(0, P))], [int(rng.integers(0, P))]
    for i in range(N - 1):
        a.append(b[i])
        b.append((a[i] + b[i] + int(k[i])) % P)
    c = [a[i] * b[i] % P * int(k[i]) % P for i in range(N)]
    pre = np.array([k], dtype=np.uint64).astype(np.uint32)
    wit = np.array([a, b, c], dtype=np.uint64).astype(np.uint32)
    return pre, wit
