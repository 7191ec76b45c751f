"""Independent big-int KoalaBear arithmetic for pinning the oracle (tests only).

F: integers mod p = 2^31 - 2^24 + 1.  EF: F[X]/(X^8 - 3), coefficient lists.
Nothing here is imported by the product.
"""
P = 0x7F000001
W = 3
D = 8


def fmul(a, b):
    return a * b % P


def finv(a):
    return pow(a, P - 2, P)


def ef(x):
    """Normalise an int or sequence into an EF coefficient list."""
    if isinstance(x, int):
        return [x % P] + [0] * (D - 1)
    return [int(v) % P for v in x]


def eadd(a, b):
    return [(x + y) % P for x, y in zip(a, b)]


def esub(a, b):
    return [(x - y) % P for x, y in zip(a, b)]


def emul(a, b):
    prod = [0] * (2 * D - 1)
    for i, x in enumerate(a):
        if x:
            for j, y in enumerate(b):
                prod[i + j] += x * y
    out = prod[:D]
    for k in range(D, 2 * D - 1):
        out[k - D] += W * prod[k]
    return [v % P for v in out]


def epow(a, e):
    r = ef(1)
    while e:
        if e & 1:
            r = emul(r, a)
        a = emul(a, a)
        e >>= 1
    return r


def einv(a):
    return epow(a, P ** D - 2)


def eq_table(point):
    """eq(point, .) big-endian."""
    t = [ef(1)]
    for p in point:
        n = []
        for v in t:
            hi = emul(v, p)
            n.append(esub(v, hi))
            n.append(hi)
        t = n
    return t


def interpolate(xs, ys):
    """Lagrange interpolation over F nodes, EF values; coefficients low first."""
    n = len(xs)
    coeffs = [ef(0) for _ in range(n)]
    for i in range(n):
        basis = [1]
        den = 1
        for j in range(n):
            if j == i:
                continue
            basis = [(c1 - xs[j] * c0) % P for c0, c1 in zip(basis + [0], [0] + basis)]
            den = den * (xs[i] - xs[j]) % P
        s = [v * finv(den) % P for v in ys[i]]
        for d in range(n):
            coeffs[d] = eadd(coeffs[d], [v * basis[d] % P for v in s])
    return coeffs


def to_u32(vals):
    """list of EF -> flat list of ints"""
    out = []
    for v in vals:
        out.extend(v)
    return out


def from_u32(arr):
    arr = [int(v) for v in arr]
    return [arr[i:i + D] for i in range(0, len(arr), D)]
