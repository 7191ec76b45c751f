"""Pins the CPU oracle against closed forms, brute force and independent restatements.

Each test checks an oracle function through a DIFFERENT computation than the one
the oracle uses (dense matrices instead of addition chains, brute-force sums over
the hypercube instead of folds, direct polynomial evaluation instead of butterflies),
so that a dropped term, a wrong index or a transposed operand fails here.
"""
import ctypes
import math
import itertools

import numpy as np
import pytest

import pyfield as pf

P = pf.P


def _u32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.uint64) % P, dtype=np.uint32)


def _rand_f(rng, n):
    return [int(x) for x in rng.integers(0, P, n)]


def _rand_ef(rng):
    return _rand_f(rng, 8)


# ------------------------------------------------------------------ Poseidon2 --
M4 = [[2, 3, 1, 1], [1, 2, 3, 1], [1, 1, 2, 3], [3, 1, 1, 2]]


def _external_matrix(w):
    """Poseidon2 external layer for t = 4k: circ(2*M4, M4, ..., M4) as a dense matrix."""
    m = [[0] * w for _ in range(w)]
    for bi in range(w // 4):
        for bj in range(w // 4):
            f = 2 if bi == bj else 1
            for i in range(4):
                for j in range(4):
                    m[4 * bi + i][4 * bj + j] = f * M4[i][j]
    return m


def _inv2k(k):
    return pow(pow(2, k, P), P - 2, P)


# KoalaBear internal-layer diagonals as published with Plonky3's KoalaBear Poseidon2
# (the values themselves are a dependency constant: parity unpinned).
DIAG = {
    16: [-2, 1, 2, _inv2k(1), 3, 4, -_inv2k(1), -3, -4, _inv2k(8), _inv2k(3), _inv2k(24),
         -_inv2k(8), -_inv2k(3), -_inv2k(4), -_inv2k(24)],
    24: [-2, 1, 2, _inv2k(1), 3, 4, -_inv2k(1), -3, -4, _inv2k(8), _inv2k(2), _inv2k(3),
         _inv2k(4), _inv2k(5), _inv2k(6), _inv2k(24), -_inv2k(8), -_inv2k(3), -_inv2k(4),
         -_inv2k(5), -_inv2k(6), -_inv2k(7), -_inv2k(9), -_inv2k(24)],
}


def _internal_matrix(w):
    """M_I = J + diag(d): every output is the state sum plus d_i * s_i."""
    d = [x % P for x in DIAG[w]]
    return [[(1 + (d[i] if i == j else 0)) % P for j in range(w)] for i in range(w)]


def _matvec(m, v):
    return [sum(a * b for a, b in zip(row, v)) % P for row in m]


def _dense_permute(w, consts, state):
    """Poseidon2 (x^3 S-box, 4+4 full rounds, R_P partial) from dense matrices."""
    rp = 20 if w == 16 else 23
    ext_init = [consts[r * w:(r + 1) * w] for r in range(4)]
    ext_term = [consts[4 * w + r * w: 4 * w + (r + 1) * w] for r in range(4)]
    internal = consts[8 * w: 8 * w + rp]
    ME, MI = _external_matrix(w), _internal_matrix(w)
    s = _matvec(ME, state)
    for r in range(4):
        s = _matvec(ME, [pow((x + c) % P, 3, P) for x, c in zip(s, ext_init[r])])
    for r in range(rp):
        s[0] = pow((s[0] + internal[r]) % P, 3, P)
        s = _matvec(MI, s)
    for r in range(4):
        s = _matvec(ME, [pow((x + c) % P, 3, P) for x, c in zip(s, ext_term[r])])
    return s


def _det_mod_p(m):
    m = [row[:] for row in m]
    n = len(m)
    det = 1
    for c in range(n):
        piv = next((r for r in range(c, n) if m[r][c] % P), None)
        if piv is None:
            return 0
        if piv != c:
            m[c], m[piv] = m[piv], m[c]
            det = -det
        det = det * m[c][c] % P
        inv = pow(m[c][c], P - 2, P)
        for r in range(c + 1, n):
            f = m[r][c] * inv % P
            if f:
                m[r] = [(x - f * y) % P for x, y in zip(m[r], m[c])]
    return det % P


def test_sbox_is_a_permutation():
    from math import gcd
    assert gcd(3, P - 1) == 1


def test_m4_is_mds_and_layers_invertible():
    for k in range(1, 5):
        for rows in itertools.combinations(range(4), k):
            for cols in itertools.combinations(range(4), k):
                sub = [[M4[r][c] for c in cols] for r in rows]
                assert _det_mod_p(sub) != 0
    for w in (16, 24):
        assert _det_mod_p(_external_matrix(w)) != 0
        assert _det_mod_p(_internal_matrix(w)) != 0


@pytest.mark.parametrize("width", [16, 24])
def test_poseidon2_matches_dense_restatement(oracle, width):
    b = oracle.Bench(4, seed=3)
    consts = [int(x) for x in b.p2_consts(width)]
    rng = np.random.default_rng(width)
    for trial in range(3):
        state = _rand_f(rng, width) if trial else [0] * width
        got = b.permute(width, np.array(state, dtype=np.uint32))
        assert [int(x) for x in got] == _dense_permute(width, consts, state)
    b.close()


def test_sponge_and_compression_modes(oracle):
    """PaddingFreeSponge<P24, 24, 16, 8> (overwrite absorption, 8-element squeeze) and
    TruncatedPermutation<P16, 2, 8, 16>, restated over the oracle's permutations."""
    b = oracle.Bench(4, seed=5)
    L = oracle.lib()
    rng = np.random.default_rng(1)
    for n in (1, 5, 16, 17, 32, 128, 165):
        row = np.array(_rand_f(rng, n), dtype=np.uint32)
        s = np.zeros(24, dtype=np.uint32)
        for off in range(0, n, 16):
            chunk = row[off:off + 16]
            s[:chunk.size] = chunk
            s = b.permute(24, s)
        out = np.zeros(8, dtype=np.uint32)
        L.or_api_hash_leaf(b.h, oracle.ptr(row), n, oracle.ptr(out))
        assert np.array_equal(out, s[:8]), n
    l_, r_ = np.array(_rand_f(rng, 8), dtype=np.uint32), np.array(_rand_f(rng, 8), dtype=np.uint32)
    out = np.zeros(8, dtype=np.uint32)
    L.or_api_compress(b.h, oracle.ptr(l_), oracle.ptr(r_), oracle.ptr(out))
    assert np.array_equal(out, b.permute(16, np.concatenate([l_, r_]))[:8])
    b.close()


# ------------------------------------------------------------------------ AIR --
def _air_eval(oracle, b, local, alpha_pows):
    out = np.zeros(8, dtype=np.uint32)
    oracle.lib().or_api_air_eval_f(b.h, oracle.ptr(_u32(local)), oracle.ptr(_u32(alpha_pows)),
                                   oracle.ptr(out))
    return [int(x) for x in out]


def test_trace_rows_are_poseidon2_of_inputs(oracle):
    n = 5
    b = oracle.Bench(n, seed=9)
    tr = b.trace().reshape(165, 1 << n)
    ac = [int(x) for x in b.air_consts()]
    # AIR constants as a Poseidon2 instance: beginning = initial external, ending = terminal
    consts = ac[:64] + ac[84:148] + ac[64:84]
    inputs = b.inputs().reshape(1 << n, 16)
    for row in (0, 7, (1 << n) - 1):
        assert np.array_equal(tr[1:17, row], inputs[row])
        assert int(tr[0, row]) == 1
        want = _dense_permute(16, consts, [int(x) for x in inputs[row]])
        assert [int(x) for x in tr[101 + 48:165, row]] == want
    b.close()


def test_constraints_vanish_on_trace_and_catch_changes(oracle):
    n = 4
    b = oracle.Bench(n, seed=2)
    tr = b.trace().reshape(165, 1 << n)
    rng = np.random.default_rng(0)
    alphas = np.array([_rand_ef(rng) for _ in range(149)], dtype=np.uint64).reshape(-1)
    for row in range(1 << n):
        assert _air_eval(oracle, b, tr[:, row], alphas) == [0] * 8
    # one constraint at a time: unit alpha vectors
    local = tr[:, 3].astype(np.uint64)
    for i in (0, 1, 64, 65, 84, 85, 148):
        e = np.zeros(149 * 8, dtype=np.uint64)
        e[8 * i] = 1
        assert _air_eval(oracle, b, local, e) == [0] * 8
    # changing any column of a row breaks some constraint
    for col in (0, 1, 16, 17, 80, 81, 100, 101, 164):
        bad = local.copy()
        bad[col] = (bad[col] + 1) % P
        assert _air_eval(oracle, b, bad, alphas) != [0] * 8, col
    b.close()


# --------------------------------------------------- eq, evaluation, folding --
def test_eq_table_brute_force(oracle):
    rng = np.random.default_rng(4)
    n = 4
    pt = [_rand_ef(rng) for _ in range(n)]
    out = np.zeros((1 << n) * 8, dtype=np.uint32)
    oracle.lib().or_api_eval_eq(oracle.ptr(_u32(pf.to_u32(pt))), n, oracle.ptr(out))
    got = pf.from_u32(out)
    for x in range(1 << n):
        bits = [(x >> (n - 1 - i)) & 1 for i in range(n)]  # big-endian: pt[0] pairs with the MSB
        v = pf.ef(1)
        for p_, bit in zip(pt, bits):
            v = pf.emul(v, p_ if bit else pf.esub(pf.ef(1), p_))
        assert got[x] == v


def test_mle_eval_and_folds_agree_with_brute_force(oracle):
    rng = np.random.default_rng(5)
    n = 5
    f = _rand_f(rng, 1 << n)
    pt = [_rand_ef(rng) for _ in range(n)]
    eq = pf.eq_table(pt)
    want = pf.ef(0)
    for x in range(1 << n):
        want = pf.eadd(want, [c * f[x] % P for c in eq[x]])
    out = np.zeros(8, dtype=np.uint32)
    oracle.lib().or_api_mle_eval_f(oracle.ptr(_u32(f)), n, oracle.ptr(_u32(pf.to_u32(pt))),
                                   oracle.ptr(out))
    assert [int(x) for x in out] == want
    # folding by all n variables with eq(pt) scalars is the same evaluation
    full = oracle.fold_f_large(_u32(f), _u32(pf.to_u32(eq)), n)
    assert [int(x) for x in full] == want
    # two single-variable folds (MSB first) equal one 2-variable fold with eq((c0, c1))
    c0, c1 = pt[0], pt[1]
    s1 = _u32(pf.to_u32([pf.esub(pf.ef(1), c0), c0]))
    step = oracle.fold_f_large(_u32(f), s1, 1)
    s2 = _u32(pf.to_u32([pf.esub(pf.ef(1), c1), c1]))
    two = oracle.fold_ef_large(step, s2, 1)
    once = oracle.fold_f_large(_u32(f), _u32(pf.to_u32(pf.eq_table([c0, c1]))), 2)
    assert np.array_equal(two, once)
    # k = 1 with (1 - c, c) binds the most significant variable
    h = 1 << (n - 1)
    for i in (0, 3, h - 1):
        lo, hi = f[i], f[i + h]
        want_i = pf.eadd(pf.ef(lo), [(c * (hi - lo)) % P for c in c0])
        assert [int(x) for x in step[8 * i: 8 * i + 8]] == want_i


def test_small_field_fold_and_linear_combination(oracle):
    """fold_multilinear_in_small_field against Python integers, and against the
    large-field fold with the same scalars embedded in EF (another oracle path);
    multilinears_linear_combination against Python integers."""
    rng = np.random.default_rng(21)
    n, k = 6, 2
    m = _rand_f(rng, 1 << n)
    sc = _rand_f(rng, 1 << k)
    got = oracle.fold_small(_u32(m), _u32(sc), k)
    M = 1 << (n - k)
    want = [sum(sc[j] * m[i + j * M] for j in range(1 << k)) % P for i in range(M)]
    assert [int(x) for x in got] == want
    emb = _u32(pf.to_u32([pf.ef(v) for v in sc]))
    large = oracle.fold_f_large(_u32(m), emb, k)
    assert [int(x) for x in large[0::8]] == want and not large.reshape(-1, 8)[:, 1:].any()
    me = _rand_f(rng, (1 << n) * 8)
    gote = oracle.fold_small(_u32(me), _u32(sc), k, is_ef=True)
    for i in (0, 5, M - 1):
        for c in range(8):
            assert int(gote[8 * i + c]) == sum(sc[j] * me[8 * (i + j * M) + c] for j in range(1 << k)) % P
    cols = [_rand_f(rng, 1 << 4) for _ in range(3)]
    scal = [_rand_ef(rng) for _ in range(3)]
    lc = oracle.lincomb(_u32(cols), _u32(scal))
    for x in (0, 7, 15):
        want_x = [sum(scal[c][d] * cols[c][x] for c in range(3)) % P for d in range(8)]
        assert [int(v) for v in lc[8 * x: 8 * x + 8]] == want_x


def test_selectors_closed_form(oracle):
    rng = np.random.default_rng(6)
    for k in (1, 2, 4):
        K = 1 << k
        for x in [pf.ef(j) for j in range(K)] + [_rand_ef(rng)]:
            out = np.zeros(K * 8, dtype=np.uint32)
            oracle.lib().or_api_selectors_at_ef(k, oracle.ptr(_u32(x)), oracle.ptr(out))
            sel = pf.from_u32(out)
            # Lagrange basis on {0..K-1}: prod_{m != j} (x - m) / (j - m)
            for j in range(K):
                num, den = pf.ef(1), 1
                for m in range(K):
                    if m != j:
                        num = pf.emul(num, pf.esub(x, pf.ef(m)))
                        den = den * (j - m) % P
                assert sel[j] == [c * pf.finv(den) % P for c in num]
            total, moment = pf.ef(0), pf.ef(0)
            for j in range(K):
                total = pf.eadd(total, sel[j])
                moment = pf.eadd(moment, [c * j % P for c in sel[j]])
            assert total == pf.ef(1) and moment == x


def test_lagrange_interpolation_round_trip(oracle):
    rng = np.random.default_rng(7)
    n = 9
    coeffs = [_rand_ef(rng) for _ in range(n)]
    xs = list(range(3, 3 + n))
    ys = []
    for x in xs:
        acc = pf.ef(0)
        for c in reversed(coeffs):
            acc = pf.eadd([v * x % P for v in acc], c)
        ys.append(acc)
    out = np.zeros(n * 8, dtype=np.uint32)
    oracle.lib().or_api_lagrange(oracle.ptr(_u32(xs)), oracle.ptr(_u32(pf.to_u32(ys))), n,
                                 oracle.ptr(out))
    assert pf.from_u32(out) == coeffs


# ----------------------------------------------------- Reed-Solomon encoding --
def test_ntt_is_natural_order_evaluation(oracle):
    rng = np.random.default_rng(8)
    for L in (1, 3, 6):
        N = 1 << L
        c = _rand_f(rng, N)
        w = int(oracle.lib().or_api_two_adic_generator(L))
        a = _u32(c)
        oracle.lib().or_ntt(oracle.ptr(a), L)
        for i in range(N):
            assert int(a[i]) == sum(cj * pow(w, i * j, P) for j, cj in enumerate(c)) % P


def test_mobius_gives_multilinear_coefficients(oracle):
    """or_mobius turns evaluations on {0,1}^m (big-endian) into the coefficients of the
    multilinear polynomial; the univariate embedding X_i -> X^(2^(m-1-i)) then makes
    coefficient index = monomial bit mask."""
    rng = np.random.default_rng(9)
    m = 4
    ev = _rand_f(rng, 1 << m)
    co = _u32(ev)
    oracle.lib().or_mobius(oracle.ptr(co), m)
    co = [int(x) for x in co]
    for x in range(1 << m):
        # f(x) = sum over monomials S subset of bits(x) of coeff[S]
        s = sum(co[S] for S in range(1 << m) if S & ~x == 0) % P
        assert s == ev[x]


# ------------------------------------------------------------------ sumcheck --
def _fold_msb(tab, c):
    h = len(tab) // 2
    return [pf.eadd(tab[i], pf.emul(c, pf.esub(tab[i + h], tab[i]))) for i in range(h)]


def test_product_sumcheck_rounds_brute_force(oracle):
    """sumcheck::prove(1, [a, b], a^2 b, degree 3, eq_factor z): round r's polynomial is
    P_r(X) = sum_x eq(z, (c_<r, X, x)) a(c_<r, X, x)^2 b(c_<r, X, x)."""
    n = 4
    rng = np.random.default_rng(10)
    a = _rand_f(rng, 1 << n)
    bb = _rand_f(rng, 1 << n)
    z = [_rand_ef(rng) for _ in range(n)]
    bench = oracle.Bench(4, seed=11)
    res = bench.sc_a2b(_u32(a), _u32(bb), _u32(pf.to_u32(z)))
    proof = [int(x) for x in res["proof"]]
    assert len(proof) == n * 5 * 8  # degree 4 after the eq factor, no grinding witness
    chal = [[int(v) for v in row] for row in res["challenges"]]
    A, B = [pf.ef(x) for x in a], [pf.ef(x) for x in bb]
    E = pf.eq_table(z)
    for r in range(n):
        co = pf.from_u32(proof[40 * r: 40 * r + 40])
        h = len(A) // 2
        for X in range(5):
            ax = [pf.eadd(A[i], [(v * X) % P for v in pf.esub(A[i + h], A[i])]) for i in range(h)]
            bx = [pf.eadd(B[i], [(v * X) % P for v in pf.esub(B[i + h], B[i])]) for i in range(h)]
            ex = [pf.eadd(E[i], [(v * X) % P for v in pf.esub(E[i + h], E[i])]) for i in range(h)]
            want = pf.ef(0)
            for i in range(h):
                want = pf.eadd(want, pf.emul(ex[i], pf.emul(pf.emul(ax[i], ax[i]), bx[i])))
            got = pf.ef(0)
            for cc in reversed(co):
                got = pf.eadd([v * X % P for v in got], cc)
            assert got == want, (r, X)
        c = chal[r]
        A, B, E = _fold_msb(A, c), _fold_msb(B, c), _fold_msb(E, c)
    assert [[int(v) for v in row] for row in res["final_evals"]] == [A[0], B[0]]
    assert [int(v) for v in res["final_sum"]] == pf.emul(E[0], pf.emul(pf.emul(A[0], A[0]), B[0]))
    bench.close()


# ---------------------------------------------------------------------- WHIR --
# The WHIR schedule (WhirConfig::new) is pinned by the soundness statements it must
# satisfy, evaluated here from the error terms of the WHIR paper (round-by-round
# soundness of each step, in bits) rather than from the oracle's formulas: every
# step reaches the security level with its proof of work, the counts are the
# smallest that do, and the proof of work is the shortfall.
FIELD_BITS = 248  # |EF| = p^8
SEC, MAX_POW = 128, 16


def _eta_bits(sa, r):
    """-log2(eta), the proximity slack: rho/2 at capacity, sqrt(rho)/20 under the
    Johnson bound (r = log2(1/rho))."""
    return {2: r + 1, 1: r / 2 + math.log2(20)}[sa]


def _list_bits(sa, d, r):
    """log2 of the list size: 2^d / (rho eta) at capacity, 1 / (2 eta sqrt(rho))
    under the Johnson bound, 1 in unique decoding."""
    if sa == 0:
        return 0.0
    if sa == 1:
        return r / 2 - 1 + _eta_bits(1, r)
    return d + r + _eta_bits(2, r)


def _query_bits(sa, r):
    """-log2 of the chance that a query misses a far word: 1 - delta = rho (capacity),
    sqrt(rho) (Johnson), (1 + rho) / 2 (unique decoding)."""
    rho = 2.0 ** -r
    return {2: r, 1: r / 2, 0: -math.log2((1 + rho) / 2)}[sa]


def _ood_bits(sa, d, r, s):
    """OOD step: two list codewords agree on s random points w.p. (2^d / |EF|)^s,
    union over the L^2 / 2 pairs."""
    return s * (FIELD_BITS - d) - (2 * _list_bits(sa, d, r) - 1)


def _check_round(sa, d, r, q, ood, pow_bits, next_r):
    psec = SEC - MAX_POW
    qb = _query_bits(sa, r)
    assert q * qb >= psec and (q - 1) * qb < psec, (d, r, q)
    if sa:
        assert _ood_bits(sa, d, next_r, ood) >= SEC and _ood_bits(sa, d, next_r, ood - 1) < SEC
    else:
        assert ood == 0
    # combination of the ood + q constraints: a random linear combination of the
    # list's errors, (ood + q) L / |EF|; the PoW covers the shortfall of the weaker term
    comb = FIELD_BITS - (math.log2(ood + q) + _list_bits(sa, d, next_r) + 1)
    assert pow_bits == max(0, math.ceil(SEC - min(q * qb, comb)))


@pytest.mark.parametrize("sa", [0, 1, 2])
@pytest.mark.parametrize("v", [13, 20, 25])
def test_whir_schedule_meets_the_soundness_bounds(oracle, v, sa):
    f0, f, rate, rs0 = 7, 4, 1, 5
    cfg = oracle.whir_config(v, sa=sa)
    # rounds: fold f0, then f per round, until at most 2^6 coefficients remain
    assert cfg["final_sumcheck_rounds"] <= 6
    assert cfg["final_sumcheck_rounds"] == v - f0 - f * cfg["n_rounds"]
    if cfg["n_rounds"]:
        assert v - f0 - f * (cfg["n_rounds"] - 1) > 6
    if sa:
        assert _ood_bits(sa, v, rate, cfg["commit_ood"]) >= SEC > _ood_bits(sa, v, rate, cfg["commit_ood"] - 1)
    else:
        assert cfg["commit_ood"] == 0
    d, r, logd = v - f0, rate, v + rate
    for i, rd in enumerate(cfg["rounds"]):
        fold = f0 if i == 0 else f
        next_r = r + fold - (rs0 if i == 0 else 1)
        assert (rd["num_variables"], rd["log_inv_rate"], rd["log_domain_size"]) == (d, r, logd)
        _check_round(sa, d, r, rd["num_queries"], rd["ood_samples"], rd["pow_bits"], next_r)
        # folding step: proximity gaps of the 2-function combination, and the sumcheck's
        # degree-2 polynomial vs the list
        if sa == 2:
            prox = FIELD_BITS - (d + 2 * next_r + _eta_bits(2, next_r))
        elif sa == 1:
            prox = FIELD_BITS - (2 * d + 7 * (min(1 + math.log2(10) + next_r / 2, _eta_bits(1, next_r)) - 1))
        else:
            prox = FIELD_BITS - (d + next_r)
        fsum = FIELD_BITS - (_list_bits(sa, d, next_r) + 1)
        assert rd["folding_pow_bits"] == max(0, math.ceil(SEC - min(prox, fsum)))
        d -= f
        r = next_r
        logd -= rs0 if i == 0 else 1
    fq = cfg["final_queries"]
    qb = _query_bits(sa, r)
    assert fq * qb >= SEC - MAX_POW > (fq - 1) * qb
    assert cfg["final_pow"] == max(0, math.ceil(SEC - fq * qb))
    assert cfg["final_log_domain"] == logd
    assert all(0 <= x["pow_bits"] <= MAX_POW for x in cfg["rounds"])


def test_poseidon2_round_numbers_from_the_security_rule():
    """The round counts of Poseidon2KoalaBear<16>/<24> (8 full, 20 / 23 partial) and of
    the reference's BabyBear alternative (x^7, 13 partial rounds at width 16,
    src/examples/poseidon2.rs:38-45) follow from the Poseidon / Poseidon2 round-number
    rule at 128 bits: the smallest R_F * t + R_P satisfying the statistical,
    interpolation and Groebner-basis bounds (with the 2023/537 correction), plus the
    margin of 2 full rounds and 7.5 % partial rounds."""
    def log2binom(n, k):
        return (math.lgamma(n + 1) - math.lgamma(k + 1) - math.lgamma(n - k + 1)) / math.log(2)

    def secure(p, t, rf, rp, a, m):
        n = math.ceil(math.log2(p))
        bounds = [6 if m <= math.floor(math.log2(p) - (a - 1) / 2) * (t + 1) else 10,
                  1 + math.ceil(math.log(2, a) * min(m, n)) + math.ceil(math.log(t, a)) - rp,
                  math.log(2, a) * min(m, math.log2(p)) - rp,
                  t - 1 + math.log(2, a) * min(m / (t + 1), math.log2(p) / 2) - rp,
                  (t - 2 + m / (2 * math.log2(a)) - rp) / (t - 1)]
        third = math.floor(t / 3)
        over = (rf - 1) * t + rp + third + third * rf / 2 + rp + a
        under = third * rf / 2 + rp + a
        return rf >= max(math.ceil(x) for x in bounds) and math.ceil(2 * log2binom(over, under)) >= m

    def rounds(p, t, a):
        best = None
        for rp in range(1, 200):
            for rf in range(4, 40, 2):
                if secure(p, t, rf, rp, a, 128):
                    cand = (rf + 2, math.ceil(rp * 1.075))
                    key = (cand[0] * t + cand[1], cand[0])
                    if best is None or key < best[0]:
                        best = (key, cand)
        return best[1]

    assert rounds(P, 16, 3) == (8, 20)
    assert rounds(P, 24, 3) == (8, 23)
    assert rounds(2**31 - 2**27 + 1, 16, 7) == (8, 13)
    assert math.gcd(3, P - 1) == 1  # x^3 permutes KoalaBear (not BabyBear: 3 | p - 1 there)
    assert math.gcd(3, 2**31 - 2**27) == 3


# ---------------------------------------------------------------- whole proof --
def test_proof_verifies_and_tampering_is_rejected(oracle):
    b = oracle.Bench(8, seed=4)
    proof = b.prove()
    assert b.verify(proof) == 0
    rng = np.random.default_rng(12)
    for pos in sorted(set(rng.integers(0, proof.size, 24).tolist()) | {0, proof.size - 1}):
        bad = proof.copy()
        bad[pos] = (int(bad[pos]) + 1) % P
        assert b.verify(bad) != 0, pos
    assert b.verify(proof[:-1]) != 0
    b.close()
