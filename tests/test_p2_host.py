"""Host build of the prover's Poseidon2 field code (whirlaway_amd/csrc/kb.h,
the templates the kernels use) on the CPU.

The internal linear layer of the partial rounds keeps some entries in lazy
ranges ([0, 2p) for +-2^-k, [0, p + 2^23) for 2^-k with k >= 8); this checks
that layer on inputs at the edges of those ranges against the definition
s_i <- sum(s) + d_i s_i mod p, with d the internal diagonal of Plonky3's
KoalaBear Poseidon2 (widths 16 and 24, as the reference's Poseidon2 AIR and
Merkle hashing use it, src/examples/poseidon2.rs), and checks that every output
stays in its entry's range.  The whole permutation is then compared with the
oracle on states built from field extremes.  Needs hipcc (host-only build).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import pyfield as pf

P = pf.P
HERE = os.path.dirname(os.path.abspath(__file__))

# (sign, kind, value): kind 0 is the integer value, kind 1 is 2^-value
DIAG = {
    16: [(-1, 0, 2), (1, 0, 1), (1, 0, 2), (1, 1, 1), (1, 0, 3), (1, 0, 4), (-1, 1, 1), (-1, 0, 3),
         (-1, 0, 4), (1, 1, 8), (1, 1, 3), (1, 1, 24), (-1, 1, 8), (-1, 1, 3), (-1, 1, 4), (-1, 1, 24)],
    24: [(-1, 0, 2), (1, 0, 1), (1, 0, 2), (1, 1, 1), (1, 0, 3), (1, 0, 4), (-1, 1, 1), (-1, 0, 3),
         (-1, 0, 4), (1, 1, 8), (1, 1, 2), (1, 1, 3), (1, 1, 4), (1, 1, 5), (1, 1, 6), (1, 1, 24),
         (-1, 1, 8), (-1, 1, 3), (-1, 1, 4), (-1, 1, 5), (-1, 1, 6), (-1, 1, 7), (-1, 1, 9), (-1, 1, 24)],
}


def d_value(e):
    sign, kind, val = e
    d = val if kind == 0 else pow(pow(2, val, P), P - 2, P)
    return sign * d % P


def bound(e):
    """exclusive upper bound of the entry's value between internal layers"""
    sign, kind, val = e
    if kind == 1 and val >= 2 and (sign < 0 or val < 8):
        return 2 * P
    if kind == 1 and val >= 8:
        return P + (1 << 23)
    return P


@pytest.fixture(scope="module")
def host_bin(tmp_path_factory):
    hipcc = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else None)
    if hipcc is None:
        pytest.skip("hipcc not available")
    out = str(tmp_path_factory.mktemp("p2host") / "p2_host_check")
    subprocess.run([hipcc, "-x", "hip", "--offload-host-only", "-O2", "-o", out,
                    os.path.join(HERE, "native", "p2_host_check.cpp")], check=True, capture_output=True)
    return out


def run(binary, header, rows):
    text = header + "\n" + "\n".join(" ".join(str(int(v)) for v in r) for r in rows) + "\n"
    res = subprocess.run([binary], input=text, capture_output=True, text=True, check=True)
    return np.array([[int(v) for v in line.split()] for line in res.stdout.splitlines()], dtype=np.int64)


def edge_values(rng, b, k):
    """values below the bound b at its edges, at the edges of p, at the split point
    of x 2^-k (low k bits all ones or zero), plus random ones"""
    vals = [0, 1, 2, b - 1, b - 2, P - 1, P - 2, (P + 1) // 2, (1 << 23) - 1, 1 << 24, (1 << 31) - 1, 1 << 30]
    if b > P:
        vals += [P, P + 1, (b + P) // 2]
    if k:
        lows = (1 << k) - 1
        vals += [lows, (b - 1) | lows, (b - 1) & ~lows, ((b - 1) >> k << k) - 1]
    vals += list(rng.integers(0, b, 8))
    return [v for v in vals if 0 <= v < b]


@pytest.mark.parametrize("width", [16, 24])
def test_internal_layer_edges(host_bin, width):
    rng = np.random.default_rng(width)
    diag = DIAG[width]
    cands = [edge_values(rng, bound(e), e[2] if e[1] == 1 else 0) for e in diag]
    rows = [[c[0] for c in cands], [c[3] for c in cands], [max(c) for c in cands]]
    rows += [[c[rng.integers(len(c))] for c in cands] for _ in range(3000)]
    got = run(host_bin, f"mix {width} {len(rows)}", rows)
    d = [d_value(e) for e in diag]
    lim = [bound(e) for e in diag]
    for r, out in zip(rows, got):
        s = sum(r) % P
        for i in range(width):
            assert out[i] % P == (s + d[i] * r[i]) % P, (width, i, r)
            assert 0 <= out[i] < lim[i], (width, i, r, out[i])


def _field_edge_states(rng, width, n_random):
    special = [0, 1, 2, P - 1, P - 2, (P + 1) // 2, (P - 1) // 2]
    special += [(1 << k) + d for k in range(1, 31) for d in (-1, 0, 1)]
    special = np.array([v for v in special if 0 <= v < P], dtype=np.uint32)
    rows = [np.full(width, v, dtype=np.uint32) for v in special]
    rows += [rng.choice(special, width) for _ in range(200)]
    rows += list(rng.integers(0, P, (n_random, width), dtype=np.uint32))
    return np.array(rows, dtype=np.uint32)


@pytest.mark.parametrize("width", [16, 24])
def test_host_permutation_matches_oracle(host_bin, oracle, width):
    b = oracle.Bench(5, seed=3)
    consts = b.p2_consts(width)
    st = _field_edge_states(np.random.default_rng(7), width, 300)
    got = run(host_bin, f"perm {width} {st.shape[0]}\n" + " ".join(str(int(v)) for v in consts), st)
    for i in range(st.shape[0]):
        assert np.array_equal(got[i], b.permute(width, st[i]).astype(np.int64)), (width, i)
