"""Autotuner of the commitment kernels' launch shapes (wl_autotune_*).

No counterpart in the reference: its CPU prover has no launch shapes.  The
Reed-Solomon NTT split, the Merkle leaf-hashing lane mapping and block size and
the tree-level schedule each have variants that compute identical values; the
library times them on the first use of a shape and keeps the fastest.
"""
import ctypes

from . import check, lib

KINDS = {0: "rs_encode_split", 1: "merkle_leaves", 2: "merkle_tree"}
NTT, LEAVES, TREE = 0, 1, 2


def mode(m=None):
    """Set tuning on (1) or off (0); returns the previous mode.  mode() only queries."""
    return lib().wl_autotune_mode(-1 if m is None else int(m))


def force(kind, log_size, variant, width=0, device=0):
    """Pin the variant of (kind, 2^log_size rows, leaf width) on a device."""
    check(lib().wl_autotune_force(device, kind, log_size, width, variant))


def entries():
    """The tuning table: dicts with kind, log_size, width, variant (-1 while its
    variants are still being timed) and best_us (None for a forced entry)."""
    L = lib()
    n = L.wl_autotune_entries(None, 0)
    buf = (ctypes.c_int64 * (6 * max(1, n)))()
    n = min(n, L.wl_autotune_entries(buf, n))
    out = []
    for i in range(n):
        dev, kind, lg, width, var, ns = buf[6 * i: 6 * i + 6]
        out.append({"device": dev, "kind": KINDS.get(kind, kind), "log_size": lg, "width": width,
                    "variant": var, "best_us": None if ns < 0 else ns / 1e3})
    return out


def clear():
    lib().wl_autotune_clear()


def exploring():
    """Number of (kernel, shape) entries whose variants are still being timed."""
    return sum(1 for e in entries() if e["variant"] < 0)
