"""Autotuner table API (no GPU: forced entries only)."""
import pytest


def test_force_entries_clear():
    from whirlaway_amd import WhirlawayError, autotune
    autotune.clear()
    autotune.force(autotune.LEAVES, 19, 2, width=128, device=0)
    autotune.force(autotune.NTT, 17, 1, width=99)  # width counts only for leaves
    e = {(x["kind"], x["log_size"]): x for x in autotune.entries()}
    assert e[("merkle_leaves", 19)]["variant"] == 2 and e[("merkle_leaves", 19)]["width"] == 128
    assert e[("rs_encode_split", 17)]["width"] == 0 and e[("rs_encode_split", 17)]["best_us"] is None
    for bad in ((autotune.NTT, 17, 3), (autotune.TREE, 17, 6), (autotune.LEAVES, 17, -1), (7, 17, 0)):
        with pytest.raises(WhirlawayError):
            autotune.force(*bad)
    autotune.clear()
    assert autotune.entries() == []


def test_mode_switch():
    from whirlaway_amd import autotune
    prev = autotune.mode()
    assert prev in (0, 1)
    assert autotune.mode(0) == prev
    assert autotune.mode() == 0
    autotune.mode(prev)
    assert autotune.mode() == prev
