"""The committed tuning table (tupan_amd/tune/cache.json) against the
library's table as include/tupan_hip.h declares it: every entry names a
kernel id and arithmetic slot the header defines, lies inside the table
(TUPAN_TUNE_SHAPES shapes, TUPAN_TUNE_BUCKETS buckets, 1..64 workgroups per
CU, the range tupan_tune_set accepts) and appears once, and apply() hands
each entry to tupan_tune_set unchanged.  apply() skips an entry the library
rejects, so an out-of-range entry would otherwise go unnoticed."""
import os
import re

from tupan_amd import tune

HEADER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..",
                      "include", "tupan_hip.h")


def defines():
    with open(HEADER) as f:
        text = f.read()
    return {m.group(1): int(m.group(2))
            for m in re.finditer(r"#define (TUPAN_\w+) (\d+)", text)}


def test_kernel_and_slot_ids_match_header():
    d = defines()
    for name, kid in tune.KERNEL_IDS.items():
        assert d["TUPAN_K_" + name.upper()] == kid, name
    for mode, slot in (("fp64", "FP64"), ("split32", "SPLIT32"),
                       ("ieee64", "IEEE64")):
        assert d["TUPAN_MODE_" + slot] == tune.SLOTS[mode], mode
    # the float build's table is slot 3 (tupan_tune_set: slot 0..3)
    assert tune.SLOTS["float32"] == 3


def test_cache_entries_in_range_and_unique():
    d = defines()
    entries = tune.load().get("pairwise", [])
    assert entries
    seen = set()
    for e in entries:
        assert e["kernel"] in tune.KERNEL_IDS, e
        assert e["mode"] in tune.SLOTS, e
        assert 0 <= int(e["bucket"]) < d["TUPAN_TUNE_BUCKETS"], e
        assert 0 <= int(e["shape"]) < d["TUPAN_TUNE_SHAPES"], e
        assert 1 <= int(e["wg_per_cu"]) <= 64, e
        key = (e["kernel"], e["mode"], int(e["bucket"]))
        assert key not in seen, key
        seen.add(key)


def test_apply_passes_every_entry():
    calls = []

    class Lib(object):
        def tupan_tune_set(self, *args):
            calls.append(args)
            return 0

    tune.apply(Lib())
    entries = tune.load()["pairwise"]
    assert calls == [(tune.KERNEL_IDS[e["kernel"]], tune.SLOTS[e["mode"]],
                      int(e["bucket"]), int(e["shape"]),
                      int(e["wg_per_cu"])) for e in entries]


def test_ensemble_lanes_are_a_wave_fraction():
    assert tune.tuned_ensemble_lanes() in (8, 16, 32, 64)
