"""Autotuned launch parameters.

``cache.json`` (next to this file) holds the choices measured by
``python -m tupan_amd.tune`` on the target GPU; the library falls back to
its built-in defaults for anything the cache does not list.
"""
import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))
CACHE = os.path.join(HERE, "cache.json")
DEFAULT_ENSEMBLE_LANES = 8


def load():
    if os.path.exists(CACHE):
        with open(CACHE) as f:
            return json.load(f)
    return {}


def save(data):
    with open(CACHE, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def tuned_ensemble_lanes():
    return int(load().get("ensemble", {}).get("min_lanes",
                                              DEFAULT_ENSEMBLE_LANES))


SLOTS = {"fp64": 0, "split32": 1, "ieee64": 2, "float32": 3}
KERNEL_IDS = {"phi": 0, "acc": 1, "acc_jerk": 2, "snap_crackle": 3,
              "tstep": 4, "nreg_x": 5, "nreg_v": 6, "pnacc": 7}


def apply(lib):
    """Load the cached pairwise choices into the library's tuning table."""
    for e in load().get("pairwise", []):
        try:
            lib.tupan_tune_set(KERNEL_IDS[e["kernel"]], SLOTS[e["mode"]],
                               int(e["bucket"]), int(e["shape"]),
                               int(e["wg_per_cu"]))
        except (KeyError, ValueError):
            continue
