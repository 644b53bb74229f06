"""tupan_pairwise_noalloc, the pairwise entry point that allocates nothing,
against tupan_pairwise through the C ABI.

Two calls in a row on each (kernel, mode), the second with a smaller input:
  - the first (N = 40000) needs more j-split partial sums than the fixed
    64 MiB array holds for AccJerk (the tuned plans run 50 to 64 splits x 6
    fields x 40000 doubles, 96 to 123 MB), so the no-alloc plan runs 34
    j-splits: FP64 results within 1e-13 per particle (vector norm) of
    tupan_pairwise's;
  - the second (N = 3000) fits: bit-identical in every mode.
A 1024-particle call takes the one-launch small-set kernel, which uses no
partial sums: bit-identical.  Arguments are checked before any device work.
"""
import ctypes

import numpy as np
import pytest

from conftest import vec_rel_err

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NOALLOC_BYTES = 64 * 1024 * 1024
FIELDS = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz",
          "ax", "ay", "az", "jx", "jy", "jz")


def state(n, seed):
    rng = np.random.default_rng(seed)
    d = {k: rng.standard_normal(n) for k in FIELDS}
    d["mass"] = rng.random(n) / n
    d["eps2"] = np.full(n, 1e-6)
    return {k: torch.from_numpy(v).to("cuda:0") for k, v in d.items()}


def run(fn, kernel, mode, st, n, eta=1.0 / 64):
    from tupan_amd import _native
    nin = {0: 5, 1: 5, 2: 8, 3: 14, 4: 8}[kernel]
    nout = {0: 1, 1: 3, 2: 6, 3: 6, 4: 2}[kernel]
    fields = [st[k] for k in FIELDS[:nin]]
    out = [torch.full((n,), np.nan, dtype=torch.float64, device="cuda:0")
           for _ in range(nout)]
    ip = _native.ptr_array([f.data_ptr() for f in fields])
    op = _native.ptr_array([o.data_ptr() for o in out])
    rc = fn(kernel, _native.MODES[mode], n, ip, n, ip, op, eta, None)
    _native.check(rc, fn.__name__)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


def both(kernel, mode, st, n):
    from tupan_amd import _native
    lib = _native.lib()
    print("noalloc kernel=%d mode=%s n=%d" % (kernel, mode, n), flush=True)
    ref = run(lib.tupan_pairwise, kernel, mode, st, n)
    got = run(lib.tupan_pairwise_noalloc, kernel, mode, st, n)
    return ref, got


def groups(kernel, vals):
    if kernel in (2, 3):
        return [vals[:3], vals[3:]]
    if kernel == 1:
        return [vals]
    return [[v] for v in vals]


@pytest.mark.parametrize("kernel", [2, 1, 4])
def test_noalloc_two_calls_second_smaller(kernel):
    from tupan_amd import _native
    big, small = 40000, 3000
    sb = state(big, 11)
    if kernel == 2:
        # even 35 splits of AccJerk partials overflow the array
        assert 35 * 6 * big * 8 > NOALLOC_BYTES
    ref, got = both(kernel, "fp64", sb, big)
    for r, g in zip(groups(kernel, ref), groups(kernel, got)):
        assert np.all(np.isfinite(np.stack(g)))
        assert vec_rel_err(g, r) <= 1e-13
    ss = state(small, 12)
    for mode in ("fp64", "split32", "ieee64"):
        ref, got = both(kernel, mode, ss, small)
        for r, g in zip(ref, got):
            assert np.array_equal(r, g), mode


def test_noalloc_small_set_kernel_bitwise():
    st = state(1024, 13)
    for kernel in (0, 2, 3):
        ref, got = both(kernel, "fp64", st, 1024)
        for r, g in zip(ref, got):
            assert np.array_equal(r, g)


def test_noalloc_rejects_bad_arguments():
    from tupan_amd import _native
    lib = _native.lib()
    st = state(16, 14)
    ip = _native.ptr_array([st[k].data_ptr() for k in FIELDS[:8]])
    nulls = _native.ptr_array([None] * 6)
    rc = lib.tupan_pairwise_noalloc(2, 0, 16, ip, 16, ip, nulls, 0.0, None)
    assert rc == -1 and "null" in _native.last_error()
    out = [torch.empty(16, dtype=torch.float64, device="cuda:0")
           for _ in range(3)]
    op = _native.ptr_array([o.data_ptr() for o in out])
    rc = lib.tupan_pairwise_noalloc(7, 0, 16, ip, 16, ip, op, 0.0, None)
    assert rc == -1 and "PNAcc" in _native.last_error()
    assert lib.tupan_pairwise_noalloc(99, 0, 16, ip, 16, ip, op, 0.0,
                                      None) == -1
