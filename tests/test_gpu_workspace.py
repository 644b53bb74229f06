"""The library's growable scratch workspace (j-split partial sums) shared by
calls on different streams, and its release.

  - Two large AccJerk calls on two torch streams, enqueued back to back
    without host synchronisation, each with its own inputs: the second
    stream's call is ordered after the first's use of the shared workspace,
    so both equal the same calls made one after the other on one stream,
    bit for bit.
  - tupan_release_workspaces frees the workspaces (current and outgrown);
    the next call allocates afresh and gives the same results.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FIELDS = ("mass", "rx", "ry", "rz", "eps2", "vx", "vy", "vz")


def state(n, seed):
    rng = np.random.default_rng(seed)
    d = {k: rng.standard_normal(n) for k in FIELDS}
    d["mass"] = rng.random(n) / n
    d["eps2"] = np.full(n, 1e-6)
    return [torch.from_numpy(d[k]).to("cuda:0") for k in FIELDS]


def enqueue(fields, n, out):
    """AccJerk (FP64) of `fields` on torch's current stream into `out`."""
    from tupan_amd import _native
    from tupan_amd.lib.hip_backend import current_stream_handle
    ip = _native.ptr_array([f.data_ptr() for f in fields])
    op = _native.ptr_array([o.data_ptr() for o in out])
    rc = _native.lib().tupan_pairwise(2, _native.MODES["fp64"], n, ip, n, ip, op,
                                      0.0, current_stream_handle())
    _native.check(rc, "tupan_pairwise")


def fresh(n):
    return [torch.full((n,), np.nan, dtype=torch.float64, device="cuda:0")
            for _ in range(6)]


def test_calls_on_two_streams_share_the_workspace_in_order():
    n = 32768
    a, b = state(n, 1), state(n, 2)
    ra, rb = fresh(n), fresh(n)
    enqueue(a, n, ra)
    enqueue(b, n, rb)
    torch.cuda.synchronize()
    ref = [[x.cpu().numpy() for x in r] for r in (ra, rb)]
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    for rep in range(3):
        print("two streams, repetition %d, n = %d" % (rep, n), flush=True)
        oa, ob = fresh(n), fresh(n)
        s1.wait_stream(torch.cuda.current_stream())
        s2.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s1):
            enqueue(a, n, oa)
        with torch.cuda.stream(s2):
            enqueue(b, n, ob)
        torch.cuda.synchronize()
        for got, want in zip((oa, ob), ref):
            for g, w in zip(got, want):
                np.testing.assert_array_equal(g.cpu().numpy(), w)


def test_release_workspaces_then_call_again():
    from tupan_amd import _native
    n = 40000
    a = state(n, 3)
    r1 = fresh(n)
    enqueue(a, n, r1)
    torch.cuda.synchronize()
    _native.release_workspaces()
    _native.release_workspaces()         # nothing left to free: still fine
    r2 = fresh(n)
    enqueue(a, n, r2)
    torch.cuda.synchronize()
    for x, y in zip(r1, r2):
        np.testing.assert_array_equal(x.cpu().numpy(), y.cpu().numpy())
