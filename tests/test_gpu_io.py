"""Worldlines and snapshots of GPU runs in the reference's HDF5 layout
(tupan/io/hdf5io.py:22-146, the drivers' dumpper calls at
hermite.py:402-405 and sia.py:1114-1117): host-driven steps with HIP forces
on Plummer N = 256 write every particle's state after each of its steps;
the run with a dumpper ends in the same state, bit for bit, as the run
without; tupan_amd.evolve reads its initial conditions from a snapshot and
writes its worldline and final snapshot."""
import json

import numpy as np
import pytest
import torch

from tupan_amd.io import h5lib

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not h5lib.available(),
                                 reason="HDF5 C library not found")]

BASE = ["id", "mass", "eps2", "rx", "ry", "rz", "vx", "vy", "vz", "time",
        "nstep", "tstep"]


def plummer(n=256, seed=3):
    from tupan_amd.ics.plummer import make_plummer
    ps = make_plummer(n, 1.0 / n, ("equalmass",), seed=seed, device="cuda:0")
    type(ps).t_curr = 0.0
    return ps


def run(method, eta, steps, dumpper=None, **kw):
    from tupan_amd.integrator import Integrator
    from tupan_amd.lib import extensions
    extensions.install("HIP", "float64")
    ps = plummer()
    it = Integrator(eta, 0.0, ps, method=method, dumpper=dumpper, **kw)
    t_end = steps * eta
    while it.time < t_end:
        it.evolve_step(t_end)
    torch.cuda.synchronize()
    return it.particle_system


def check_worldline(fn, first, final):
    from tupan_amd.io import IO
    wl = IO(fn, "r").load_worldline(device="cpu")
    ids = wl.id.numpy()
    nst = wl.nstep.numpy()
    for row in final:
        mine = np.nonzero(ids == row["id"])[0]
        assert len(mine) == row["nstep"] + 1
        assert nst[mine].tolist() == list(range(int(row["nstep"]) + 1))
        start = first[first["id"] == row["id"]][0]
        for name in BASE:
            col = getattr(wl, name).numpy()
            assert col[mine[0]] == start[name], name
            assert col[mine[-1]] == row[name], name
    return wl


@pytest.mark.parametrize("method,eta,kw", [
    ("ahermite4", 1.0 / 64, {}),
    ("sia21h.dkd", 1.0 / 32, {"device_subtree": False}),
])
def test_gpu_worldline(tmp_path, method, eta, kw):
    from tupan_amd.io import IO
    print("worldline", method, "N = 256", flush=True)
    fn = str(tmp_path / "wl.hdf5")
    ps = run(method, eta, 4, dumpper=IO(fn, "a"), **kw)
    ref = run(method, eta, 4, **kw)
    oa, ob = torch.argsort(ps.id), torch.argsort(ref.id)
    for name in BASE:
        assert torch.equal(getattr(ps, name)[oa], getattr(ref, name)[ob]), \
            name
    check_worldline(fn, plummer().get_state(), ps.get_state())


def test_evolve_reads_and_writes_hdf5(tmp_path, capsys):
    from tupan_amd import evolve
    from tupan_amd.io import IO
    ic, wl, fin = (str(tmp_path / f) for f in ("ic.hdf5", "wl.hdf5",
                                               "fin.hdf5"))
    print("evolve --input --output --snapshot, hermite4 N = 256", flush=True)
    IO(ic, "w").dump_snapshot(plummer())
    args = ["--input", ic, "--output", wl, "--snapshot", fin,
            "--method", "hermite4", "--eta", str(1.0 / 64),
            "--t-end", str(4.0 / 64), "--chunk", str(4.0 / 64),
            "--graph-steps", "0"]
    assert evolve.main(args) == 0
    out = json.loads(capsys.readouterr().out.strip().splitlines()[-1])
    assert out["n"] == 256 and out["steps"] == 4 and out["t"] == 4.0 / 64
    final = IO(fin, "r").load_snapshot(device="cpu").get_state()
    check_worldline(wl, IO(ic, "r").load_snapshot(device="cpu").get_state(),
                    final)


def test_gpu_diagnostic_report(tmp_path):
    """The reference's report file (simulation.py:40-129) from device
    Hermite steps: the header, then one 13-column line per report_freq
    reports, te = ke + pe as printed, eerr = (te - te0) / (-pe) against the
    energies the system reports at the same time."""
    from tupan_amd.diagnostics import Diagnostic
    from tupan_amd.integrator import Integrator
    from tupan_amd.lib import extensions
    extensions.install("HIP", "float64")
    print("Diagnostic from device ahermite4 steps, N = 256", flush=True)
    ps = plummer()
    fname = str(tmp_path / "diag.txt")
    rep = Diagnostic(fname, 0.0, ps, report_freq=2)
    it = Integrator(1.0 / 64, 0.0, ps, method="ahermite4", reporter=rep,
                    device_steps=True)
    for k in range(6):
        it.evolve_step(1.0)
    torch.cuda.synchronize()
    lines = open(fname).read().splitlines()
    assert lines[0].split() == list(Diagnostic.HEADER)
    body = [[float(x) for x in ln.split()] for ln in lines[1:]]
    assert len(body) == 4 and all(len(r) == 13 for r in body)
    assert body[0][0] == 0.0 and body[0][6] == 0.0
    for r in body:
        t, dt, ke, pe, te, vir, eerr = r[:7]
        assert abs(te - (ke + pe)) < 1e-3 * abs(te)
    last = rep.rows[-1]
    q = it.particle_system
    assert last["time"] == it.time
    assert abs(last["te"] - (q.kinetic_energy + q.potential_energy)) < \
        1e-12 * abs(last["te"])
    assert last["eerr"] == (last["te"] - rep.te0) / (-last["pe"])
    assert abs(last["eerr"]) < 1e-6
