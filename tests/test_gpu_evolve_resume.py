"""tupan_amd.evolve --save/--resume: a run stopped at a chunk boundary and
resumed from its saved state ends in the same state, bit for bit, as the
same run made in one go (ahermite4 device steps replayed as graphs,
split-FP32 forces, Plummer N = 1024)."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

pytest.importorskip("torch")

COMMON = ["--n", "1024", "--method", "ahermite4", "--eta", "0.01",
          "--mode", "split32", "--chunk", str(1.0 / 64)]


def run(args, capsys):
    from tupan_amd import evolve
    assert evolve.main(COMMON + args) == 0
    return json.loads(capsys.readouterr().out.strip().splitlines()[-1])


def test_resumed_run_equals_uninterrupted(tmp_path, capsys):
    half, whole, once = (str(tmp_path / f) for f in ("half.npz", "whole.npz",
                                                     "once.npz"))
    print("evolve: first half, resume, one go", flush=True)
    a = run(["--t-end", str(1.0 / 64), "--save", half], capsys)
    b = run(["--t-end", str(1.0 / 32), "--resume", half, "--save", whole],
            capsys)
    c = run(["--t-end", str(1.0 / 32), "--save", once], capsys)
    assert a["t"] == 1.0 / 64 and b["t"] == c["t"] == 1.0 / 32
    assert b["steps"] == c["steps"] > a["steps"] > 0
    assert b["dE_E"] == c["dE_E"]
    with np.load(whole) as x, np.load(once) as y:
        assert sorted(x.files) == sorted(y.files)
        for k in x.files:
            if k != "__meta__":
                np.testing.assert_array_equal(x[k], y[k], err_msg=k)
