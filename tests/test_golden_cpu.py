"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py): small inputs of the hot path and the oracle's
FP64 outputs after a fixed number of server-sequence steps.  The oracle must
reproduce them exactly (it is built without floating-point contraction), and
the fixtures must be self-consistent."""
import glob
import importlib.util
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))
OUTPUTS = ("sched", "t_end", "dt_end", "h_out", "qx_out", "qy_out", "h_max_out")


def load_golden_module():
    spec = importlib.util.spec_from_file_location("make_golden", os.path.join(HERE, "golden", "make_golden.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def load(path):
    with np.load(path, allow_pickle=False) as f:
        d = {k: f[k] for k in f.files}
    inputs = {k: v for k, v in d.items() if k not in OUTPUTS}
    return inputs, {k: d[k] for k in OUTPUTS}


def test_fixtures_present():
    names = {os.path.basename(p) for p in FIXTURES}
    assert {"dam_break_32_order2.npz", "rainfall_runoff_40_order1.npz", "boundary_mix_24x31_order2.npz",
            "cell0_stations_24_order2.npz"} <= names


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_golden(oracle_lib, path):
    inputs, expect = load(path)
    got = load_golden_module().oracle_run(inputs)
    for k in OUTPUTS:
        assert np.array_equal(np.asarray(got[k]), expect[k]), k


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_golden_fixtures_are_consistent(path):
    inputs, out = load(path)
    n_valid = int((inputs["mask"] > 0).sum())
    assert out["h_out"].shape == (n_valid,) and out["sched"].shape == (int(inputs["steps"]), 2)
    for k in ("h_out", "qx_out", "qy_out", "h_max_out"):
        assert np.isfinite(out[k]).all()
    assert (out["h_max_out"] >= out["h_out"]).all()
    # t advances by the next step's dt (SWE_CUDA.py:831): t_k+1 = t_k + dt_k+1
    t, dt = out["sched"][:, 0], out["sched"][:, 1]
    assert np.allclose(t[1:], t[:-1] + dt[1:], rtol=0, atol=1e-12 * max(1.0, t[-1]))
    assert out["t_end"] == pytest.approx(t[-1] + float(out["dt_end"]), rel=1e-14)
    assert (dt > 0).all()
