"""Committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py from the oracle): the oracle must reproduce them
(CPU), and the engine must match them at the parity tolerances (GPU)."""
from pathlib import Path

import numpy as np
import pytest

from helpers import keys_numpy

GOLDEN = Path(__file__).resolve().parent / "golden" / "substep_box32.npz"


def load(mpm):
    d = np.load(GOLDEN, allow_pickle=False)
    c = d["cfg"]
    cfg = mpm.make_config(SIZE_X=int(c[0]), SIZE_Y=int(c[1]), SIZE_Z=int(c[2]), PADDING=int(c[3]),
                          GRAVITY=float(c[4]), DT=float(c[5]), GAMMA=float(c[6]), COMPRESSION=float(c[7]),
                          RESOLUTION=float(c[8]))
    return cfg, d


def test_oracle_reproduces_golden(mpm, oracle):
    cfg, d = load(mpm)
    st = oracle.init_state(cfg, d["particles"], np.float32)
    np.testing.assert_allclose(oracle.steps(cfg, st, d["sdf"], 1), d["step1"], rtol=0, atol=1e-6)
    np.testing.assert_allclose(oracle.steps(cfg, st, d["sdf"], 5), d["step5"], rtol=0, atol=1e-5)
    np.testing.assert_array_equal(oracle.morton_keys(d["particles"][0:3]), d["keys"])
    np.testing.assert_array_equal(keys_numpy(d["particles"][0:3]), d["keys"])
    np.testing.assert_array_equal(oracle.sort_perm(d["keys"]), d["perm"])
    # FP32 vs FP64 restatement: one substep agrees far inside the 1e-4-cell bar
    assert float(np.abs(d["step1"][0:3] - d["step1_f64"][0:3]).max()) < 1e-5


@pytest.mark.gpu
def test_engine_matches_golden(gpu):
    cfg, d = load(gpu)
    n = d["particles"].shape[1]
    e = gpu.Engine(cfg)
    e.init(n, d["particles"])
    e.set_sdf(d["sdf"])
    keys, perm = e.sortParticles()
    np.testing.assert_array_equal(keys, d["keys"][d["perm"]])
    np.testing.assert_array_equal(perm, d["perm"])
    e.step()
    got = e.getParticles()
    assert float(np.abs(got[0:3] - d["step1"][0:3]).max()) < 1e-4
    assert float(np.abs(got[3:6] - d["step1"][3:6]).max()) < 1e-4
    e.steps(4)
    got = e.getParticles()
    assert float(np.abs(got[0:3] - d["step5"][0:3]).max()) < 5e-4
    e.close()
