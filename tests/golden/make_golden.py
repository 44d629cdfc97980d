"""Writes the committed golden fixtures of tests/golden/ from the oracle.

    python tests/golden/make_golden.py

The reference holds no golden vectors or fixtures; these are produced by the
oracle (oracle/oracle.hpp, a restatement of the reference's step) on small
seeded scenes and committed, so a change in the oracle or the engine shows up
as a fixture mismatch:

  substep_box32.npz  32^3 grid, jittered 6x6x6-cell block (1,728 particles) above
                     a box floor, physical parameters; state after 1 and 5
                     substeps (FP32 serial oracle) and after 1 substep (FP64),
                     Morton keys and the stable sort permutation.
"""
import importlib
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))


def main() -> None:
    m = importlib.import_module("water-sim-cuda_amd")
    from oracle import oracle

    sc = m.scenes.small_box(n_cells=(6, 6, 6), grid=32, physical=True, seed=7, dt=1e-3)
    cfg, parts, sdf = sc["cfg"], sc["particles"], sc["sdf"]
    st = oracle.init_state(cfg, parts, np.float32)
    one = oracle.steps(cfg, st, sdf, 1)
    five = oracle.steps(cfg, st, sdf, 5)
    one64 = oracle.steps(cfg, oracle.init_state(cfg, parts, np.float64), sdf.astype(np.float64), 1)
    keys = oracle.morton_keys(parts[0:3])
    perm = oracle.sort_perm(keys)
    cfg_vals = np.array([cfg.SIZE_X, cfg.SIZE_Y, cfg.SIZE_Z, cfg.PADDING, cfg.GRAVITY, cfg.DT, cfg.GAMMA,
                         cfg.COMPRESSION, cfg.RESOLUTION], dtype=np.float64)
    out = Path(__file__).resolve().parent / "substep_box32.npz"
    np.savez_compressed(out, cfg=cfg_vals, particles=parts, sdf=sdf, step1=one, step5=five, step1_f64=one64,
                        keys=keys, perm=perm)
    print(f"wrote {out} ({out.stat().st_size} bytes)")


if __name__ == "__main__":
    main()
