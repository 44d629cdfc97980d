"""GPU tests of the learned corrector's kernels (BF16 MFMA forward, backward,
Adam), its grid features, training pairs and the fallback, against the CPU
oracle (oracle/mlp.py).

Tolerances.  The GPU rounds inputs, weights, activations and back-propagated
gradients to BF16 (8 significant bits) and accumulates in FP32 in another
order than the oracle; compared with the oracle's BF16 emulation the outputs
agree to a few BF16 units of the output scale, the gradient to 3 % in norm.
"""
import sys

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rand_problem(n, seed):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, 72), np.float32)
    X[:, :66] = rng.normal(size=(n, 66)) * rng.uniform(0.2, 3.0, size=66) + rng.normal(size=66)
    mean = X[:, :66].mean(0).astype(np.float32)
    inv = (1.0 / X[:, :66].std(0)).astype(np.float32)
    return rng, X, mean, inv


def test_forward_matches_oracle(gpu):
    rng, X, mean, inv = rand_problem(1000, 1)  # not a multiple of the 32-node tile
    p = mlp.init_params(2)
    p[mlp.NPARAM - 3:] = [0.3, -0.2, 0.1]
    c = gpu.Corrector(p)
    c.set_norm(mean, inv, 2.5)
    got = c.forward(X)
    ref_q = mlp.forward(X, p, mean, inv, 2.5, emulate_bf16=True)
    ref = mlp.forward(X, p, mean, inv, 2.5, emulate_bf16=False)
    scale = float(np.abs(ref).max())
    err_q = float(np.abs(got - ref_q).max()) / scale
    err = float(np.abs(got - ref).max()) / scale
    say(f"corrector forward: max err vs bf16 oracle {err_q:.2e}, vs fp64 {err:.2e} (of max |y| {scale:.3f})")
    assert err_q < 2e-2
    assert err < 5e-2
    assert float(np.abs(got - ref_q).mean()) / scale < 2e-3
    c.close()


def test_gradient_matches_oracle(gpu):
    rng, X, mean, inv = rand_problem(777, 3)
    T = rng.normal(size=(777, 3)).astype(np.float32)
    p = mlp.init_params(4)
    c = gpu.Corrector(p)
    c.set_norm(mean, inv, 1.5)
    c.load_dataset(X, T)
    g, loss = c.gradient(0, 777)
    L, gref = mlp.loss_and_grad(X, T, p, mean, inv, 1.5, emulate_bf16=True)
    rel = float(np.linalg.norm(g - gref) / np.linalg.norm(gref))
    cos = float(g @ gref / (np.linalg.norm(g) * np.linalg.norm(gref)))
    say(f"corrector gradient: rel err {rel:.2e}, cosine {cos:.6f}, loss {loss:.5f} vs {L:.5f}")
    assert abs(loss - L) < 1e-2 * L
    assert rel < 3e-2 and cos > 0.999
    # per layer: every block agrees, not only the largest
    o = 0
    for name, shape in mlp.SHAPES:
        k = int(np.prod(shape))
        a, b = g[o:o + k], gref[o:o + k]
        assert np.linalg.norm(a - b) <= 5e-2 * np.linalg.norm(b) + 1e-7, name
        o += k
    # wrapping batch: rows [700, 700 + 200) take rows 700..776 and 0..122
    g2, _ = c.gradient(700, 200)
    idx = np.r_[700:777, 0:123]
    _, gref2 = mlp.loss_and_grad(X[idx], T[idx], p, mean, inv, 1.5, emulate_bf16=True)
    assert np.linalg.norm(g2 - gref2) / np.linalg.norm(gref2) < 3e-2
    c.close()


def test_training_fits_a_teacher(gpu):
    """Adam on the device reduces the loss on targets from a random teacher
    network by a large factor (loss-decrease test)."""
    rng, X, mean, inv = rand_problem(8192, 5)
    teacher = mlp.init_params(6, scale=0.7)
    T = mlp.forward(X, teacher, mean, inv, 1.0, emulate_bf16=False).astype(np.float32)
    scale = float(np.sqrt((T ** 2).mean()))
    c = gpu.Corrector(mlp.init_params(7))
    c.set_norm(mean, inv, scale)
    c.load_dataset(X, T)
    l0 = c.gradient(0, 8192)[1]
    losses = c.train(600, 1024, 2e-3)
    l1 = c.gradient(0, 8192)[1]
    say(f"corrector training: loss {l0:.4f} -> {l1:.4f} (steps: first {losses[0]:.4f} last {losses[-1]:.4f})")
    assert np.isfinite(losses).all()
    assert l1 < 0.1 * l0
    c.close()


def test_features_match_oracle(gpu):
    sc = gpu.scenes.sphere64(physical=True)
    e = gpu.Engine(sc["cfg"])
    e.init(sc["particles"].shape[1], sc["particles"])
    e.set_sdf(sc["sdf"])
    e.steps(30)
    nodes, X = e.debug_features(300000)
    grid = e.getGrid()
    sdf = e.get_sdf()
    assert len(nodes) > 1000
    np.testing.assert_array_equal(nodes[:, 1], 0)
    mass = grid[0]
    assert len(nodes) == int((mass > 1e-9).sum())
    rng = np.random.default_rng(0)
    pick = rng.choice(len(nodes), size=400, replace=False)
    for i in pick:
        code = int(nodes[i, 0])
        x, y, z = code & 1023, (code >> 10) & 1023, code >> 20
        ref = mlp.node_features(grid, sdf, sc["cfg"], x, y, z, float(sc["cfg"].DT))
        np.testing.assert_allclose(X[i], ref, rtol=1e-5, atol=1e-6)
    e.close()


def test_engine_correction_matches_forward(gpu):
    """The engine's fused feature + MLP kernel against the feature kernel and
    the standalone forward pass (both checked against the oracle above), with
    random weights: one substep's grid velocity with the corrector minus the
    same substep's without it equals forward(features) at every active node
    (within BF16 units of the output scale; the two engines differ only by
    float-atomic order)."""
    sc = gpu.scenes.sphere64(physical=True)
    a, b = gpu.Engine(sc["cfg"]), gpu.Engine(sc["cfg"])
    for e in (a, b):
        e.init(sc["particles"].shape[1], sc["particles"])
        e.set_sdf(sc["sdf"])
        e.steps(10)
    a.steps(1)
    nodes, X = a.debug_features(300000)
    assert len(nodes) > 1000
    c = gpu.Corrector(mlp.init_params(3, 2.0).astype(np.float32))
    mean = X[:, :66].mean(0).astype(np.float32)
    inv = (1.0 / (X[:, :66].std(0) + 1e-3)).astype(np.float32)
    c.set_norm(mean, inv, 0.5)
    c.set_clamp(1e6)
    b.set_corrector(c)
    b.set_option("CORRECT", 1)
    b.steps(1)
    want = c.forward(X)
    ga, gb = a.getGrid(), b.getGrid()
    sx, sy = int(sc["cfg"].SIZE_X), int(sc["cfg"].SIZE_Y)
    code = nodes[:, 0].astype(np.int64)
    g = (code & 1023) + sx * (((code >> 10) & 1023) + sy * (code >> 20))
    got = (gb[1:4, g] - ga[1:4, g]).T
    scale = float(np.abs(want).max())
    err = float(np.abs(got - want).max())
    say(f"engine correction vs forward: {len(nodes)} nodes, max |correction| {scale:.3g}, max error {err:.3g}")
    assert scale > 1e-2
    assert err < 0.03 * scale
    a.close(); b.close(); c.close()


def test_zero_corrector_changes_nothing(gpu):
    sc = gpu.scenes.sphere64(physical=True)
    a = gpu.Engine(sc["cfg"])
    b = gpu.Engine(sc["cfg"])
    for e in (a, b):
        e.init(sc["particles"].shape[1], sc["particles"])
        e.set_sdf(sc["sdf"])
    c = gpu.Corrector(np.zeros(gpu.CORRECTOR_PARAMS, np.float32))
    b.set_corrector(c)
    b.set_option("CORRECT", 1)
    a.steps(12)
    b.steps(12)
    pa, pb = a.getParticles(), b.getParticles()
    assert float(np.abs(pa[0:6] - pb[0:6]).max()) < 12e-4
    a.close(); b.close(); c.close()


def test_correction_is_added_to_grid_velocity(gpu):
    """A network with only an output bias adds that constant to every active
    node's velocity: one substep then moves every particle (away from the
    collider and walls) by DT * bias more than the plain step."""
    sc = gpu.scenes.sphere64(physical=True)
    bias = np.array([0.5, 0.25, -0.75], np.float32)
    p = np.zeros(gpu.CORRECTOR_PARAMS, np.float32)
    p[-3:] = bias
    c = gpu.Corrector(p)
    a = gpu.Engine(sc["cfg"])
    b = gpu.Engine(sc["cfg"])
    for e in (a, b):
        e.init(sc["particles"].shape[1], sc["particles"])
        e.set_sdf(sc["sdf"])
    b.set_corrector(c)
    b.set_option("CORRECT", 1)
    a.step()
    b.step()
    pa, pb = a.getParticles(), b.getParticles()
    dt = float(sc["cfg"].DT)
    np.testing.assert_allclose(pb[3:6] - pa[3:6], np.tile(bias[:, None], (1, pa.shape[1])), atol=2e-4)
    np.testing.assert_allclose(pb[0:3] - pa[0:3], dt * np.tile(bias[:, None], (1, pa.shape[1])), atol=2e-6)
    a.close(); b.close(); c.close()


def test_collect_pairs(gpu):
    sc = gpu.scenes.sphere64(physical=True)
    a = gpu.Engine(sc["cfg"])
    b = gpu.Engine(sc["cfg"])
    for e in (a, b):
        e.init(sc["particles"].shape[1], sc["particles"])
        e.set_sdf(sc["sdf"])
        e.steps(20)
    X, T = a.collect_pairs(4, 200000)
    b.steps(4)
    say(f"pairs: {len(X)} rows, |target| rms {float(np.sqrt((T ** 2).mean())):.3e}")
    assert len(X) > 1000 and np.isfinite(X).all() and np.isfinite(T).all()
    assert not X[:, 66:].any()
    # the engine has advanced four plain substeps
    pa, pb = a.getParticles(), b.getParticles()
    assert float(np.abs(pa[0:3] - pb[0:3]).max()) < 4e-4
    # the coarse features were taken at 4 DT: feature 62 = GRAVITY (4 DT)^2
    dt4 = 4 * float(sc["cfg"].DT)
    np.testing.assert_allclose(X[:, 62], sc["cfg"].GRAVITY * dt4 * dt4, rtol=1e-5)
    a.close(); b.close()


def test_fallback_rewinds_to_plain_step(gpu):
    """A corrector whose vertical correction grows with height (dv_y/dy of
    about 30/s) stretches F far beyond the density-drift band; the runner
    rewinds each frame and takes the plain step, so the state equals a plain
    run's."""
    from importlib import import_module

    learned = import_module("water-sim-cuda_amd.learned")
    sc = gpu.scenes.sphere64(physical=True)
    d = mlp.unpack(np.zeros(mlp.NPARAM))
    d["W1"][65, 0] = d["W2"][0, 0] = d["W3"][0, 0] = 1.0  # feature 65 = y / SIZE_Y
    d["W4"][0, 1] = 2000.0
    c = gpu.Corrector(mlp.pack(d).astype(np.float32))
    a = learned.make_engine(sc)
    b = learned.make_engine(sc)
    # the same corrector without the fallback does drift
    probe = learned.make_engine(sc)
    probe.set_corrector(c)
    probe.set_option("CORRECT", 1)
    probe.set_dt_scale(learned.RATIO)
    probe.steps(2)
    drift = probe.check(learned.PEN_DEPTH)["mean_abs_J_minus_1"]
    say(f"fallback: corrected drift {drift:.3f} (band {learned.DRIFT_MAX})")
    assert drift > learned.DRIFT_MAX
    probe.close()
    runner = learned.LearnedRunner(b, c, check_every=1)
    for _ in range(3):
        a.steps(8)
        assert not runner.frame(2)
    assert runner.fallbacks == 3
    chk = b.check(learned.PEN_DEPTH)
    assert chk["nonfinite"] == 0 and chk["mean_abs_J_minus_1"] < learned.DRIFT_MAX
    pa, pb = a.getParticles(), b.getParticles()
    assert float(np.abs(pa[0:3] - pb[0:3]).max()) < 24e-4
    a.close(); b.close(); c.close()


def test_fallback_checks_blocks_of_frames(gpu):
    """With checks every third frame, the drifting corrector's frames run
    corrected until the block ends; the block then falls back as a whole and
    the state equals a plain run's."""
    from importlib import import_module

    learned = import_module("water-sim-cuda_amd.learned")
    sc = gpu.scenes.sphere64(physical=True)
    d = mlp.unpack(np.zeros(mlp.NPARAM))
    d["W1"][65, 0] = d["W2"][0, 0] = d["W3"][0, 0] = 1.0
    d["W4"][0, 1] = 2000.0
    c = gpu.Corrector(mlp.pack(d).astype(np.float32))
    a = learned.make_engine(sc)
    b = learned.make_engine(sc)
    runner = learned.LearnedRunner(b, c, check_every=3)
    results = []
    for _ in range(3):
        a.steps(8)
        results.append(runner.frame(2))
    assert results == [True, True, False]
    assert runner.fallbacks == 3 and runner.checks == 1
    assert runner.flush()  # nothing pending
    pa, pb = a.getParticles(), b.getParticles()
    assert float(np.abs(pa[0:3] - pb[0:3]).max()) < 24e-4
    a.close(); b.close(); c.close()


def test_dead_zone_matches_oracle(gpu):
    """Rows whose network output magnitude is below the dead zone give no
    correction; the others are unchanged (checked away from the threshold,
    where BF16 rounding could move a row across it)."""
    _, X, mean, inv = rand_problem(4096, 7)
    p = mlp.init_params(2)
    p[mlp.NPARAM - 3:] = [0.3, -0.2, 0.1]
    c = gpu.Corrector(p)
    c.set_norm(mean, inv, 2.5)
    raw = mlp.forward(X, p, mean, inv, 1.0, emulate_bf16=True)
    mag = np.sqrt((raw ** 2).sum(1))
    dead = float(np.median(mag))
    c.set_deadzone(dead)
    got = c.forward(X)
    ref = mlp.forward(X, p, mean, inv, 2.5, emulate_bf16=True, dead=dead)
    clear = np.abs(mag - dead) > 0.05 * dead
    say(f"dead zone {dead:.3f}: {int((mag < dead).sum())} of {len(X)} rows dropped")
    assert clear.sum() > 0.8 * len(X)
    np.testing.assert_array_equal(got[clear & (mag < dead)], 0.0)
    live = clear & (mag >= dead)
    assert np.abs(got[live] - ref[live]).max() < 0.05 * 2.5 * float(mag.max())
    c.close()


CHECKPOINTS = (0.25, 0.5, 0.75, 1.0)  # simulated seconds


def test_trained_corrector_gains_at_every_checkpoint(gpu):
    """The committed dam-break corrector on held-out seed 1000 (neither a
    training nor a validation seed), 4xDT: the RMS position error against the
    plain DT run, at each checkpoint, is no larger with the correction than
    without it, and after one simulated second it is at least 10 % smaller.
    Both stay far inside 0.5 cells."""
    from importlib import import_module
    from pathlib import Path

    learned = import_module("water-sim-cuda_amd.learned")
    wpath = Path(gpu.__file__).parent / "corrector_dambreak128.npz"
    corr, meta = learned.load(wpath)
    sc = learned.dambreak_family(1000)
    dt = float(sc["cfg"].DT)
    fine = [int(round(t / dt)) for t in CHECKPOINTS]
    ref = learned.make_engine(sc, RESORT_EVERY=learned.PLAIN_RESORT)
    refs, done = [], 0
    for f in fine:
        ref.steps(f - done)
        done = f
        refs.append(ref.getParticles())
    ref.close()
    errs = {}
    for name, use in (("coarse", False), ("learned", True)):
        e = learned.make_engine(sc, RESORT_EVERY=learned.COARSE_RESORT)
        e.set_dt_scale(learned.RATIO)
        if use:
            e.set_corrector(corr)
            e.set_option("CORRECT", 1)
        vals, done = [], 0
        for f, r in zip(fine, refs):
            e.steps(f // learned.RATIO - done)
            done = f // learned.RATIO
            vals.append(learned.rms(e.getParticles(), r))
        errs[name] = vals
        e.close()
    corr.close()
    say(f"trained corrector, held-out seed 1000, RMS cells at {CHECKPOINTS} s: coarse "
        f"{[round(v, 5) for v in errs['coarse']]}, corrected {[round(v, 5) for v in errs['learned']]}")
    for t, c_, l_ in zip(CHECKPOINTS, errs["coarse"], errs["learned"]):
        assert l_ <= c_ * (1.0 + 1e-3) + 1e-6, f"worse with the correction at {t} s: {l_} vs {c_}"
    assert errs["learned"][-1] < 0.9 * errs["coarse"][-1]
    assert max(errs["learned"]) < 0.5
