"""GPU tests of the chunk layout built by the re-sort (no reference
counterpart: the reference scatters particles in Morton order, engine.cu:54-70,
without chunks or rounds).

A chunk holds particles of one 8 x 4 x 4 block of base cells (trunc(x - 0.5)
per axis, clamped into the grid).  With CHUNK large enough that no block is
split, the chunk count is the number of non-empty blocks and a chunk has
max(largest cell, ceil(count / 64)) rounds: the fewest that keep the base
cells of each round distinct within 64 lanes, or, where the fullest cell sets
that count and allowing two members of a cell per round (paired in adjacent
lanes by the transfer) saves at least 30 % of the rounds,
max(ceil(largest cell / 2), ceil(count / 64)).  Both are computed here with
numpy from the initial positions.  Right after a re-sort the lanes of a round
that claim a base cell (all but the second lane of each same-cell pair) hold
distinct cells, so the first P2G needs no extra claim pass and no particle
leaves its tile; a step() also runs the next substep's P2G (left in
the grid for the next call), where only particles whose base cell changed can
lose a claim and only those whose block changed can leave their tile.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def expected_layout(cfg, pos, cap=None):
    """(chunks, rounds) of one scene's particles; pos is 3 x n float32.
    With cap, blocks split into ceil(count / cap) chunks and rounds is None
    (the split of a block depends on atomic order)."""
    size = np.array([cfg.SIZE_X, cfg.SIZE_Y, cfg.SIZE_Z], dtype=np.int64)[:, None]
    base = np.trunc(pos.astype(np.float32) - np.float32(0.5)).astype(np.int64)
    base = np.clip(base, 0, size - 1)
    nbx, nby = (cfg.SIZE_X + 7) // 8, (cfg.SIZE_Y + 3) // 4
    block = (base[0] >> 3) + nbx * ((base[1] >> 2) + nby * (base[2] >> 2))
    cell = (base[0] & 7) + 8 * (base[1] & 3) + 32 * (base[2] & 3)
    keys, cnt = np.unique(block * 128 + cell, return_counts=True)
    blocks, inv = np.unique(keys // 128, return_inverse=True)
    n_b = np.bincount(inv, weights=cnt).astype(np.int64)
    if cap is not None:
        return int(((n_b + cap - 1) // cap).sum()), None
    max_b = np.zeros(len(blocks), dtype=np.int64)
    np.maximum.at(max_b, inv, cnt)
    one = np.maximum(max_b, (n_b + 63) // 64)
    # two members of a cell per round where that saves at least 30 % of the rounds
    two = np.maximum((max_b + 1) // 2, (n_b + 63) // 64)
    return len(blocks), int(np.where(100 * two <= 70 * one, two, one).sum())


def moved_particles(cfgs, counts, before, after):
    """(particles whose base cell changed, those whose 8 x 4 x 4 block
    changed) between two 26 x n buffers of consecutive scenes."""
    moved = left = lo = 0
    for cfg, cnt in zip(cfgs, counts):
        hi = lo + int(cnt)
        size = np.array([cfg.SIZE_X, cfg.SIZE_Y, cfg.SIZE_Z], dtype=np.int64)[:, None]
        b0, b1 = (np.clip(np.trunc(x[0:3, lo:hi].astype(np.float32) - np.float32(0.5)).astype(np.int64), 0, size - 1)
                  for x in (before, after))
        blk = np.array([[8], [4], [4]])
        moved += int(np.any(b0 != b1, axis=0).sum())
        left += int(np.any(b0 // blk != b1 // blk, axis=0).sum())
        lo = hi
    return moved, left


def test_layout_single_scene(gpu):
    sc = gpu.scenes.sphere64(physical=False)
    p = sc["particles"]
    e = gpu.Engine(sc["cfg"], CHUNK=2048)
    e.init(p.shape[1], p)
    e.set_sdf(sc["sdf"])
    chunks, rounds = expected_layout(sc["cfg"], p[0:3])
    got = e.counters()
    say(f"sphere64 layout: n={p.shape[1]} chunks {got['chunks']} (expect {chunks}), rounds {got['rounds']} "
        f"(expect {rounds}), lane use {p.shape[1] / (64 * rounds):.3f}")
    assert got["chunks"] == chunks
    assert got["rounds"] == rounds
    e.step()
    got = e.counters()
    moved, left = moved_particles([sc["cfg"]], [p.shape[1]], p, e.getParticles())
    say(f"after one step: claim losers {got['claim_losers']} (base cell changed: {moved}), "
        f"outside their tile {got['tile_overflow_particles']} (block changed: {left})")
    # of the lanes claiming a base cell at most one kept its cell (the layout
    # gave each claiming lane of a round its own), so each extra lane is a
    # moved one
    assert got["claim_losers"] <= moved
    assert got["tile_overflow_particles"] <= left
    e.close()


@pytest.mark.parametrize("cap", [64, 2048])
def test_layout_odd_size_batch(gpu, cap):
    """Scenes with an odd number of 4^3 blocks in x: their last block per
    row is a single 4^3 block."""
    b = gpu.scenes.batch_scenes(count=5, seed=23, n_min=2000, n_max=40000, edges=(21, 37, 45))
    eb = gpu.Batch(b["cfgs"], CHUNK=cap)
    eb.init(b["counts"], b["particles"])
    for s, sdf in enumerate(b["sdfs"]):
        eb.set_scene_sdf(s, sdf)
    chunks = rounds = 0
    lo = 0
    for s, cfg in enumerate(b["cfgs"]):
        hi = lo + int(b["counts"][s])
        c, r = expected_layout(cfg, b["particles"][0:3, lo:hi], cap if cap < 2048 else None)
        chunks += c
        rounds += r or 0
        lo = hi
    got = eb.counters()
    say(f"odd batch layout, CHUNK {cap}: chunks {got['chunks']} (expect {chunks}), rounds {got['rounds']}"
        + ("" if cap < 2048 else f" (expect {rounds})"))
    assert got["chunks"] == chunks
    if cap == 2048:
        assert got["rounds"] == rounds
    eb.step()
    moved, _ = moved_particles(b["cfgs"], b["counts"], b["particles"], eb.getParticles())
    assert eb.counters()["claim_losers"] <= moved
    eb.close()


def test_layout_many_chunks_per_workgroup(gpu):
    """The re-sort's reorder runs as many workgroups as fit on the device, each
    taking every gridDim-th chunk and loading the next one ahead.  The 128^3 dam
    break (1.77 M particles) at CHUNK 64 has about 28 k chunks, many per
    workgroup: the full re-sort of init must return every particle to the
    caller unchanged, the chunk count must match the numpy layout, and steps
    with a re-sort every substep (the reorder without the planes) must agree
    with whole-block chunks within float-atomic order."""
    sc = gpu.scenes.dambreak128()
    p = sc["particles"]
    n = p.shape[1]
    say(f"many chunks: dambreak128 n={n}, CHUNK 64 vs 2048, init round trip then 3 substeps")
    small = gpu.Engine(sc["cfg"], CHUNK=64, RESORT_EVERY=1)
    small.init(n, p)
    expect = p.copy()
    expect[[15, 19, 23]] = 1.0  # init sets F = I (the reference's initFMatrices)
    np.testing.assert_array_equal(small.getParticles(), expect)
    chunks, _ = expected_layout(sc["cfg"], p[0:3], 64)
    got = small.counters()
    say(f"chunks {got['chunks']} (expect {chunks})")
    assert got["chunks"] == chunks
    whole = gpu.Engine(sc["cfg"], CHUNK=2048, RESORT_EVERY=1)
    whole.init(n, p)
    for e in (small, whole):
        e.initSDF(sc["triangles"])
        e.steps(3)
    a, b = small.getParticles(), whole.getParticles()
    assert np.isfinite(a).all()
    err = float(np.abs(a[0:3].astype(np.float64) - b[0:3].astype(np.float64)).max())
    say(f"max position difference {err:.2e}")
    assert err < 3e-4
    small.close()
    whole.close()


def test_round_count_hysteresis(gpu):
    """A block keeps the round count of its previous layout while that is at
    most one above the fewest it needs (a different count re-deals every
    particle of the chunk, scattering the next transfer's reads): one 8x4x4
    block of base cells with 8 particles per cell plus one (1025: 17 rounds),
    then the 1024 of a full block (16 needed, 17 kept), then 7 per cell (14
    needed, 17 dropped).  A fresh engine lays out the fewest."""
    cfg = gpu.make_config(SIZE_X=32, SIZE_Y=32, SIZE_Z=32)
    rng = np.random.default_rng(3)

    def block_state(per_cell, extra):
        cells = np.stack(np.meshgrid(np.arange(8, 16), np.arange(8, 12), np.arange(8, 12), indexing="ij"),
                         0).reshape(3, -1)
        pos = np.repeat(cells, per_cell, axis=1).astype(np.float64)
        if extra:
            pos = np.concatenate([pos, np.array([[8.0], [8.0], [8.0]])], axis=1)
        pos = pos + 1.0 + rng.uniform(-0.4, 0.4, size=pos.shape)  # base cell trunc(x - 0.5) = the cell
        n = pos.shape[1]
        st = np.zeros((26, n), dtype=np.float32)
        st[0:3] = pos
        st[[15, 19, 23]] = 1.0
        st[24] = st[25] = 1.0
        return st

    e = gpu.Engine(cfg, CHUNK=2048)
    e.load_state(block_state(8, True))
    r1025 = e.counters()["rounds"]
    e.load_state(block_state(8, False))
    r1024 = e.counters()["rounds"]
    e.load_state(block_state(7, False))
    r896 = e.counters()["rounds"]
    e.steps(2)
    assert np.isfinite(e.getParticles()).all()
    e.close()
    fresh = gpu.Engine(cfg, CHUNK=2048)
    fresh.load_state(block_state(8, False))
    r_fresh = fresh.counters()["rounds"]
    fresh.close()
    say(f"rounds: 1025 -> {r1025}, then 1024 -> {r1024}, then 896 -> {r896}; fresh 1024 -> {r_fresh}")
    assert (r1025, r1024, r896, r_fresh) == (17, 17, 14, 16)
