"""GPU tests of batched fitting (lig_amd.batch): several images of uneven size
fitted in one launch per stage give every image the parameters, gradients and
loss its plain per-image fit (lig_amd.fit.Level, fused path) gives it."""
import pytest
import torch

pytestmark = pytest.mark.gpu

SIZES = [(100, 130), (48, 40), (256, 200), (17, 300), (64, 64), (33, 47)]
COUNTS = [600, 150, 2500, 300, 400, 90]
SEEDS = [11, 12, 13, 14, 15, 16]


def test_batched_matches_plain(device):
    from lig_amd.batch import BatchedFit, make_targets
    from lig_amd.fit import Level
    targets = make_targets(SIZES, SEEDS, device)
    bf = BatchedFit(targets, COUNTS, SEEDS, scale=1.5, lr=1e-2)
    levels = [Level(c, h, w, s, 1.5, lr=1e-2, device=device)
              for (h, w), c, s in zip(SIZES, COUNTS, SEEDS)]
    for step in range(3):
        bf.step()
        plain = [float(lv.step(t)) for lv, t in zip(levels, targets)]
        batched = bf.last_losses.tolist()
        for i, (a, b) in enumerate(zip(batched, plain)):
            assert a == pytest.approx(b, rel=1e-6), (step, i)
    o = 0
    for i, (lv, c) in enumerate(zip(levels, COUNTS)):
        for name in ("means", "cholesky", "rgbs"):
            assert torch.equal(getattr(bf, name)[o:o + c], getattr(lv, name)), (i, name)
        o += c
    psnr = bf.psnr()
    assert psnr.shape == (len(SIZES),) and (psnr > 0).all()


def _state(bf):
    st = bf.optimizer.state
    return [t.detach().clone() for p in (bf.rgbs, bf.means, bf.cholesky)
            for t in (p, st[p]["exp_avg"], st[p]["exp_avg_sq"])]


def _assert_same(a, b, where):
    for k, (x, y) in enumerate(zip(_state(a), _state(b))):
        assert torch.equal(x, y), (where, k)
    for p, q in zip((a.rgbs, a.means, a.cholesky), (b.rgbs, b.means, b.cholesky)):
        assert a.optimizer.state[p]["step"] == b.optimizer.state[q]["step"], where


@pytest.mark.parametrize("capacity", [None, 64])
def test_graph_step_matches_eager(device, capacity):
    """The captured step replayed k times gives the bits of k eager steps,
    also when the first capacity is far too small (the skipped steps change
    nothing, the capacity grows and the graph is recorded again).  Replays are
    checked one by one, with eager work and synchronisation between them."""
    from lig_amd.batch import BatchedFit, BatchedStepGraph, make_targets
    targets = make_targets(SIZES, SEEDS, device)
    eager = BatchedFit(targets, COUNTS, SEEDS, scale=1.5, lr=1e-2)
    graphed = BatchedFit(targets, COUNTS, SEEDS, scale=1.5, lr=1e-2)
    gs = BatchedStepGraph(graphed, capacity=capacity)
    print(f"capacity {capacity} -> {gs.capacity}", flush=True)
    for step in range(3):
        le = eager.step()
        lg = gs.run(1)
        torch.cuda.synchronize()
        assert torch.equal(le.reshape(()), lg), step
        assert torch.equal(eager.last_losses, graphed.last_losses), step
        _assert_same(eager, graphed, step)
    if capacity is not None:
        assert gs.captures >= 2 and gs.skipped >= 1 and gs.capacity > capacity
    # several replays per run, then an eager step on the graphed model, then more replays
    for _ in range(3):
        eager.step()
    gs.run(3)
    _assert_same(eager, graphed, "run(3)")
    eager.step()
    graphed.step()
    _assert_same(eager, graphed, "interleaved eager")
    eager.step()
    eager.step()
    gs.run(2)
    _assert_same(eager, graphed, "after interleave")


def test_graph_step_bias_table_extension(device):
    """A bias table shorter than the run: the step past its end is skipped,
    the table is extended and the result still equals the eager steps."""
    from lig_amd.batch import BatchedFit, BatchedStepGraph, make_targets
    targets = make_targets(SIZES[:3], SEEDS[:3], device)
    eager = BatchedFit(targets, COUNTS[:3], SEEDS[:3], scale=1.5, lr=1e-2)
    graphed = BatchedFit(targets, COUNTS[:3], SEEDS[:3], scale=1.5, lr=1e-2)
    gs = BatchedStepGraph(graphed, table_steps=2)
    for _ in range(5):
        eager.step()
    gs.run(5)
    _assert_same(eager, graphed, "table")
    assert gs.table_len >= 5 and gs.captures >= 2


@pytest.mark.parametrize("skewed", [False, True])
def test_persistent_queue_matches_grid(device, skewed):
    """The persistent raster stages (per-XCD tile queue with stealing) give
    the bits of the one-tile-per-workgroup launches, eagerly and captured; the
    skewed batch puts 20x the Gaussians per pixel into two images so that the
    XCD shares are uneven."""
    from lig_amd.batch import BatchedFit, BatchedStepGraph, make_targets
    counts = list(COUNTS)
    if skewed:
        counts[0] *= 20
        counts[2] *= 20
    targets = make_targets(SIZES, SEEDS, device)
    grid = BatchedFit(targets, counts, SEEDS, scale=1.5, lr=1e-2)
    queue = BatchedFit(targets, counts, SEEDS, scale=1.5, lr=1e-2)
    queue.persistent = True
    for step in range(3):
        a = grid.step()
        b = queue.step()
        assert torch.equal(a, b), step
        assert torch.equal(grid.last_losses, queue.last_losses), step
    _assert_same(grid, queue, "eager")
    gs = BatchedStepGraph(queue)
    for _ in range(2):
        grid.step()
    gs.run(2)
    _assert_same(grid, queue, "graph")


def test_autotune_batched_picks_form_and_depth(device, tmp_path, monkeypatch):
    """tune_batched times every (launch form, depth) pair, sets the fastest
    on the model and caches it; a second call reads the cache."""
    from lig_amd import autotune
    from lig_amd.batch import BatchedFit, make_targets
    monkeypatch.setenv("LIG_AUTOTUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_cache", {})
    targets = make_targets(SIZES, SEEDS, device)
    bf = BatchedFit(targets, COUNTS, SEEDS, scale=1.5, lr=1e-2)
    res = autotune.tune_batched(bf, reps=1)
    assert set(res["ms"]) == {"grid", "persistent"}
    assert all(set(row) == set(autotune.DEPTHS) for row in res["ms"].values())
    best = min(t for row in res["ms"].values() for t in row.values())
    form = "persistent" if bf.persistent else "grid"
    assert res["ms"][form][bf.batch_depth] == best
    again = autotune.tune_batched(bf, reps=1)
    assert again["cached"] and again["choice"] == res["choice"]
    assert again["persistent"] == res["persistent"]
    bf.step()   # the tuned configuration runs


def test_batched_images_without_intersections(device):
    """An image with no Gaussians and one whose Gaussians all lie off its
    canvas render ones, as each does fitted alone (rasterize.py:80-90), and
    the other images are unaffected: losses and parameters match the plain
    per-image fits."""
    from lig_amd.batch import BatchedFit, make_targets
    from lig_amd.fit import Level
    sizes = [(40, 56), (33, 47), (64, 64), (20, 31)]
    counts = [200, 0, 150, 80]
    seeds = [21, 22, 23, 24]
    targets = make_targets(sizes, seeds, device)
    bf = BatchedFit(targets, counts, seeds, scale=1.5, lr=1e-2)
    levels = [Level(c, h, w, s, 1.5, lr=1e-2, device=device)
              for (h, w), c, s in zip(sizes, counts, seeds)]
    with torch.no_grad():
        bf.means[200:350] += 1.0e4
        levels[2].means += 1.0e4
    for step in range(2):
        bf.step()
        plain = [float(lv.step(t)) for lv, t in zip(levels, targets)]
        for i, (a, b) in enumerate(zip(bf.last_losses.tolist(), plain)):
            assert a == pytest.approx(b, rel=1e-6), (step, i)
    for i in (1, 2):
        want = float(((1.0 - targets[i]) ** 2).mean())
        assert bf.last_losses[i].item() == pytest.approx(want, rel=1e-6)
    o = 0
    for i, (lv, c) in enumerate(zip(levels, counts)):
        for name in ("means", "cholesky", "rgbs"):
            assert torch.equal(getattr(bf, name)[o:o + c], getattr(lv, name)), (i, name)
        o += c


@pytest.mark.parametrize("persistent", [False, True])
def test_batched_dense_tiles_match_api_chain(device, persistent):
    """Dense tiles (over a thousand Gaussians per tile, so the fit backward
    groups several 512-Gaussian chunks by their rows, with partial edge rows
    on the 100 x 130 image) in the batched grid and persistent-queue launches:
    every image's loss and its parameters after two Adam steps equal those of
    its plain fit through the reference API chain (Level(fused=False))."""
    from lig_amd.batch import BatchedFit, make_targets
    from lig_amd.fit import Level
    sizes = [(100, 130), (96, 96), (40, 56)]
    counts = [20000, 12000, 2000]
    seeds = [31, 32, 33]
    print(f"batched dense persistent={persistent}", flush=True)
    targets = make_targets(sizes, seeds, device)
    bf = BatchedFit(targets, counts, seeds, scale=3.0, lr=1e-2)
    bf.persistent = persistent
    levels = [Level(c, h, w, s, 3.0, lr=1e-2, device=device, fused=False)
              for (h, w), c, s in zip(sizes, counts, seeds)]
    with torch.no_grad():
        bf.rgbs.mul_(0.01)   # keep part of each image inside (0, 1)
        for lv in levels:
            lv.rgbs.mul_(0.01)
        nt = levels[0].project()[3]
    assert int(nt.sum()) > 63 * 2 * 512
    for step in range(2):
        bf.step()
        plain = [float(lv.step(t)) for lv, t in zip(levels, targets)]
        for i, (a, b) in enumerate(zip(bf.last_losses.tolist(), plain)):
            assert a == pytest.approx(b, rel=1e-6), (step, i)
    o = 0
    for i, (lv, c) in enumerate(zip(levels, counts)):
        for name in ("means", "cholesky", "rgbs"):
            assert torch.equal(getattr(bf, name)[o:o + c], getattr(lv, name)), (i, name)
        o += c
