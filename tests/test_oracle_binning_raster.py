"""Pins the oracle's binning and rasteriser:
  * bins: brute-force (tile, gid) enumeration from each Gaussian's tile
    rectangle, permutation / sortedness / bin-size invariants, stable sort
    against numpy's stable argsort (hypothesis);
  * forward: single-Gaussian closed form and a tile-free brute-force sum over
    every Gaussian for every pixel (numpy FP64), plus the GA forward's
    dx/dy/dxy against finite differences of the image in pixel position;
  * backward: FP64 central finite differences of the forward with the set of
    contributing pairs held fixed (plain and gradient-aware).
"""
import numpy as np
import pytest
from hypothesis import given, settings, strategies as st

import oracle
from tests.scenes import random_scene


def _scene_proj(s, bw):
    xys, ext, con, nt = oracle.project_gaussians_forward(s["par"], s["means"], s["H"], s["W"],
                                                         bw, s["opac"], cholesky=s["cholesky"])
    return xys, ext, con, nt


@pytest.mark.parametrize("bw", [16, 8, 5])
def test_bins_brute_force(bw):
    s = random_scene(150, 50, 70, 1, opacities=True, margin=10.0)
    s["opac"][::9] = 0.002      # phantom slots
    xys, ext, con, nt = _scene_proj(s, bw)
    m, cum, isect, gids, ks, gs, bins = oracle.bin_and_sort(xys, ext, nt, s["H"], s["W"], bw)
    tx, ty = -(-s["W"] // bw), -(-s["H"] // bw)
    rec = []
    for g in range(150):
        if nt[g] == 0:
            continue
        if ext[g, 0] <= 0 or ext[g, 1] <= 0:
            rec += [(0, 0)] * int(nt[g])           # zero-filled slot, sorts as (tile 0, gid 0)
            continue
        x0 = min(max(0, int(np.float32(xys[g, 0] / bw) - np.float32(ext[g, 0] / bw))), tx)
        x1 = min(max(0, int(np.float32(xys[g, 0] / bw) + np.float32(ext[g, 0] / bw) + 1)), tx)
        y0 = min(max(0, int(np.float32(xys[g, 1] / bw) - np.float32(ext[g, 1] / bw))), ty)
        y1 = min(max(0, int(np.float32(xys[g, 1] / bw) + np.float32(ext[g, 1] / bw) + 1)), ty)
        rec += [(y * tx + x, g) for y in range(y0, y1) for x in range(x0, x1)]
    assert len(rec) == m
    # stable sort of the emission order by tile
    order = sorted(range(m), key=lambda i: (rec[i][0], i))
    np.testing.assert_array_equal(gs, [rec[i][1] for i in order])
    tiles = np.array([rec[i][0] for i in order])
    for t in range(tx * ty):
        sel = np.nonzero(tiles == t)[0]
        if len(sel):
            assert bins[t, 0] == sel[0] and bins[t, 1] == sel[-1] + 1
        else:
            assert bins[t, 0] == 0 and bins[t, 1] == 0
    assert (bins[:, 1] - bins[:, 0]).sum() == m


def test_bins_cover_every_contributing_pixel():
    """Every pixel where a Gaussian reaches alpha >= 1/255 sits in a tile that
    lists it (SnugBox + tile bbox never drop a contributor)."""
    s = random_scene(120, 64, 64, 2, opacities=True, scale=(0.5, 6.0), margin=8.0)
    bw = 16
    xys, ext, con, nt = _scene_proj(s, bw)
    m, cum, isect, gids, ks, gs, bins = oracle.bin_and_sort(xys, ext, nt, 64, 64, bw)
    ys, xs = np.mgrid[0:64, 0:64] + 0.5
    for g in range(120):
        if nt[g] == 0 or ext[g, 0] <= 0:
            continue
        dx, dy = xys[g, 0] - xs, xys[g, 1] - ys
        a, b, c = con[g].astype(np.float64)
        sig = 0.5 * (a * dx * dx + c * dy * dy) + b * dx * dy
        hit = (s["opac"][g] * np.exp(-sig) >= 1 / 255) & (sig >= 0)
        for i, j in zip(*np.nonzero(hit)):
            t = (i // bw) * 4 + j // bw
            assert g in gs[bins[t, 0]:bins[t, 1]], (g, i, j)


@settings(max_examples=50, deadline=None)
@given(st.lists(st.tuples(st.integers(-2**40, 2**40), st.integers(0, 2**31 - 1)),
                min_size=0, max_size=300))
def test_sort_pairs_is_stable_sort(pairs):
    keys = np.array([p[0] for p in pairs], np.int64)
    vals = np.array([p[1] for p in pairs], np.int32)
    ks, vs = oracle.sort_pairs(keys, vals)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(ks, keys[order])
    np.testing.assert_array_equal(vs, vals[order])


# --------------------------------------------------------------------------
# forward
# --------------------------------------------------------------------------

def _brute_forward(means, conics, colors, opac, H, W, clamp=0.999, ga=False):
    """Tile-free FP64 sum over all Gaussians for every pixel (d = mu - p)."""
    ys, xs = np.mgrid[0:H, 0:W] + 0.5
    img = np.zeros((H, W, 3))
    wsum = np.zeros((H, W))
    dxo = np.zeros((H, W, 3))
    dyo = np.zeros((H, W, 3))
    dxyo = np.zeros((H, W, 3))
    for g in range(means.shape[0]):
        dx = means[g, 0] - xs
        dy = means[g, 1] - ys
        a, b, c = conics[g]
        sig = 0.5 * (a * dx * dx + c * dy * dy) + b * dx * dy
        o = 1.0 if opac is None else opac[g]
        al = np.minimum(o * np.exp(-sig), clamp)
        ok = ~((sig < 0) | (al < 1 / 255))
        al = np.where(ok, al, 0.0)
        img += al[..., None] * colors[g]
        wsum += al
        if ga:
            u = a * dx + b * dy
            v = b * dx + c * dy
            dxo += (al * u)[..., None] * colors[g]
            dyo += (al * v)[..., None] * colors[g]
            dxyo += (al * (u * v - b))[..., None] * colors[g]
    return img, wsum, dxo, dyo, dxyo


def test_single_gaussian_closed_form():
    H = W = 32
    cov = np.array([[9.0, 2.0, 4.0]], np.float32)
    means = np.array([[15.3, 16.7]], np.float32)
    col = np.array([[0.2, 0.5, 0.9]], np.float32)
    xys, ext, con, nt = oracle.project_gaussians_forward(cov, means, H, W, 16)
    m, cum, isect, gids, ks, gs, bins = oracle.bin_and_sort(xys, ext, nt, H, W, 16)
    out = oracle.rasterize_forward(gs, bins, xys, con, col, None, H, W, 16)
    Si = np.linalg.inv(np.array([[9.0, 2.0], [2.0, 4.0]]))
    ys, xs = np.mgrid[0:H, 0:W] + 0.5
    d = np.stack([xs - 15.3, ys - 16.7], -1)
    alpha = np.exp(-0.5 * np.einsum("hwi,ij,hwj->hw", d, Si, d))
    alpha = np.where(alpha >= 1 / 255, np.minimum(alpha, 0.999), 0)
    np.testing.assert_allclose(out["img"], alpha[..., None] * col[0], atol=2e-6)
    np.testing.assert_allclose(out["wsum"], alpha, atol=2e-6)


@pytest.mark.parametrize("bw,seed", [(16, 0), (8, 1), (4, 2), (13, 3)])
def test_forward_equals_tile_free_sum(bw, seed):
    s = random_scene(80, 40, 56, seed, opacities=bool(seed % 2), margin=6.0)
    xys, ext, con, nt = _scene_proj(s, bw)
    m, cum, isect, gids, ks, gs, bins = oracle.bin_and_sort(xys, ext, nt, 40, 56, bw)
    out = oracle.rasterize_forward(gs, bins, xys, con, s["colors"], s["opac"], 40, 56, bw,
                                   gradient_aware=True)
    on = nt > 0
    ref = _brute_forward(xys[on].astype(np.float64), con[on].astype(np.float64),
                         s["colors"][on].astype(np.float64),
                         None if s["opac"] is None else s["opac"][on].astype(np.float64),
                         40, 56, ga=True)
    clean = out["ambig"] == 0
    for k, r in zip(("img", "wsum", "dx", "dy", "dxy"), ref):
        np.testing.assert_allclose(out[k][clean], r[clean], atol=1e-9, err_msg=k)


def test_ga_forward_is_spatial_derivative():
    """dx, dy, dxy of the gradient-aware forward are the derivatives of the
    image w.r.t. pixel position: I(p + h e) = image with every mean moved by -h."""
    s = random_scene(30, 24, 24, 4, opacities=True, scale=(2.0, 4.0))
    xys, ext, con, nt = _scene_proj(s, 16)
    on = nt > 0
    mu = xys[on].astype(np.float64)
    args = (con[on].astype(np.float64), s["colors"][on].astype(np.float64),
            s["opac"][on].astype(np.float64), 24, 24)
    img, w, dx, dy, dxy = _brute_forward(mu, *args, ga=True)
    h = 1e-4
    sh = lambda ox, oy: _brute_forward(mu - np.array([ox, oy]), *args)[0]  # noqa: E731
    fdx = (sh(h, 0) - sh(-h, 0)) / (2 * h)
    fdy = (sh(0, h) - sh(0, -h)) / (2 * h)
    fdxy = (sh(h, h) - sh(h, -h) - sh(-h, h) + sh(-h, -h)) / (4 * h * h)
    # compare away from the 1/255 cut, where the image is smooth
    smooth = np.ones((24, 24), bool)
    ys, xs = np.mgrid[0:24, 0:24] + 0.5
    for g in range(mu.shape[0]):
        ddx, ddy = mu[g, 0] - xs, mu[g, 1] - ys
        a, b, c = args[0][g]
        sg = 0.5 * (a * ddx * ddx + c * ddy * ddy) + b * ddx * ddy
        al = args[2][g] * np.exp(-sg)
        smooth &= np.abs(al * 255 - 1) > 0.05
    np.testing.assert_allclose(dx[smooth], fdx[smooth], atol=1e-6)
    np.testing.assert_allclose(dy[smooth], fdy[smooth], atol=1e-6)
    np.testing.assert_allclose(dxy[smooth], fdxy[smooth], atol=1e-4)


# --------------------------------------------------------------------------
# backward: FD with the pair set fixed
# --------------------------------------------------------------------------

def _loss(means, conics, colors, opac, H, W, w_img, w_ws, ga_w=None):
    img, ws, dx, dy, dxy = _brute_forward(means, conics, colors, opac, H, W,
                                          ga=ga_w is not None)
    L = (img * w_img).sum() + (ws * w_ws).sum()
    if ga_w is not None:
        L += (dx * ga_w[0]).sum() + (dy * ga_w[1]).sum() + (dxy * ga_w[2]).sum()
    return L


def _pairs(means, conics, opac, H, W):
    ys, xs = np.mgrid[0:H, 0:W] + 0.5
    out = []
    for g in range(means.shape[0]):
        dx, dy = means[g, 0] - xs, means[g, 1] - ys
        a, b, c = conics[g]
        sg = 0.5 * (a * dx * dx + c * dy * dy) + b * dx * dy
        out.append(~((sg < 0) | (opac[g] * np.exp(-sg) < 1 / 255)))
    return np.stack(out)


@pytest.mark.parametrize("ga", [False, True])
def test_backward_matches_finite_differences(ga):
    H, W = 24, 28
    s = random_scene(12, H, W, 8 + ga, opacities=True, scale=(1.5, 4.0))
    s["opac"] = np.clip(s["opac"], 0.1, 0.9).astype(np.float32)   # never clamped
    xys, ext, con, nt = _scene_proj(s, 16)
    assert np.all(nt > 0)
    m, cum, isect, gids, ks, gs, bins = oracle.bin_and_sort(xys, ext, nt, H, W, 16)
    fwd = oracle.rasterize_forward(gs, bins, xys, con, s["colors"], s["opac"], H, W, 16)
    rng = np.random.default_rng(3)
    w_img = rng.normal(size=(H, W, 3)).astype(np.float32)
    w_ws = rng.normal(size=(H, W)).astype(np.float32)
    ga_w = [rng.normal(size=(H, W, 3)).astype(np.float32) for _ in range(3)] if ga else None
    got = oracle.rasterize_backward(gs, bins, xys, con, s["colors"], s["opac"],
                                    fwd["final_idx"], w_img, w_ws, H, W, 16,
                                    *(ga_w or [None] * 3))
    mu = xys.astype(np.float64)
    cn = con.astype(np.float64)
    col = s["colors"].astype(np.float64)
    op = s["opac"].astype(np.float64)
    base_pairs = _pairs(mu, cn, op, H, W)
    h = 1e-6
    checked = 0
    for g in range(12):
        for name, arr, k, slot in (("v_xy", mu, 0, 0), ("v_xy", mu, 1, 1),
                                   ("v_conic", cn, 0, 0), ("v_conic", cn, 1, 1),
                                   ("v_conic", cn, 2, 2), ("v_rgb", col, 0, 0),
                                   ("v_rgb", col, 2, 2), ("v_opacity", op, None, None)):
            if name == "v_opacity" and ga:
                continue    # the reference's GA backward ignores d(dx, dy, dxy)/d opacity
            plus, minus = arr.copy(), arr.copy()
            if k is None:
                plus[g] += h
                minus[g] -= h
            else:
                plus[g, k] += h
                minus[g, k] -= h
            vals = []
            for a in (plus, minus):
                args = {"v_xy": (a, cn, col, op), "v_conic": (mu, a, col, op),
                        "v_rgb": (mu, cn, a, op), "v_opacity": (mu, cn, col, a)}[name]
                if not np.array_equal(_pairs(args[0], args[1], args[3], H, W), base_pairs):
                    break
                vals.append(_loss(*args, H, W, w_img, w_ws, ga_w))
            if len(vals) < 2:
                continue
            fd = (vals[0] - vals[1]) / (2 * h)
            an = got[name][g] if slot is None else got[name][g, slot]
            assert abs(an - fd) <= 1e-5 * max(1.0, abs(fd)), (g, name, k, an, fd)
            checked += 1
    assert checked > 60


def _pixel_vxy_fd(mu, cn, col, op, H, W, w_img, w_ws, ga_w, h=1e-6):
    """Per-pixel, per-Gaussian position gradient [n, H, W, 2] by FP64 central
    differences of each Gaussian's own tile-free forward (the forward is a sum
    over Gaussians, so pixel p's gradient w.r.t. mean g involves g alone)."""
    n = mu.shape[0]
    out = np.zeros((n, H, W, 2))
    ga = ga_w is not None

    def per_pixel(m, g):
        img, ws, dx, dy, dxy = _brute_forward(m, cn[g:g + 1], col[g:g + 1], op[g:g + 1], H, W,
                                              ga=ga)
        v = (img * w_img).sum(-1) + ws * w_ws
        if ga:
            v = v + (dx * ga_w[0]).sum(-1) + (dy * ga_w[1]).sum(-1) + (dxy * ga_w[2]).sum(-1)
        return v

    for g in range(n):
        base = _pairs(mu[g:g + 1], cn[g:g + 1], op[g:g + 1], H, W)
        for k in range(2):
            plus, minus = mu[g:g + 1].copy(), mu[g:g + 1].copy()
            plus[0, k] += h
            minus[0, k] -= h
            for m in (plus, minus):
                assert np.array_equal(_pairs(m, cn[g:g + 1], op[g:g + 1], H, W), base)
            out[g, :, :, k] = (per_pixel(plus, g) - per_pixel(minus, g)) / (2 * h)
    return out


def _group_abs(vxy_px, H, W, bw):
    """sum over tiles and 32-thread groups (thread rank ty*bw + tx) of
    |sum of the group's per-pixel gradients|; pixels outside the image add 0."""
    n = vxy_px.shape[0]
    tx, ty = (W + bw - 1) // bw, (H + bw - 1) // bw
    out = np.zeros((n, 2))
    ranks = np.arange(bw * bw)
    for t in range(tx * ty):
        i0, j0 = (t // tx) * bw, (t % tx) * bw
        for r0 in range(0, bw * bw, 32):
            r = ranks[r0:r0 + 32]
            i, j = i0 + r // bw, j0 + r % bw
            inside = (i < H) & (j < W)
            out += np.abs(vxy_px[:, i[inside], j[inside], :].sum(1))
    return out


@pytest.mark.parametrize("bw", [16, 8, 12, 5])
def test_absgrad_group_semantics(bw):
    """xys.absgrad pinned independently of the oracle's pair math: per-pixel
    position gradients by FD, then the plain kernel's sum of per-pixel |v_xy|
    (backward.cu:144-150) and the gradient-aware kernel's sum over 32-thread
    groups of |group sum| (gradient_aware_rasterizer.cu:320,339-341).  The scene
    is chosen so that the two differ by far more than the tolerance."""
    H, W = 26, 30
    s = random_scene(10, H, W, 30 + bw, opacities=True, scale=(1.5, 4.0))
    s["opac"] = np.clip(s["opac"], 0.1, 0.9).astype(np.float32)   # never clamped
    xys, ext, con, nt = _scene_proj(s, bw)
    assert np.all(nt > 0)
    m, cum, isect, gids, ks, gs, bins = oracle.bin_and_sort(xys, ext, nt, H, W, bw)
    fwd = oracle.rasterize_forward(gs, bins, xys, con, s["colors"], s["opac"], H, W, bw)
    rng = np.random.default_rng(bw)
    w_img = rng.normal(size=(H, W, 3)).astype(np.float32)
    w_ws = rng.normal(size=(H, W)).astype(np.float32)
    ga_w = [rng.normal(size=(H, W, 3)).astype(np.float32) for _ in range(3)]
    mu, cn = xys.astype(np.float64), con.astype(np.float64)
    col, op = s["colors"].astype(np.float64), s["opac"].astype(np.float64)
    for ga in (False, True):
        gw = ga_w if ga else None
        got = oracle.rasterize_backward(gs, bins, xys, con, s["colors"], s["opac"],
                                        fwd["final_idx"], w_img, w_ws, H, W, bw,
                                        *(gw or [None] * 3))
        px = _pixel_vxy_fd(mu, cn, col, op, H, W, w_img.astype(np.float64),
                           w_ws.astype(np.float64),
                           None if gw is None else [a.astype(np.float64) for a in gw])
        per_pixel = np.abs(px).sum((1, 2))
        grouped = _group_abs(px, H, W, bw)
        want = grouped if ga else per_pixel
        np.testing.assert_allclose(got["v_xy_abs"], want, rtol=1e-5, atol=1e-6,
                                   err_msg=f"ga={ga} bw={bw}")
        np.testing.assert_allclose(got["v_xy"], px.sum((1, 2)), rtol=1e-5, atol=1e-6)
        if ga:
            # the two semantics must be far apart for this pin to mean anything
            assert grouped.sum() < 0.9 * per_pixel.sum(), (grouped.sum(), per_pixel.sum())


@pytest.mark.parametrize("bw,op", [(16, True), (8, False), (5, True)])
def test_final_idx_is_last_contributing_bin_position(bw, op):
    """final_idx (forward.cu:217,269,274): for each pixel, the largest position
    k of its tile's bin whose Gaussian contributes (tile-free predicate), else 0."""
    H, W = 30, 34
    s = random_scene(60, H, W, 40 + bw, opacities=op, margin=5.0)
    xys, ext, con, nt = _scene_proj(s, bw)
    m, cum, isect, gids, ks, gs, bins = oracle.bin_and_sort(xys, ext, nt, H, W, bw)
    fwd = oracle.rasterize_forward(gs, bins, xys, con, s["colors"], s["opac"], H, W, bw)
    opv = np.ones(60) if s["opac"] is None else s["opac"].astype(np.float64)
    pairs = _pairs(xys.astype(np.float64), con.astype(np.float64), opv, H, W)
    tx = (W + bw - 1) // bw
    want = np.zeros((H, W), np.int32)
    for i in range(H):
        for j in range(W):
            t = (i // bw) * tx + j // bw
            for k in range(bins[t, 0], bins[t, 1]):
                if pairs[gs[k], i, j]:
                    want[i, j] = k
    clean = fwd["ambig"] == 0
    np.testing.assert_array_equal(fwd["final_idx"][clean], want[clean])
    assert clean.mean() > 0.95 and want.max() > 0
