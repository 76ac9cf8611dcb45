"""The no-allocation binning entry point (lig_bin_and_sort_gaussians_noalloc)
against the workspace entry point (lig_bin_and_sort_gaussians) and the oracle,
through the C-ABI: two calls in a row on one device, the second smaller, so the
second reuses the scratch the first left behind."""
import numpy as np
import pytest
import torch

from tests.scenes import random_scene

pytestmark = pytest.mark.gpu


def _t(a, dev, dtype=torch.float32):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)


def _np(t):
    return t.detach().cpu().numpy()


# n, H, W, block width, seed: a 2048^2 level of 300 k Gaussians (about 0.95 M
# intersections), then the 512^2 x 20 k scene, then an odd-sized 8-px-tile case
CALLS = [(300_000, 2048, 2048, 16, 21), (20_000, 512, 512, 16, 22), (3_000, 300, 217, 8, 23)]


def _scene(oracle_lib, n, H, W, bw, seed, dev):
    s = random_scene(n, H, W, seed, cholesky=True, margin=20.0)
    xys, ext, con, nt = oracle_lib.project_gaussians_forward(s["par"], s["means"], H, W, bw,
                                                             None, cholesky=True)
    m, cum, isect, gids, ks, gs, bins = oracle_lib.bin_and_sort(xys, ext, nt, H, W, bw)
    return dict(m=m, gs=gs, bins=bins, xys=_t(xys, dev), ext=_t(ext, dev),
                cum=_t(cum, dev, torch.int32))


def test_noalloc_binning_matches_workspace_entry_point(device, oracle_lib):
    from lig_amd.gsplat2d import cuda as C
    for n, H, W, bw, seed in CALLS:
        for with_eids in (False, True):
            print(f"noalloc binning n={n} H={H} W={W} bw={bw} seed={seed} eids={with_eids}",
                  flush=True)
            sc = _scene(oracle_lib, n, H, W, bw, seed, device)
            m = sc["m"]
            nt = ((W + bw - 1) // bw) * ((H + bw - 1) // bw)
            gids = torch.full((m,), -7, dtype=torch.int32, device=device)
            bins = torch.full((nt, 2), -7, dtype=torch.int32, device=device)
            eids = torch.full((m,), -7, dtype=torch.int32, device=device) if with_eids else None
            C.bin_and_sort_noalloc(n, m, sc["xys"], sc["ext"], sc["cum"], H, W, bw, gids, bins,
                                   eids)
            ref = C.bin_and_sort_fast(n, m, sc["xys"], sc["ext"], sc["cum"], H, W, bw,
                                      with_emission_ids=with_eids)
            torch.cuda.synchronize()
            assert torch.equal(gids, ref[0])
            assert torch.equal(bins, ref[1])
            if with_eids:
                assert torch.equal(eids, ref[2])
            np.testing.assert_array_equal(_np(gids), sc["gs"])
            np.testing.assert_array_equal(_np(bins), sc["bins"])


def test_noalloc_binning_rejects_sizes_beyond_scratch(device):
    """A size whose workspace exceeds the fixed scratch is refused on the host
    (LIG_ERR_TOO_LARGE) before anything is launched; the outputs are untouched."""
    import lig_amd._lib as L
    from lig_amd.gsplat2d import cuda as C
    m = 3_000_000
    assert L.size_query("lig_bin_workspace_size", m, 16384) > C.BIN_NOALLOC_SCRATCH_BYTES
    xys = torch.zeros((4, 2), device=device)
    ext = torch.zeros((4, 2), device=device)
    cum = torch.zeros((4,), dtype=torch.int32, device=device)
    gids = torch.zeros((4,), dtype=torch.int32, device=device)
    bins = torch.full((16384, 2), 5, dtype=torch.int32, device=device)
    with pytest.raises(L.LigError, match="fixed scratch size"):
        C.bin_and_sort_noalloc(4, m, xys, ext, cum, 2048, 2048, 16, gids, bins)
    torch.cuda.synchronize()
    assert int(bins.min()) == 5 and int(bins.max()) == 5
