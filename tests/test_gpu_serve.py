"""GPU tests of the crop server (lig_amd.serve): continuous batching over
several HIP streams gives every request the same bits as rendering it alone,
and HBM admission holds the in-flight reservations to the budget."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _setup(device):
    from lig_amd.serve import make_representations, make_requests
    reps = make_representations(3, seed=7, device=device, sizes=(512, 768),
                                n_range=(20_000, 60_000), scale_px=2.0)
    reqs = make_requests(reps, 24, seed=11, crop_range=(32, 384))
    return reps, reqs


def test_results_same_alone_and_in_the_mix(device):
    from lig_amd.serve import CropServer, Request
    reps, reqs = _setup(device)
    alone = [reps[r.rep].render_crop(r.x0, r.y0, r.w, r.h, r.scale) for r in reqs]
    torch.cuda.synchronize()
    srv = CropServer(reps, slots=4, keep_results=True, device=device)
    mixed = [Request(r.rep, r.x0, r.y0, r.w, r.h, r.scale) for r in reqs]
    st = srv.run_poisson(mixed, qps=2000.0)
    assert st.requests == len(reqs)
    for r, a in zip(mixed, alone):
        assert r.result is not None and torch.equal(r.result, a), (r.rep, r.x0, r.y0, r.w, r.h)
    s = st.summary()
    print("serve summary", s)
    assert s["p99_ms"] >= s["p50_ms"] > 0


def test_admission_by_budget(device):
    from lig_amd.serve import CropServer, Request
    reps, reqs = _setup(device)
    # a budget smaller than any request admits one request at a time
    srv = CropServer(reps, slots=4, hbm_budget=1, device=device)
    mixed = [Request(r.rep, r.x0, r.y0, r.w, r.h, r.scale) for r in reqs]
    st = srv.run_poisson(mixed, qps=2000.0)
    assert st.peak_reserved == max(r.reserved for r in mixed)
    # an ample budget lets requests overlap
    srv = CropServer(reps, slots=4, hbm_budget=64 << 30, device=device)
    mixed = [Request(r.rep, r.x0, r.y0, r.w, r.h, r.scale) for r in reqs]
    st = srv.run_poisson(mixed, qps=2000.0)
    assert st.peak_reserved >= max(r.reserved for r in mixed)
