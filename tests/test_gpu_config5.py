"""The batched workload at its full size: eight seeded scans of uneven size
(cubic phantoms of 128-512 voxels with 1-3 materials, detectors 256x256 to
1024x768, 60-360 views, 80-140 kVp, grid on or off) in one persistent,
graph-captured batch with FP16 volumes.  Items carry different launch
options (LDS budgets, staging chunks, ray-tile shapes), so the batch holds
several persistent launches.  Every item's outputs equal its own forward bit
for bit, and one view of every item is checked against the FP64 oracle
(1e-3, the FP16 bar).  The projection each item runs is the reference's
(Source/CBCTPolyForwardProjGrid.cpp:108-203)."""
import numpy as np
import pytest

import ct_forward_projection_amd as P
from ct_forward_projection_amd import synthetic as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TUNINGS = [dict(lds_bytes=40 * 1024, chunk=16, wave_u=16, block_u=1),
           dict(lds_bytes=64 * 1024, chunk=16, wave_u=8, block_u=2),
           dict(lds_bytes=40 * 1024, chunk=32, wave_u=16, block_u=2),
           dict(lds_bytes=24 * 1024, chunk=16, wave_u=32, block_u=1)]


def rel_err(got, ref, floor_frac):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    return float((np.abs(got - ref) / np.maximum(np.abs(ref), floor_frac * np.abs(ref).max())).max())


def test_config5_batch_against_oracle(require_gpu):
    import torch
    dev = torch.device("cuda")
    specs = S.config5_specs(2024)
    assert len(specs) == 8
    rng = np.random.default_rng(5)
    scans, items = [], []
    for k, sp in enumerate(specs):
        s, vols = S.config5_setup(sp)
        kw = P.spectral_inputs(s, *s.det)
        pr = P.Projector(P.scan_params(**s.scan_kwargs()))
        pr.set_volumes(vols, storage_dtype=P.CTFP_F16)
        pr.set_spectral(**kw)
        pr.set_tuning(**TUNINGS[k % len(TUNINGS)])
        I = torch.empty((s.views, s.det[1], s.det[0]), device=dev)
        items.append((pr, 0, s.views, I, torch.empty_like(I)))
        scans.append((s, vols, kw, int(rng.integers(0, s.views))))
        print(f"config5 scan {k}: {sp}", flush=True)
    samples = 0
    with P.Batch(items) as batch:
        samples = batch.samples
        batch.run()
        torch.cuda.synchronize()
    assert samples == sum(it[2] * it[0].samples_per_view for it in items)
    for k, ((pr, _, nviews, I, IA), (s, vols, kw, view)) in enumerate(zip(items, scans)):
        rI, rIA = pr.forward()
        I0 = pr.flat_field()
        torch.cuda.synchronize()
        assert torch.equal(I, rI) and torch.equal(IA, rIA), k
        g = O.geom_from_setup(s)
        p = O.paras(g)
        L64 = O.march_f64(g, p, vols, view, 1)
        det = np.stack([O.det_response(kw["det_factor"], float(m), kw["scint_thickness"])
                        for m in kw["scint_mu"]])
        I64, IA64, I064 = O.spectrum_f64(L64, kw["atten"], kw["spectrum"], kw["i0_val"], kw.get("grid"), det)
        Iv = I[view:view + 1].cpu().numpy()
        IAv = IA[view:view + 1].cpu().numpy()
        I0n = I0.cpu().numpy().astype(np.float64)
        e = dict(I=rel_err(Iv, I64, 1e-3), IA=rel_err(IAv, IA64, 1e-3),
                 p=rel_err(np.log(I0n[None] / Iv), np.log(I064[None] / I64), 1e-2))
        print(f"config5 scan {k}, view {view} vs FP64: {e}", flush=True)
        for key, v in e.items():
            assert v < 1e-3, (k, key, e)
        assert (IAv > 0).mean() > 0.1  # the phantom is in the field of view
    for it in items:
        it[0].close()
