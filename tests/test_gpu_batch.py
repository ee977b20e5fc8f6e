"""Batched persistent launch (ctfp_batch_*) and the autotuner on the GPU.

A batch runs each item through the same tile code as a single-scan launch,
so its outputs must equal each projector's own forward bit for bit; the
captured graph is replayed several times with other GPU work in between.
The autotuner may only change speed: projections before and after tuning
are bit-identical."""
import numpy as np
import pytest

import ct_forward_projection_amd as P
from ct_forward_projection_amd import synthetic as S

pytestmark = pytest.mark.gpu


def make_item(n, n_mat, det, views, kvp, grid, seed, dtype):
    w, b, i = S.three_material_phantom(n, seed=seed)
    vols = [w, b, i][:n_mat]
    s = S.config2()
    s.n_vox, s.length_mm, s.det, s.views = (n, n, n), (n * 0.5,) * 3, det, views
    s.d_size = 1.2 * n * 0.5 * 1.5 / det[1]
    s.kvp = kvp
    s.materials = s.materials[:n_mat]
    if not grid:
        s.grid = None
    proj = P.Projector(P.scan_params(**s.scan_kwargs()))
    proj.set_volumes(vols, storage_dtype=dtype)
    proj.set_spectral(**P.spectral_inputs(s, *s.det))
    return proj


@pytest.mark.parametrize("dtype", [P.CTFP_F32, P.CTFP_F16])
def test_batch_equals_individual_forwards(require_gpu, dtype):
    import torch
    dev = torch.device("cuda")
    specs = [(40, 3, (72, 56), 7, 120.0, True), (24, 1, (40, 40), 5, 80.0, False),
             (56, 2, (96, 64), 9, 100.0, True), (32, 3, (48, 80), 3, 140.0, False)]
    projs = [make_item(*sp, seed=k + 3, dtype=dtype) for k, sp in enumerate(specs)]
    # uneven view ranges, one not starting at view 0
    ranges = [(0, 7), (1, 4), (2, 7), (0, 3)]
    tunings = [dict(), dict(wave_u=16, block_u=1), dict(chunk=32, lds_bytes=24 * 1024),
               dict(wave_u=4, block_u=4)]
    refs, items = [], []
    for proj, (vb, vc), tn in zip(projs, ranges, tunings):
        proj.set_tuning(**tn)
        refs.append(proj.forward(vb, vc))
        I = torch.full((vc, proj.nv, proj.nu), float("nan"), device=dev)
        items.append((proj, vb, vc, I, torch.full_like(I, float("nan"))))
    with P.Batch(items) as batch:
        assert batch.samples == sum(vc * p.samples_per_view for p, (vb, vc) in zip(projs, ranges))
        for rep in range(3):
            for _, _, _, I, IA in items:
                I.fill_(float("nan"))
                IA.fill_(float("nan"))
            batch.run()
            torch.cuda.synchronize()
            for k, ((_, _, _, I, IA), (rI, rIA)) in enumerate(zip(items, refs)):
                assert torch.equal(I, rI) and torch.equal(IA, rIA), (rep, k)
            # eager work between replays
            projs[0].forward(0, 2)
            torch.cuda.synchronize()
    for p in projs:
        p.close()


def test_batch_rejects_changed_inputs(require_gpu):
    """Volumes or spectral tables replaced after the batch was built (same
    sizes, so the volume buffer is reused) make ctfp_batch_run fail."""
    import torch
    proj = make_item(24, 2, (32, 32), 3, 90.0, False, seed=2, dtype=P.CTFP_F32)
    I = torch.empty((3, 32, 32), device="cuda")
    w, b, _ = S.three_material_phantom(24, seed=9)
    with P.Batch([(proj, 0, 3, I, torch.empty_like(I))]) as batch:
        batch.run()
        torch.cuda.synchronize()
        proj.set_volumes([w, b], storage_dtype=P.CTFP_F16)
        with pytest.raises(P.CTFPError):
            batch.run()
    with P.Batch([(proj, 0, 3, I, torch.empty_like(I))]) as batch:
        batch.run()
        torch.cuda.synchronize()
        s = S.config2()
        s.n_vox, s.length_mm, s.det, s.views, s.kvp = (24,) * 3, (12.0,) * 3, (32, 32), 3, 90.0
        s.materials, s.grid = s.materials[:2], None
        proj.set_spectral(**P.spectral_inputs(s, 32, 32))
        with pytest.raises(P.CTFPError):
            batch.run()
    proj.close()


def test_batch_rejects_volume_past_4gib(require_gpu):
    """A scan whose padded volume spans 4 GiB or more (2040 x 2040 x 520,
    8.76 GB) runs through ctfp_forward but is refused as a batch item."""
    import torch
    s = S.config2()
    s.materials = (S.WATER,)
    s.n_vox, s.length_mm = (2040, 2040, 520), (204.0, 204.0, 52.0)
    s.det, s.d_size, s.views = (32, 32), 0.075, 2
    proj = P.Projector(P.scan_params(**s.scan_kwargs()))
    vol = torch.zeros((520, 2040, 2040), dtype=torch.float32, device="cuda")
    proj.set_volumes([vol])
    del vol
    proj.set_spectral(**P.spectral_inputs(s, *s.det))
    I, IA = proj.forward(0, 2)
    I0 = proj.flat_field()
    torch.cuda.synchronize()
    assert torch.equal(I, I0.expand_as(I))  # an empty volume: the flat field
    with pytest.raises(P.CTFPError):
        P.Batch([(proj, 0, 2, torch.empty_like(I), torch.empty_like(IA))])
    proj.close()


def test_autotune_keeps_results(require_gpu):
    import torch
    proj = make_item(64, 3, (128, 96), 12, 120.0, True, seed=4, dtype=P.CTFP_F16)
    ref = proj.forward()
    torch.cuda.synchronize()
    tuned = proj.autotune()
    assert tuned["march"] == "lds" and tuned["chunk"] % 16 == 0
    assert tuned == proj.get_tuning()
    got = proj.forward()
    torch.cuda.synchronize()
    assert torch.equal(got[0], ref[0]) and torch.equal(got[1], ref[1]), tuned
    proj.close()
