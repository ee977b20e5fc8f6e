"""ctfp_backward_noalloc (the adjoint with its workspace in a static device
array, for callers that cannot allocate) against ctfp_backward, through the C
ABI: two calls in a row, the second on a smaller scan, must give the same
gradients bit for bit under the voxel gather (both run one group of views),
and agree to FP32 rounding with ctfp_backward's default cell gather (the
no-allocation call runs the voxel gather); a scan whose dl/dL needs several
groups of the static workspace agrees to FP32 rounding; a scan one view of
which exceeds the workspace is rejected."""
import numpy as np
import pytest

import ct_forward_projection_amd as P
from ct_forward_projection_amd import _lib as L
from ct_forward_projection_amd import synthetic as S

pytestmark = pytest.mark.gpu


def make_case(n, det, views, d_size, seed, n_mat=3):
    w, b, i = S.three_material_phantom(n, seed=seed)
    s = S.config2()
    s.n_vox, s.length_mm, s.det, s.d_size, s.views = (n, n, n), (n * 0.5,) * 3, det, d_size, views
    s.materials = s.materials[:n_mat]
    s.scint_thickness_err = 0.05
    return s, [w, b, i][:n_mat]


def grads(proj, s, seed):
    """(ctfp_backward, ctfp_backward_noalloc) for the same seeded upstream
    gradients, both through the C ABI on the current stream."""
    import torch
    dev = torch.device("cuda")
    rng = np.random.default_rng(seed)
    shape = (s.views, s.det[1], s.det[0])
    gI = torch.from_numpy(rng.normal(size=shape).astype(np.float32)).to(dev)
    gIA = torch.from_numpy(rng.normal(size=shape).astype(np.float32)).to(dev)
    gshape = (proj.n_mat, s.n_vox[2], s.n_vox[1], s.n_vox[0])
    out = []
    stream = torch.cuda.current_stream().cuda_stream
    for name in ("ctfp_backward", "ctfp_backward_noalloc"):
        g = torch.zeros(gshape, dtype=torch.float32, device=dev)
        fn = getattr(L.lib(), name)
        L.check(fn(proj._ctx, 0, s.views, gI.data_ptr(), gIA.data_ptr(), None, g.data_ptr(), stream))
        torch.cuda.synchronize()
        out.append(g.cpu().numpy())
    return out


def test_noalloc_backward_equals_backward_twice(require_gpu):
    for k, (n, det, views, seed) in enumerate([(48, (80, 64), 10, 7), (32, (48, 40), 6, 8)]):
        print(f"noalloc call {k + 1}: {n}^3, detector {det}, {views} views", flush=True)
        s, vols = make_case(n, det, views, 1.0, seed)
        with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
            proj.set_volumes(vols)
            proj.set_spectral(**P.spectral_inputs(s, *s.det))
            cell, got_cell = grads(proj, s, seed=20 + k)
            proj.set_tuning(adjoint="gather")
            ref, got = grads(proj, s, seed=20 + k)
        assert np.abs(ref).max() > 0
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(got_cell, ref)
        assert np.abs(cell - ref).max() <= 1e-5 * np.abs(ref).max()


def test_noalloc_backward_view_groups(require_gpu):
    """1024 x 768 pixels x 3 materials is 9.4 MB of dl/dL per view: the 64 MiB
    workspace takes 7 views per group, so 16 views run in 3 groups (one
    group in ctfp_backward).  The groups add to the gradient in another
    order, so the two agree to FP32 rounding of the sum."""
    s, vols = make_case(40, (1024, 768), 16, 0.1, 9)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**P.spectral_inputs(s, *s.det))
        ref, got = grads(proj, s, seed=31)
    scale = np.abs(ref).max()
    assert scale > 0
    assert np.abs(got - ref).max() <= 1e-5 * scale


def test_noalloc_backward_rejects_oversized_view(require_gpu):
    """4 materials x 2100 x 2100 pixels x 4 B = 70.6 MB > 64 MiB for one view."""
    import torch
    s, vols = make_case(8, (2100, 2100), 1, 0.05, 3, n_mat=3)
    vols = vols + [vols[0]]
    s.materials = tuple(s.materials) + (S.WATER,)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**P.spectral_inputs(s, *s.det, with_response=False))
        dev = torch.device("cuda")
        gI = torch.zeros((1, 2100, 2100), dtype=torch.float32, device=dev)
        g = torch.zeros((4, 8, 8, 8), dtype=torch.float32, device=dev)
        rc = L.lib().ctfp_backward_noalloc(proj._ctx, 0, 1, gI.data_ptr(), None, None, g.data_ptr(),
                                           torch.cuda.current_stream().cuda_stream)
        assert rc == L.CTFP_ERR_UNSUPPORTED
        assert b"static workspace" in L.lib().ctfp_last_error()
