"""The adjoint (ctfp_backward) on the GPU: against the FP64 CPU adjoint of the
oracle (spectrum chain rule + exact backprojection), against central finite
differences of the GPU forward on 64 random voxels (1e-3), and with LDS
accumulation against the global-atomic fallback."""
import numpy as np
import pytest

import ct_forward_projection_amd as P
from ct_forward_projection_amd import synthetic as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def rel_err(got, ref, floor_frac=1e-3):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    floor = floor_frac * np.abs(ref).max()
    return float((np.abs(got - ref) / np.maximum(np.abs(ref), floor)).max())


def setup_case(n=40, views=6, det=(72, 56), d_size=1.0, seed=5):
    w, b, i = S.three_material_phantom(n, seed=seed)
    s = S.config2()
    s.n_vox, s.length_mm, s.det, s.d_size, s.views = (n, n, n), (n * 0.5,) * 3, det, d_size, views
    s.scint_thickness_err = 0.05
    return s, [w, b, i]


def det_table(kw):
    return np.stack([O.det_response(kw["det_factor"], float(m), kw["scint_thickness"])
                     for m in kw["scint_mu"]])


@pytest.mark.parametrize("mode", ["gather", "cell", "scatter"])
@pytest.mark.parametrize("with_L", [False, True])
def test_adjoint_matches_fp64_cpu_adjoint(require_gpu, with_L, mode):
    import torch
    s, vols = setup_case()
    kw = P.spectral_inputs(s, *s.det)
    rng = np.random.default_rng(3)
    gI = rng.normal(size=(s.views, s.det[1], s.det[0])).astype(np.float32)
    gIA = rng.normal(size=gI.shape).astype(np.float32)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**kw)
        proj.set_tuning(adjoint=mode)
        _, _, Lg = proj.forward(want_L=True)
        dev = torch.device("cuda")
        grad = proj.backward(torch.from_numpy(gI).to(dev), torch.from_numpy(gIA).to(dev),
                             L_in=Lg if with_L else None)
        torch.cuda.synchronize()
        grad = grad.cpu().numpy()
    g = O.geom_from_setup(s)
    p = O.paras(g)
    L64 = O.march_f64(g, p, vols)
    gL = O.spectrum_grad_f64(L64, kw["atten"], kw["spectrum"], kw["i0_val"], gI, gIA, kw["grid"],
                             det_table(kw))
    ref = O.backproject_f64(g, p, gL)
    # Each voxel's gradient sums hundreds of rays' contributions of both
    # signs, so its rounding scales with the sum of their magnitudes: the
    # backprojection of |dl/dL|.
    scale = O.backproject_f64(g, p, np.abs(gL))
    err = np.abs(grad - ref) / np.maximum(scale, 1e-3 * scale.max())
    assert err.max() < 1e-4, err.max()


@pytest.mark.parametrize("n_mat", [1, 2, 4])
@pytest.mark.parametrize("mode", ["gather", "cell", "scatter"])
def test_adjoint_anisotropic_material_counts(require_gpu, n_mat, mode):
    """A non-cubic volume (40 x 32 x 24) with 0.6 / 0.45 / 0.8 mm voxels and
    1, 2 or 4 materials: the adjoint against the FP64 CPU adjoint."""
    import torch
    w, b, i = S.three_material_phantom(40, seed=13)
    base = [w, b, i, 0.5 * w]
    vols = [np.ascontiguousarray(v[:24, :32, :40], dtype=np.float32) for v in base[:n_mat]]
    s = S.config2()
    s.materials = (S.WATER, S.BONE, S.IODINE, S.WATER)[:n_mat]
    s.n_vox = (40, 32, 24)
    s.length_mm = (40 * 0.6, 32 * 0.45, 24 * 0.8)
    s.det, s.d_size, s.views = (72, 56), 0.9, 6
    kw = P.spectral_inputs(s, *s.det)
    rng = np.random.default_rng(17)
    gI = rng.normal(size=(s.views, s.det[1], s.det[0])).astype(np.float32)
    gIA = rng.normal(size=gI.shape).astype(np.float32)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**kw)
        proj.set_tuning(adjoint=mode)
        dev = torch.device("cuda")
        grad = proj.backward(torch.from_numpy(gI).to(dev), torch.from_numpy(gIA).to(dev))
        torch.cuda.synchronize()
        grad = grad.cpu().numpy()
    assert grad.shape == (n_mat, 24, 32, 40)
    g = O.geom_from_setup(s)
    p = O.paras(g)
    L64 = O.march_f64(g, p, vols)
    gL = O.spectrum_grad_f64(L64, kw["atten"], kw["spectrum"], kw["i0_val"], gI, gIA, kw["grid"],
                             det_table(kw))
    ref = O.backproject_f64(g, p, gL)
    scale = O.backproject_f64(g, p, np.abs(gL))
    err = np.abs(grad - ref) / np.maximum(scale, 1e-3 * scale.max())
    assert err.max() < 1e-4, err.max()


def test_adjoint_finite_differences(require_gpu):
    """dl/d(rho) at 64 random voxels against finite differences of the GPU
    forward, l = sum(c * I) + sum(d * IAbsorb).  The step per material is
    h_m = 0.4 / max_E A[E][m] over the bins with spectral weight (one voxel
    then moves any weighted tau by at most ~0.2), and the difference quotient
    is Richardson-extrapolated from steps h and h/2, which leaves a truncation
    error ~(A dL)^4 / 100 < 1e-5 while keeping dL large against the FP32
    rounding of the marched sums.  The denominators are the perturbations as
    stored in float32, not h itself (at h ~ 1e-3 the rounding of rho + h
    alone would be a 1e-4 error)."""
    import torch
    s, vols = setup_case(n=32, views=4, det=(56, 44))
    kw = P.spectral_inputs(s, *s.det)
    rng = np.random.default_rng(9)
    c = rng.normal(size=(s.views, s.det[1], s.det[0]))
    d = rng.normal(size=c.shape)
    dev = torch.device("cuda")
    proj = P.Projector(P.scan_params(**s.scan_kwargs()))
    proj.set_spectral(**kw)

    def loss(vs):
        proj.set_volumes(vs)
        I, IA = proj.forward()
        torch.cuda.synchronize()
        return float((I.cpu().numpy().astype(np.float64) * c).sum()
                     + (IA.cpu().numpy().astype(np.float64) * d).sum())

    def quotient(m, z, y, x, h):
        vp = [v.copy() for v in vols]
        vm = [v.copy() for v in vols]
        vp[m][z, y, x] += h
        vm[m][z, y, x] -= h
        return (loss(vp) - loss(vm)) / (float(vp[m][z, y, x]) - float(vm[m][z, y, x]))

    proj.set_volumes(vols)
    grad = proj.backward(torch.from_numpy(c.astype(np.float32)).to(dev),
                         torch.from_numpy(d.astype(np.float32)).to(dev))
    torch.cuda.synchronize()
    grad = grad.cpu().numpy().astype(np.float64)
    active = np.asarray(kw["spectrum"]) > 0
    step = 0.4 / np.asarray(kw["atten"], np.float64)[active].max(axis=0)
    n = vols[0].shape[0]
    errs = []
    while len(errs) < 64:
        m = int(rng.integers(0, 3))
        z, y, x = (int(v) for v in rng.integers(4, n - 4, size=3))
        if grad[m, z, y, x] == 0.0:
            continue
        h = float(np.float32(step[m]))
        fd = (4 * quotient(m, z, y, x, h / 2) - quotient(m, z, y, x, h)) / 3
        errs.append(abs(fd - grad[m, z, y, x]) / max(abs(fd), 1e-3 * np.abs(grad[m]).max()))
    proj.close()
    assert max(errs) < 1e-3, (max(errs), np.median(errs))


def test_gather_adjoint_is_deterministic_and_matches_scatter(require_gpu):
    """The voxel and cell gathers sum every voxel in a fixed order
    (bit-identical on repeat); the scatter adjoint reaches the same gradient
    through atomics.
    The volume (45 x 37 x 30) does not divide into 8 x 8 x 4 bricks, and the
    view range does not start at view 0."""
    import torch
    w, b, i = S.three_material_phantom(48, seed=6)
    vols = [np.ascontiguousarray(v[:30, :37, :45]) for v in (w, b, i)]
    s = S.config2()
    s.n_vox, s.length_mm, s.det, s.d_size, s.views = (45, 37, 30), (22.5, 18.5, 15.0), (84, 60), 0.9, 9
    kw = P.spectral_inputs(s, *s.det)
    rng = np.random.default_rng(8)
    dev = torch.device("cuda")
    gI = torch.from_numpy(rng.normal(size=(6, s.det[1], s.det[0])).astype(np.float32)).to(dev)
    gIA = torch.from_numpy(rng.normal(size=gI.shape).astype(np.float32)).to(dev)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**kw)
        proj.set_tuning(adjoint="gather")
        a = proj.backward(gI, gIA, view_begin=2, view_count=6)
        b2 = proj.backward(gI, gIA, view_begin=2, view_count=6)
        torch.cuda.synchronize()
        assert torch.equal(a, b2)
        proj.set_tuning(adjoint="cell")
        e = proj.backward(gI, gIA, view_begin=2, view_count=6)
        e2 = proj.backward(gI, gIA, view_begin=2, view_count=6)
        torch.cuda.synchronize()
        assert torch.equal(e, e2)
        proj.set_tuning(adjoint="scatter")
        c = proj.backward(gI, gIA, view_begin=2, view_count=6)
        torch.cuda.synchronize()
    g = O.geom_from_setup(s)
    p = O.paras(g)
    L64 = O.march_f64(g, p, vols, 2, 6)
    gL = O.spectrum_grad_f64(L64, kw["atten"], kw["spectrum"], kw["i0_val"], gI.cpu().numpy(),
                             gIA.cpu().numpy(), kw["grid"], det_table(kw))
    ref = O.backproject_f64(g, p, gL, view0=2)
    scale = O.backproject_f64(g, p, np.abs(gL), view0=2)
    for got in (a, e, c):
        err = np.abs(got.cpu().numpy() - ref) / np.maximum(scale, 1e-3 * scale.max())
        assert err.max() < 1e-4, err.max()


def test_adjoint_lds_equals_global_atomics(require_gpu):
    """Scatter adjoint: LDS-accumulated chunks against the all-global-atomics fallback (an LDS
    budget too small for any chunk box).  Both sum in FP32 atomics in an
    order that varies from run to run, so they agree to rounding, not
    bitwise.  With random-sign upstream gradients the sums cancel, and that
    rounding scales with the sum of the absolute contributions per voxel,
    which is the backprojection of |gI| (every ray factor is non-negative):
    the difference is measured against it, as in the FP64 adjoint test."""
    import torch
    s, vols = setup_case(n=48, views=5, det=(90, 70), d_size=0.9)
    kw = P.spectral_inputs(s, *s.det)
    rng = np.random.default_rng(4)
    dev = torch.device("cuda")
    gI = torch.from_numpy(rng.normal(size=(s.views, s.det[1], s.det[0])).astype(np.float32)).to(dev)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**kw)
        proj.set_tuning(march="lds", chunk=64, lds_bytes=1024, adjoint="scatter")
        ref = proj.backward(gI, None).cpu().numpy()
        scale = proj.backward(gI.abs(), None).cpu().numpy().astype(np.float64)
        floor = 1e-3 * scale.max()
        for march, chunk, lds in [("lds", 16, 40 * 1024), ("lds", 48, 12 * 1024), ("lds", 32, 2048),
                                  ("global", 16, 40 * 1024)]:
            proj.set_tuning(march=march, chunk=chunk, lds_bytes=lds)
            got = proj.backward(gI, None).cpu().numpy()
            err = float((np.abs(got.astype(np.float64) - ref) / np.maximum(scale, floor)).max())
            assert err < 1e-4, (march, chunk, lds, err)


@pytest.mark.parametrize("mode", ["gather", "cell", "scatter"])
def test_adjoint_with_focal_spot(require_gpu, mode):
    import torch
    s, vols = setup_case(n=36, views=5, det=(64, 48))
    kw = P.spectral_inputs(s, *s.det)
    focal = P.focal_spot_offsets(s.views, 1.5, seed=10).astype(np.float64)
    rng = np.random.default_rng(12)
    gI = rng.normal(size=(s.views, s.det[1], s.det[0])).astype(np.float32)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**kw)
        proj.set_focal_spot(focal)
        proj.set_tuning(adjoint=mode)
        grad = proj.backward(torch.from_numpy(gI).cuda(), None)
        torch.cuda.synchronize()
        grad = grad.cpu().numpy()
    g = O.geom_from_setup(s)
    p = O.paras(g)
    L64 = O.march_f64(g, p, vols, focal=focal)
    gL = O.spectrum_grad_f64(L64, kw["atten"], kw["spectrum"], kw["i0_val"], gI, None, kw["grid"],
                             det_table(kw))
    ref = O.backproject_f64(g, p, gL, focal=focal)
    scale = O.backproject_f64(g, p, np.abs(gL), focal=focal)
    err = np.abs(grad - ref) / np.maximum(scale, 1e-3 * scale.max())
    assert err.max() < 1e-4, err.max()


@pytest.mark.parametrize("mode", ["gather", "cell"])
@pytest.mark.parametrize("case", range(6))
def test_adjoint_random_geometries(require_gpu, case, mode):
    """Seeded random scans (volume shape and pitch per axis, SOD and SDD,
    rotation direction, detector size and pitch, 1-3 materials of balls and
    boxes): the voxel gather (bricks staging per-pixel records, 2 x 2 voxel
    quads) and the cell gather (each sample weighted once, into its cell's
    corners, by the plane crossings) against the FP64 CPU adjoint."""
    import torch
    rng = np.random.default_rng(2000 + case)
    nx, ny, nz = (int(v) for v in rng.integers(10, 40, 3))
    pitch = rng.uniform(0.3, 1.2, 3)
    n_mat = int(rng.integers(1, 4))
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    vols = [np.zeros((nz, ny, nx), np.float32) for _ in range(n_mat)]
    for k in range(int(rng.integers(2, 5))):
        m = int(rng.integers(0, n_mat))
        c = rng.uniform(0, 1, 3) * (nx, ny, nz)
        r = rng.uniform(1.5, 0.4 * min(nx, ny, nz) + 1.5)
        vols[m][(xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 <= r * r] = rng.uniform(0.2, 2.0)
    s = S.config2()
    s.materials = (S.WATER, S.BONE, S.IODINE)[:n_mat]
    s.n_vox = (nx, ny, nz)
    s.length_mm = tuple(float(n * p_) for n, p_ in zip((nx, ny, nz), pitch))
    s.sod = float(rng.uniform(300.0, 1000.0))
    s.sdd = float(s.sod * rng.uniform(1.2, 2.5))
    s.rot = int(rng.choice([-1, 1]))
    s.det = (int(rng.integers(16, 64)), int(rng.integers(12, 48)))
    s.d_size = float(rng.uniform(0.4, 1.5)) * s.sdd / s.sod * float(pitch.min())
    s.views = 4
    s.scint_thickness_err = 0.05
    print(f"adjoint random geometry {case}: {nx}x{ny}x{nz} pitch {np.round(pitch, 2)} M {n_mat} "
          f"SOD {s.sod:.0f} SDD {s.sdd:.0f} det {s.det} {s.d_size:.2f} mm rot {s.rot}", flush=True)
    kw = P.spectral_inputs(s, *s.det)
    gI = rng.normal(size=(s.views, s.det[1], s.det[0])).astype(np.float32)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**kw)
        proj.set_tuning(adjoint=mode)
        grad = proj.backward(torch.from_numpy(gI).to(torch.device("cuda")))
        torch.cuda.synchronize()
        grad = grad.cpu().numpy()
    g = O.geom_from_setup(s)
    p = O.paras(g)
    L64 = O.march_f64(g, p, vols)
    gL = O.spectrum_grad_f64(L64, kw["atten"], kw["spectrum"], kw["i0_val"], gI, None, kw["grid"],
                             det_table(kw))
    ref = O.backproject_f64(g, p, gL)
    scale = O.backproject_f64(g, p, np.abs(gL))
    err = np.abs(grad - ref) / np.maximum(scale, 1e-3 * scale.max())
    assert err.max() < 1e-4, err.max()


@pytest.mark.parametrize("mode", ["gather", "cell", "scatter"])
def test_adjoint_high_magnification(require_gpu, mode):
    """A short source distance (150 mm, SDD 600 mm: magnification 4),
    anisotropic voxels (1.5 x 0.5 x 1.0 mm) and fine pixels (0.6 mm): a
    voxel step spans up to 10 pixels and the projection departs from its
    linearisation at a brick's centre by about a pixel at the brick's edge.
    The pixel windows must still hold every ray through each voxel's
    support (voxel gather) or cell (cell gather): both, and the scatter
    (which needs no windows), against the FP64 CPU adjoint."""
    import torch
    nx, ny, nz = 24, 40, 28
    rng = np.random.default_rng(77)
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    vols = [np.zeros((nz, ny, nx), np.float32) for _ in range(2)]
    vols[0][(xx - 11.5) ** 2 / 100 + (yy - 19.5) ** 2 / 300 + (zz - 13.5) ** 2 / 150 <= 1] = 1.0
    vols[1][(xx - 8) ** 2 + (yy - 25) ** 2 / 4 + (zz - 10) ** 2 <= 16] = 0.7
    s = S.config2()
    s.materials = (S.WATER, S.BONE)
    s.n_vox = (nx, ny, nz)
    s.length_mm = (nx * 1.5, ny * 0.5, nz * 1.0)
    s.sod, s.sdd = 150.0, 600.0
    s.det, s.d_size, s.views = (260, 150), 0.6, 4
    s.scint_thickness_err = 0.05
    kw = P.spectral_inputs(s, *s.det)
    gI = rng.normal(size=(s.views, s.det[1], s.det[0])).astype(np.float32)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**kw)
        proj.set_tuning(adjoint=mode)
        grad = proj.backward(torch.from_numpy(gI).to(torch.device("cuda")))
        torch.cuda.synchronize()
        grad = grad.cpu().numpy()
    g = O.geom_from_setup(s)
    p = O.paras(g)
    L64 = O.march_f64(g, p, vols)
    gL = O.spectrum_grad_f64(L64, kw["atten"], kw["spectrum"], kw["i0_val"], gI, None, kw["grid"],
                             det_table(kw))
    ref = O.backproject_f64(g, p, gL)
    scale = O.backproject_f64(g, p, np.abs(gL))
    err = np.abs(grad - ref) / np.maximum(scale, 1e-3 * scale.max())
    assert err.max() < 1e-4, err.max()


@pytest.mark.parametrize("mode", ["gather", "cell"])
@pytest.mark.parametrize("case", range(4))
def test_adjoint_random_high_magnification(require_gpu, case, mode):
    """Seeded random scans close to the source (SOD 100-300 mm,
    magnification 2-5, voxel pitch 0.3-2 mm per axis, pixels down to 0.3 of
    a magnified voxel): where a brick's projection bends most from its
    linearisation, the voxel and cell gathers against the FP64 CPU adjoint."""
    import torch
    rng = np.random.default_rng(3100 + case)
    nx, ny, nz = (int(v) for v in rng.integers(12, 36, 3))
    pitch = rng.uniform(0.3, 2.0, 3)
    zz, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    vols = [np.zeros((nz, ny, nx), np.float32)]
    for k in range(3):
        c = rng.uniform(0.2, 0.8, 3) * (nx, ny, nz)
        r = rng.uniform(2.0, 0.4 * min(nx, ny, nz) + 2.0)
        vols[0][(xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2 <= r * r] = rng.uniform(0.3, 1.5)
    s = S.config2()
    s.materials = (S.WATER,)
    s.n_vox = (nx, ny, nz)
    s.length_mm = tuple(float(n * p_) for n, p_ in zip((nx, ny, nz), pitch))
    half_diag = 0.5 * float(np.hypot(s.length_mm[0], s.length_mm[1]))
    s.sod = float(max(rng.uniform(100.0, 300.0), 1.5 * half_diag))
    s.sdd = float(s.sod * rng.uniform(2.0, 5.0))
    s.rot = int(rng.choice([-1, 1]))
    s.d_size = float(rng.uniform(0.3, 1.0)) * s.sdd / s.sod * float(pitch.min())
    s.det = (int(rng.integers(48, 160)), int(rng.integers(32, 120)))
    s.views = 3
    s.scint_thickness_err = 0.05
    print(f"high-magnification case {case}: {nx}x{ny}x{nz} pitch {np.round(pitch, 2)} SOD {s.sod:.0f} "
          f"SDD {s.sdd:.0f} det {s.det} {s.d_size:.2f} mm", flush=True)
    kw = P.spectral_inputs(s, *s.det)
    gI = rng.normal(size=(s.views, s.det[1], s.det[0])).astype(np.float32)
    with P.Projector(P.scan_params(**s.scan_kwargs())) as proj:
        proj.set_volumes(vols)
        proj.set_spectral(**kw)
        proj.set_tuning(adjoint=mode)
        grad = proj.backward(torch.from_numpy(gI).to(torch.device("cuda")))
        torch.cuda.synchronize()
        grad = grad.cpu().numpy()
    g = O.geom_from_setup(s)
    p = O.paras(g)
    L64 = O.march_f64(g, p, vols)
    gL = O.spectrum_grad_f64(L64, kw["atten"], kw["spectrum"], kw["i0_val"], gI, None, kw["grid"],
                             det_table(kw))
    ref = O.backproject_f64(g, p, gL)
    scale = O.backproject_f64(g, p, np.abs(gL))
    err = np.abs(grad - ref) / np.maximum(scale, 1e-3 * scale.max())
    assert err.max() < 1e-4, err.max()
