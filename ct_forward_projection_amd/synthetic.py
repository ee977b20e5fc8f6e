"""Synthetic inputs: spectra, mass-attenuation tables, seeded phantoms and the
two scan set-ups the projector is measured on.

No measured physics tables ship with this package.  Mass attenuation is a
two-term model, mu/rho(E) = (Z/A) N_A sigma_KN(E) + c E^-p (with K-edge jumps),
whose coefficients are fitted by hand to textbook values at a few energies; it
is smooth, positive and has the K-edges in the right places, which is all the
projector needs.  Energy bin k (0-based) is (k+1) keV, the reference's indexing
(Source/CBCTPolyForwardProjGrid.cpp:40, 387), so a spectrum starting at 20 keV
has 19 leading zero bins.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

AVOGADRO = 0.602214076  # 1e24 / mol, so (Z/A) * N_A * sigma[barn] is cm^2/g
R_E2_2PI_BARN = 2.0 * math.pi * 0.0794079  # 2 pi r_e^2 in barn


def klein_nishina_barn(e_kev: np.ndarray) -> np.ndarray:
    """Total Klein-Nishina cross-section per electron (barn)."""
    k = np.asarray(e_kev, dtype=np.float64) / 510.99895
    l2 = np.log1p(2.0 * k)
    return R_E2_2PI_BARN * ((1 + k) / k**2 * (2 * (1 + k) / (1 + 2 * k) - l2 / k)
                            + l2 / (2 * k) - (1 + 3 * k) / (1 + 2 * k) ** 2)


@dataclass(frozen=True)
class Material:
    name: str
    z_over_a: float
    pe_coef: float          # photoelectric coefficient above the highest edge
    pe_power: float
    edges: tuple = ()       # ((edge keV, jump ratio), ...)
    density: float = 1.0    # g/cm^3 (nominal)

    def mass_atten(self, e_kev) -> np.ndarray:
        e = np.asarray(e_kev, dtype=np.float64)
        compton = self.z_over_a * AVOGADRO * klein_nishina_barn(e)
        pe = self.pe_coef * e ** (-self.pe_power)
        for edge, jump in self.edges:
            pe = np.where(e < edge, pe / jump, pe)
        return compton + pe


WATER = Material("water", 0.5551, 4752.0, 3.0, (), 1.0)
BONE = Material("bone", 0.5148, 31100.0, 3.0, (), 1.92)
ALUMINIUM = Material("aluminium", 0.4818, 25500.0, 3.0, (), 2.699)
IODINE = Material("iodine", 0.4177, 7.1e5, 2.8, ((33.17, 6.5),), 4.93)
CAESIUM = Material("caesium", 0.4139, 7.9e5, 2.8, ((35.98, 6.3),), 1.93)
LEAD = Material("lead", 0.3958, 2.33e6, 2.8, ((88.0, 5.0),), 11.35)


def csi_mass_atten(e_kev) -> np.ndarray:
    """CsI by mass fraction (Cs 0.5115, I 0.4885)."""
    return 0.5115 * CAESIUM.mass_atten(e_kev) + 0.4885 * IODINE.mass_atten(e_kev)


def bin_energies(n_bins: int) -> np.ndarray:
    """Energy (keV) of each bin: bin k is (k+1) keV."""
    return np.arange(1, n_bins + 1, dtype=np.float64)


def kramers_spectrum(kvp: float, e_min: float, al_mm: float) -> np.ndarray:
    """Kramers' spectrum (kVp - E)/E integrated over each 1-keV bin
    [E-0.5, E+0.5] (clipped at kVp), filtered by `al_mm` of aluminium,
    normalised to sum 1.  Returns int(kVp) bins; bins below e_min are zero."""
    n = int(round(kvp))
    out = np.zeros(n, dtype=np.float64)
    for k in range(n):
        e = k + 1.0
        if e < e_min:
            continue
        lo, hi = e - 0.5, min(e + 0.5, kvp)
        if hi <= lo:
            continue
        val = kvp * math.log(hi / lo) - (hi - lo)   # integral of (kVp - E)/E
        mu_al = ALUMINIUM.mass_atten(np.array([e]))[0] * ALUMINIUM.density
        out[k] = val * math.exp(-mu_al * al_mm / 10.0)
    return (out / out.sum()).astype(np.float32)


# ----------------------------------------------------------------- phantoms

def _ellipsoid_mask(shape, centre, semi, rot):
    """Boolean mask of a rotated ellipsoid in index space, evaluated over its
    bounding box; returns (slices, mask)."""
    r = float(max(semi))
    lo = [max(0, int(math.floor(c - r - 1))) for c in centre]
    hi = [min(s, int(math.ceil(c + r + 2))) for c, s in zip(centre, shape)]
    if any(h <= l for l, h in zip(lo, hi)):
        return None, None
    z, y, x = np.meshgrid(*[np.arange(l, h, dtype=np.float32) for l, h in zip(lo, hi)],
                          indexing="ij")
    p = np.stack([z - centre[0], y - centre[1], x - centre[2]], axis=-1) @ rot.astype(np.float32)
    q = (p[..., 0] / semi[0]) ** 2 + (p[..., 1] / semi[1]) ** 2 + (p[..., 2] / semi[2]) ** 2
    return tuple(slice(l, h) for l, h in zip(lo, hi)), q <= 1.0


def _random_rotation(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    a, b, c, d = q
    return np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)],
                     [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                     [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]])


def cylinder_phantom(n: int = 128, radius: float = 55.0, height: float = 110.0,
                     n_ellipsoids: int = 10, seed: int = 20231) -> np.ndarray:
    """Single-material density volume [z][y][x] (g/cm^3): a water cylinder
    (1.0 g/cc, axis along z, centred) holding `n_ellipsoids` seeded rotated
    ellipsoids with semi-axes U[3,20] voxels and density U[0.2,1.9], drawn in
    order with later ones overwriting earlier ones."""
    rng = np.random.default_rng(seed)
    vol = np.zeros((n, n, n), dtype=np.float32)
    c = (n - 1) / 2.0
    yy, xx = np.meshgrid(np.arange(n) - c, np.arange(n) - c, indexing="ij")
    disk = (yy**2 + xx**2) <= radius**2
    z0 = int(round(c - height / 2)); z1 = int(round(c + height / 2))
    vol[max(z0, 0):min(z1, n)][:, disk] = 1.0
    for _ in range(n_ellipsoids):
        while True:  # centre inside the cylinder
            cz = rng.uniform(z0 + 10, z1 - 10)
            cy, cx = rng.uniform(-radius + 10, radius - 10, size=2)
            if cy * cy + cx * cx <= (radius - 10) ** 2:
                break
        semi = rng.uniform(3.0, 20.0, size=3)
        rho = rng.uniform(0.2, 1.9)
        sl, m = _ellipsoid_mask(vol.shape, (cz, cy + c, cx + c), semi, _random_rotation(rng))
        if sl is not None:
            vol[sl][m] = rho
    return vol


def three_material_phantom(n: int = 512, seed: int = 7122) -> tuple:
    """Water / bone / iodine density channels [z][y][x] (g/cm^3).  An elliptical
    water body (semi-axes 0.45n x 0.35n, height 0.86n) with 6 seeded bone
    ellipsoids (1.2-1.9 g/cc, water removed where bone is) and 8 seeded iodine
    inserts (2-20 mg/mL of iodine in water)."""
    rng = np.random.default_rng(seed)
    water = np.zeros((n, n, n), dtype=np.float32)
    bone = np.zeros_like(water)
    iodine = np.zeros_like(water)
    c = (n - 1) / 2.0
    ay, ax = 0.35 * n, 0.45 * n
    yy, xx = np.meshgrid(np.arange(n) - c, np.arange(n) - c, indexing="ij")
    body = (yy / ay) ** 2 + (xx / ax) ** 2 <= 1.0
    z0, z1 = int(round(0.07 * n)), int(round(0.93 * n))
    water[z0:z1][:, body] = 1.0

    def inside_centre(margin):
        while True:
            cz = rng.uniform(z0 + margin, z1 - margin)
            cy = rng.uniform(-ay + margin, ay - margin)
            cx = rng.uniform(-ax + margin, ax - margin)
            if (cy / (ay - margin)) ** 2 + (cx / (ax - margin)) ** 2 <= 1.0:
                return cz, cy + c, cx + c

    for _ in range(6):
        semi = rng.uniform(0.03 * n, 0.09 * n, size=3)
        centre = inside_centre(float(max(semi)) + 2)
        sl, m = _ellipsoid_mask(water.shape, centre, semi, _random_rotation(rng))
        if sl is not None:
            bone[sl][m] = rng.uniform(1.2, 1.9)
            water[sl][m] = 0.0
    for _ in range(8):
        semi = rng.uniform(0.015 * n, 0.04 * n, size=3)
        centre = inside_centre(float(max(semi)) + 2)
        sl, m = _ellipsoid_mask(water.shape, centre, semi, _random_rotation(rng))
        if sl is not None:
            iodine[sl][m] = rng.uniform(0.002, 0.020)
            water[sl][m] = 1.0
            bone[sl][m] = 0.0
    return water, bone, iodine


def label_phantom(n: int, n_mat: int, seed: int = 5) -> np.ndarray:
    """uint8 label volume with labels 0..n_mat (and a few labels above n_mat,
    which the reference treats as empty) in nested seeded ellipsoids."""
    rng = np.random.default_rng(seed)
    lab = np.zeros((n, n, n), dtype=np.uint8)
    c = (n - 1) / 2.0
    zz, yy, xx = np.meshgrid(*(np.arange(n) - c,) * 3, indexing="ij")
    lab[(yy / (0.42 * n)) ** 2 + (xx / (0.38 * n)) ** 2 + (zz / (0.45 * n)) ** 2 <= 1] = 1
    for i in range(2 * n_mat + 2):
        semi = rng.uniform(0.06 * n, 0.16 * n, size=3)
        centre = rng.uniform(0.3 * n, 0.7 * n, size=3)
        sl, m = _ellipsoid_mask(lab.shape, tuple(centre), semi, _random_rotation(rng))
        if sl is not None:
            lab[sl][m] = (i % (n_mat + 1)) + 1   # includes n_mat + 1 (an unknown label)
    return lab


# ------------------------------------------------------------- scan set-ups

@dataclass
class ScanSetup:
    """Everything one polyenergetic scan needs, in the reference's units."""
    name: str
    n_vox: tuple            # (nx, ny, nz)
    length_mm: tuple        # (lx, ly, lz)
    det: tuple              # (nu, nv)
    d_size: float
    views: int
    sod: float = 1000.0
    sdd: float = 1500.0
    rot: int = 1
    kvp: float = 80.0
    e_min: float = 20.0
    al_mm: float = 2.0
    i0_val: float = 1.0e5
    materials: tuple = (WATER,)
    grid: dict | None = None
    scint_density: float = 4.51
    scint_thickness: float = 0.6
    scint_thickness_err: float = 0.0
    det_factor: float = 1.0
    energy_weighting: bool = False
    extra: dict = field(default_factory=dict)

    def scan_kwargs(self) -> dict:
        return dict(p_num_x=self.n_vox[0], p_num_y=self.n_vox[1], p_num_z=self.n_vox[2],
                    p_length_x=self.length_mm[0], p_length_y=self.length_mm[1],
                    p_length_z=self.length_mm[2], proj_num=self.views,
                    d_num_u=self.det[0], d_num_v=self.det[1], sod=self.sod, sdd=self.sdd,
                    rotated_direction=self.rot, d_size=self.d_size)

    @property
    def n_energy(self) -> int:
        return int(round(self.kvp))

    def spectrum(self) -> np.ndarray:
        return kramers_spectrum(self.kvp, self.e_min, self.al_mm)

    def atten(self) -> np.ndarray:
        """[n_energy][n_mat] attenuation per unit density, 1/mm per g/cm^3."""
        e = bin_energies(self.n_energy)
        return np.stack([m.mass_atten(e) / 10.0 for m in self.materials], axis=1).astype(np.float32)

    def scint_mu(self) -> np.ndarray:
        """[n_energy] CsI linear attenuation (1/mm) (CBCTForwardProj.cpp:247)."""
        e = bin_energies(self.n_energy)
        return (csi_mass_atten(e).astype(np.float32) * np.float32(self.scint_density)
                / np.float32(10))

    def grid_mu(self) -> tuple:
        """[n_energy] strip and interspace linear attenuation (1/mm).  The
        reference's interspace attenuation is always zero: its grid driver
        never copies InterspaceDensity (CBCTPolyForwardProjGrid.cpp:53-63)."""
        e = bin_energies(self.n_energy)
        g = self.grid
        mu_s = (LEAD.mass_atten(e).astype(np.float32) * np.float32(g["strip_density"])
                / np.float32(10))
        return mu_s, np.zeros_like(mu_s)


def config1() -> ScanSetup:
    """128^3 single-material water/ellipsoid phantom, 1 mm voxels; 256x256
    detector, 0.8 mm; 90 views; 80 kVp Kramers, 2 mm Al, 61 bins (20-80 keV);
    0.6 mm CsI with energy-integrating E*eta weighting; no grid."""
    return ScanSetup(name="config1", n_vox=(128, 128, 128), length_mm=(128.0, 128.0, 128.0),
                     det=(256, 256), d_size=0.8, views=90, kvp=80.0, e_min=20.0, al_mm=2.0,
                     materials=(WATER,), grid=None, scint_thickness=0.6,
                     energy_weighting=True)


def config2() -> ScanSetup:
    """3 x 512^3 water/bone/iodine, 0.5 mm voxels; 1024x768 detector, 0.4 mm;
    360 views; 120 kVp, 101 bins (20-120 keV); focused Pb grid 10:1,
    40 lines/cm (36 um strips, 214 um interspace, h = 2.14 mm), FD 1500 mm;
    0.6 mm CsI response."""
    return ScanSetup(name="config2", n_vox=(512, 512, 512), length_mm=(256.0, 256.0, 256.0),
                     det=(1024, 768), d_size=0.4, views=360, kvp=120.0, e_min=20.0, al_mm=2.5,
                     materials=(WATER, BONE, IODINE),
                     grid=dict(strip_width=36, interspace_width=214, height_mm=2.14,
                               focal_mm=1500.0, strip_density=11.35),
                     scint_thickness=0.6)


def config4() -> ScanSetup:
    """Differentiable projector: 3 x 256^3 water/bone/iodine, 1 mm voxels;
    512x512 detector, 0.8 mm; 180 views; 120 kVp, 101 bins (20-120 keV);
    the config-2 Pb grid and 0.6 mm CsI response."""
    s = config2()
    s.name = "config4"
    s.n_vox, s.length_mm = (256, 256, 256), (256.0, 256.0, 256.0)
    s.det, s.d_size, s.views = (512, 512), 0.8, 180
    return s



def config5_specs(seed: int = 2024) -> list:
    """The eight uneven scans of the batched workload (seeded): cubic phantoms
    of {128, 192, 256, 384, 512} voxels at 0.5 mm with 1-3 materials,
    detectors from 256x256 to 1024x768, 60-360 views, 80-140 kVp spectra
    (61-121 active 1-keV bins from 20 keV), grid on or off."""
    rng = np.random.default_rng(seed)
    dets = [(256, 256), (512, 384), (512, 512), (768, 576), (1024, 768)]
    out = []
    for k in range(8):
        n = int(rng.choice([128, 192, 256, 384, 512]))
        out.append(dict(n=n, n_mat=int(rng.integers(1, 4)), det=dets[int(rng.integers(0, len(dets)))],
                        views=int(rng.integers(60, 361)), kvp=float(rng.choice(np.arange(80, 141, 10))),
                        grid=bool(rng.random() < 0.5), seed=100 + k))
    return out


def config5_setup(spec: dict) -> tuple:
    """(ScanSetup, volumes) of one scan of config5_specs: the config-2 scanner
    and tables with the scan's phantom size, materials, detector, views,
    tube voltage and grid; the detector spans the phantom's diagonal at 1.5x
    magnification."""
    s = config2()
    n = spec["n"]
    s.name = f"batch{spec['seed']}"
    s.n_vox, s.length_mm = (n, n, n), (n * 0.5,) * 3
    s.det, s.views, s.kvp = tuple(spec["det"]), spec["views"], spec["kvp"]
    s.d_size = float(n * 0.5 * 1.45 * 1.5 / min(spec["det"]))
    s.materials = s.materials[:spec["n_mat"]]
    if not spec["grid"]:
        s.grid = None
    vols = list(three_material_phantom(n, seed=spec["seed"]))[:spec["n_mat"]]
    return s, vols
