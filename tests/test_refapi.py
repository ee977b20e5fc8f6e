"""The reference-shaped C++ layer through its command-line driver: INI parsing
in the shipped Config/ScanSetting.ini format (tab/';' comments, backslash
paths, case-insensitive OutPutFolder.ini), configuration errors, the host-only
TestI0 mode and (on a GPU) full scans written to the reference's file formats,
checked against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

from ct_forward_projection_amd import synthetic as S
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "ct_forward_projection_amd", "bin", "ct_forward_projection")


def write_table(path, values):
    with open(path, "w") as f:
        for v in values:
            f.write(f"{float(v):.10f}\n")


def make_case(tmp, *, test_i0=0, response=1, grid=1, single=0, n=24, nu=20, nv=16, views=4,
              dsize=1.2, kvp=40, terr=0.05, materials=2, photons=60000, labels=None):
    """Writes Config/, InputData/ and a label volume the way the reference's
    Config/ScanSetting.ini lays them out; returns the inputs used."""
    os.makedirs(os.path.join(tmp, "Config"), exist_ok=True)
    os.makedirs(os.path.join(tmp, "InputData", "Tables"), exist_ok=True)
    e = S.bin_energies(kvp)
    spec = S.kramers_spectrum(kvp, 20.0, 1.0)
    write_table(os.path.join(tmp, "InputData", "Tables", "spectrum.dat"), spec)
    mats = [S.WATER, S.BONE, S.ALUMINIUM, S.IODINE][:materials]
    for i, m in enumerate(mats):
        write_table(os.path.join(tmp, "InputData", "Tables", f"mat{i + 1}.dat"), m.mass_atten(e))
    write_table(os.path.join(tmp, "InputData", "Tables", "CsI.txt"), S.csi_mass_atten(e))
    write_table(os.path.join(tmp, "InputData", "Tables", "Pb.txt"), S.LEAD.mass_atten(e))
    if labels is None:
        labels = S.label_phantom(n, materials, seed=12)
    labels.tofile(os.path.join(tmp, "InputData", "labels.raw"))
    dens = [1.0, 1.92, 2.699, 4.93][:materials]
    cfg = f"""[ScanPara]
IsTestI0={test_i0};
IsDetResponse={response};
IsGrid={grid};
IsSingleMaterial={single};
SDD=600\t\t\t\t\t;mm
SOD=400\t\t\t\t\t;mm
ScanAngleNum={views}
RotatedDirection=1;\t\t\t;-1 clockwise, 1 counter-clockwise

[RaySource]
Photons={photons}\t\t\t\t;photons per pixel
SpecEnergyNum={kvp}\t\t\t;bins
SpectrumPath=InputData\\Tables\\spectrum.dat

[PhantomVox]
PNumX={n}\t\t\t\t\t;pix
PNumY={n}
PNumZ={n}
PLengthX={n * 1.0}\t\t\t\t\t;mm
PLengthY={n * 1.0}
PLengthZ={n * 1.0}

[Detector]
DNumU={nu}\t\t\t\t\t;pix
DNumV={nv}
DSize={dsize}\t\t\t\t\t;mm

[Scintillator]
Density=4.510\t\t\t\t\t;g/cm^3
Thickness=0.4\t\t\t\t\t;mm
ThicknessErr={terr}\t\t\t\t;mm
PhotoelectConverFactor=1\t\t;Photoelectric conversion factor, 0<x<1
ScintillatorMAPath=InputData/Tables/CsI.txt

[PostGrid]
FD=600\t\t\t\t\t\t\t;mm
Height=1.5\t\t\t\t\t\t;mm
GridStripsWidth=36\t\t\t\t;um
InterspcaceWidth=214\t\t\t;um
GridStripMaterial=Pb
InterspaceMaterial=Air
GridStripDensity=11.35\t\t\t;g/cm^3
InterspaceDensity=0\t\t\t;g/cm^3
GridStripMAPath=InputData/Tables/Pb.txt
InterspaceMAPath=InputData/Tables/Pb.txt

[PhantomMaterial]
MaterialNum={materials}
StructureMarkPath=InputData/labels.raw
"""
    for i in range(materials):
        cfg += f"Density{i + 1}={dens[i]}\t\t\t\t\t;g/cm^3\nMassAttenPath{i + 1}=InputData/Tables/mat{i + 1}.dat\n"
    with open(os.path.join(tmp, "Config", "ScanSetting.ini"), "w") as f:
        f.write(cfg)
    with open(os.path.join(tmp, "Config", "OutPutFolder.ini"), "w") as f:
        for sec in ("GridNoResponse", "GridResponse", "NoGridNoResponse", "NoGridResponse",
                    "NoGridResponseTestI0"):
            f.write(f"[{sec}]\nOutFolder=OutputResult/{sec}\n\n")
    return dict(e=e, spec=spec, mats=mats, dens=np.array(dens, np.float32), labels=labels, n=n,
                nu=nu, nv=nv, views=views, dsize=dsize, kvp=kvp, terr=terr, photons=photons)


def run_cli(tmp, env=None):
    return subprocess.run([CLI, tmp], capture_output=True, text=True, timeout=600, env=env)


def read_raw(path, dtype, shape):
    return np.fromfile(path, dtype=dtype).reshape(shape)


def test_cli_exists():
    assert os.access(CLI, os.X_OK), "build the package first"


def test_testI0_matches_oracle(tmp_path):
    case = make_case(str(tmp_path), test_i0=1)
    r = run_cli(str(tmp_path))
    assert r.returncode == 0, r.stderr
    npix = case["nu"] * case["nv"]
    I0 = read_raw(os.path.join(tmp_path, "OutputResult", "NoGridResponseTestI0",
                               f"BgImgPro_NoGrid_{case['nu']}x{case['nv']}_float.raw"),
                  np.float32, (case["nv"], case["nu"]))
    # host loop of testI0 (CBCTPolyForwardProjNoGrid.cpp:188-264) in FP32 bin order
    spec = np.loadtxt(os.path.join(tmp_path, "InputData", "Tables", "spectrum.dat")).astype(np.float32)
    csi = np.loadtxt(os.path.join(tmp_path, "InputData", "Tables", "CsI.txt")).astype(np.float32)
    mu = csi * np.float32(4.51) / np.float32(10)
    t = O.scint_thickness(npix, 0.4, case["terr"])
    acc = np.zeros(npix, np.float32)
    for k in range(case["kvp"]):
        eta = O.det_response(1.0, float(mu[k]), t)
        acc = acc + (np.float32(case["photons"]) * spec[k]) * eta
    np.testing.assert_array_equal(I0.ravel(), acc)
    # addStandardI0 (CBCTForwardProj.cpp:92-95): the same sum for the nominal
    # thickness without jitter, printed with 6 significant digits
    m = re.search(r"gives I0 = ([0-9.eE+-]+) photons", r.stdout)
    assert m, r.stdout
    t0 = np.full(1, 0.4, np.float32)
    std = np.float32(0)
    for k in range(case["kvp"]):
        std = std + (np.float32(case["photons"]) * spec[k]) * O.det_response(1.0, float(mu[k]), t0)[0]
    assert float(m.group(1)) == pytest.approx(float(std), rel=1e-5)


@pytest.mark.parametrize("edit,msg", [
    (("IsGrid=1;", "IsGrid=7;"), "IsGrid"),
    (("SDD=600", "SDD=-3"), "SDD"),
    (("ScanAngleNum=4", "ScanAngleNum=-4"), "ScanAngleNum"),
    (("PNumY=24", "PNumY=abc"), "PNumX/Y/Z"),
    (("Density2=1.92", "Density2x=1.92"), "Density2"),
    (("MaterialNum=2", "MaterialNum=17"), "MaterialNum"),
    (("Thickness=0.4", "Thickness=-0.4"), "Scintillator"),
    (("FD=600", "FD=0"), "PostGrid"),
])
def test_config_errors(tmp_path, edit, msg):
    make_case(str(tmp_path), test_i0=0)
    p = os.path.join(tmp_path, "Config", "ScanSetting.ini")
    s = open(p).read()
    assert edit[0] in s
    open(p, "w").write(s.replace(edit[0], edit[1], 1))
    r = run_cli(str(tmp_path))
    assert r.returncode == 6, (r.returncode, r.stderr)
    assert msg in r.stderr and "configuration error" in r.stderr


def test_missing_table_is_io_error(tmp_path):
    make_case(str(tmp_path), test_i0=1)
    os.remove(os.path.join(tmp_path, "InputData", "Tables", "spectrum.dat"))
    r = run_cli(str(tmp_path))
    assert r.returncode == 5 and "spectrum" in r.stderr


def test_missing_config_is_reported(tmp_path):
    r = run_cli(str(tmp_path))
    assert r.returncode == 6 and "ScanSetting.ini" in r.stderr


# ------------------------------------------------------------------ GPU runs

def _oracle_scan(tmp, case, grid, response, single):
    n, nu, nv, views = case["n"], case["nu"], case["nv"], case["views"]
    g = O.geom(n, n, n, float(n), float(n), float(n), views, nu, nv, 400.0, 600.0, case["dsize"])
    p = O.paras(g)
    spec = np.loadtxt(os.path.join(tmp, "InputData", "Tables", "spectrum.dat")).astype(np.float32)
    mass = np.stack([np.loadtxt(os.path.join(tmp, "InputData", "Tables", f"mat{i + 1}.dat"))
                     for i in range(len(case["mats"]))], 1).astype(np.float32)
    det = None
    if response:
        csi = np.loadtxt(os.path.join(tmp, "InputData", "Tables", "CsI.txt")).astype(np.float32)
        mu = csi * np.float32(4.51) / np.float32(10)
        t = O.scint_thickness(nu * nv, 0.4, case["terr"])
        det = np.stack([O.det_response(1.0, float(m), t) for m in mu])
    T = None
    if grid:
        pb = np.loadtxt(os.path.join(tmp, "InputData", "Tables", "Pb.txt")).astype(np.float32)
        mus = pb * np.float32(11.35) / np.float32(10)
        d_um = int(np.float32(case["dsize"]) * np.float32(1000))
        T = np.stack([O.grid_transmission(nu, d_um, 214, 36, 1.5, 600.0, float(m), 0.0) for m in mus])
    if single:
        dens = O.labels_to_density_single(case["labels"], float(case["dens"][0]))
        L64 = O.march_f64(g, p, [dens])
        atten = mass[:, :1] / np.float32(10)
    else:
        ch = O.labels_to_density_channels(case["labels"], case["dens"])
        L64 = O.march_f64(g, p, list(ch))
        atten = mass / np.float32(10)
    I64, IA64, I064 = O.spectrum_f64(L64, atten, spec, float(case["photons"]), T, det)
    return I64, IA64, I064, det


@pytest.mark.gpu
@pytest.mark.parametrize("grid,response,single", [(1, 1, 0), (0, 1, 1), (1, 0, 1), (0, 0, 0)])
def test_cli_scan_matches_oracle(require_gpu, tmp_path, grid, response, single):
    case = make_case(str(tmp_path), grid=grid, response=response, single=single,
                     materials=1 if single else 3)
    r = run_cli(str(tmp_path))
    assert r.returncode == 0, r.stderr
    tag = "Grid" if grid else "NoGrid"
    sec = ("Grid" if grid else "NoGrid") + ("Response" if response else "NoResponse")
    out = os.path.join(tmp_path, "OutputResult", sec)
    nu, nv, v = case["nu"], case["nv"], case["views"]
    Iu = read_raw(os.path.join(out, f"PhoImgPro_{tag}_{nu}x{nv}x{v}_uint32.raw"), np.uint32, (v, nv, nu))
    IAu = read_raw(os.path.join(out, f"PhoImgAbsorb_{tag}_{nu}x{nv}x{v}_uint32.raw"), np.uint32, (v, nv, nu))
    I0 = read_raw(os.path.join(out, f"BgImgPro_{tag}_{nu}x{nv}_float.raw"), np.float32, (nv, nu))
    proj = read_raw(os.path.join(out, f"Projection_{tag}_{nu}x{nv}x{v}_float.raw"), np.float32, (v, nv, nu))
    I64, IA64, I064, det = _oracle_scan(str(tmp_path), case, grid, response, single)
    # floor(I) within one count of floor(I64) (the float path may straddle an integer)
    assert np.abs(Iu.astype(np.int64) - np.floor(I64).astype(np.int64)).max() <= 1
    assert np.abs(IAu.astype(np.int64) - np.floor(IA64).astype(np.int64)).max() <= 1
    np.testing.assert_allclose(I0, I064, rtol=1e-6)
    ok = Iu > 0
    np.testing.assert_allclose(proj[ok], np.log(I0[None] / Iu.astype(np.float32))[ok], rtol=1e-6)
    if grid:
        assert os.path.exists(os.path.join(out, f"GridPeriodIndex_{nu}.txt"))
    if single and response:
        a = np.loadtxt(os.path.join(out, "ADetResponse.txt"))
        np.testing.assert_allclose(a, det[:, (nu + nv) // 2], rtol=1e-5)
