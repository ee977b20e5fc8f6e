"""AddressSanitizer + UndefinedBehaviorSanitizer on the CPU code: the oracle
(every function, through oracle/sanitize_driver.c) and the host half of the
library with the reference-shaped classes and their driver (the host-only
TestI0 scan, the INI and table readers with their error paths).  A report
ends the sanitized program with a non-zero status.  The sanitized driver
must also write the same files as the normal one."""
import os
import subprocess

import numpy as np
import pytest

from test_refapi import make_case

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "ct_forward_projection_amd", "csrc")
SAN_CLI = os.path.join(CSRC, "_build", "ct_forward_projection_san")
CLI = os.path.join(ROOT, "ct_forward_projection_amd", "bin", "ct_forward_projection")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:exitcode=23",
           UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1:exitcode=24")


@pytest.fixture(scope="module")
def built():
    for d in (CSRC, os.path.join(ROOT, "oracle")):
        subprocess.run(["make", "-s", "-C", d, "sanitize"], check=True, capture_output=True, timeout=600)


def test_oracle_under_sanitizers(built):
    r = subprocess.run([os.path.join(ROOT, "oracle", "_build", "oracle_san")], capture_output=True, text=True,
                       timeout=300, env=ENV)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "ok" in r.stdout


@pytest.mark.parametrize("grid", [0, 1])
def test_test_i0_under_sanitizers(built, tmp_path, grid):
    san = tmp_path / "san"
    ref = tmp_path / "ref"
    for d in (san, ref):
        make_case(str(d), test_i0=1, grid=grid)
    r = subprocess.run([SAN_CLI, str(san)], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0, r.stderr[-4000:]
    r2 = subprocess.run([CLI, str(ref)], capture_output=True, text=True, timeout=600)
    assert r2.returncode == 0, r2.stderr
    out = os.path.join("OutputResult", "NoGridResponseTestI0")
    names = sorted(os.listdir(ref / out))
    assert names and names == sorted(os.listdir(san / out))
    for n in names:
        assert (san / out / n).read_bytes() == (ref / out / n).read_bytes(), n


@pytest.mark.parametrize("edit,code", [(("IsGrid=1;", "IsGrid=7;"), 6), (("MaterialNum=2", "MaterialNum=17"), 6),
                                       (("DSize=1.2", "DSize=-1"), 6)])
def test_config_errors_under_sanitizers(built, tmp_path, edit, code):
    make_case(str(tmp_path), test_i0=1)
    p = os.path.join(tmp_path, "Config", "ScanSetting.ini")
    s = open(p).read()
    assert edit[0] in s
    open(p, "w").write(s.replace(edit[0], edit[1], 1))
    r = subprocess.run([SAN_CLI, str(tmp_path)], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == code, (r.returncode, r.stderr[-4000:])


def test_bad_table_under_sanitizers(built, tmp_path):
    make_case(str(tmp_path), test_i0=1)
    path = os.path.join(tmp_path, "InputData", "Tables", "CsI.txt")
    v = np.loadtxt(path)
    with open(path, "w") as f:
        f.write("\n".join(["1.5"] * 3 + ["nan"] + [str(x) for x in v[4:]]) + "\n")
    r = subprocess.run([SAN_CLI, str(tmp_path)], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 1 and "not a finite number" in r.stderr, (r.returncode, r.stderr[-4000:])


def test_scan_needs_a_device_in_the_sanitizer_build(built, tmp_path):
    """The full scan path reaches the device stubs and reports it cleanly."""
    make_case(str(tmp_path), test_i0=0)
    r = subprocess.run([SAN_CLI, str(tmp_path)], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 7 and "no device" in r.stderr, (r.returncode, r.stderr[-4000:])
