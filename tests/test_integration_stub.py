"""The reference-side binding and the Fortran create_bins, on the CPU.

* The "Option 2" stub of INTEGRATION.md (the bodies of tangential_shear_tpcf
  and scalar_kappa_tpcf replaced by calls through the C-ABI) compiles inside
  the reference's own module tpcf, with its unmodified wrappers shear_2pcf
  and kappa_2pcf (src/tpcf.f90:134-188), and a program calling them links
  against lib2pcf_amd.so (tests/fortran/check_integration_stub.sh).
* create_bins of module tpcf_gpu accepts and refuses the argument forms the
  reference's create_bins (src/utils.f90:139-230) does, with the same n_bin,
  x_max and bin_size and edges and centres within FP32 rounding (the library
  computes the edges in FP64 closed form, the reference in FP32): the
  (x_min, x_max, n_bin) and (x_min, bin_size, n_bin) forms for all four
  spacings; the (x_min, x_max, bin_size) form stops with "Binning is
  ill-defined." in both (src/utils.f90:163-206).

Needs the reference tree and flang (present where the reference is built);
skipped elsewhere.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = os.path.join(ROOT, "tests", "fortran", "check_integration_stub.sh")
FC = os.environ.get("FC", "/opt/rocm/llvm/bin/flang")
REF = os.environ.get("TPCF_REFERENCE", "/root/reference")

pytestmark = pytest.mark.skipif(
    not (os.path.exists(os.path.join(REF, "src", "tpcf.f90")) and os.path.exists(FC)
         and os.path.exists(os.path.join(ROOT, "2pcf_amd", "lib2pcf_amd.so"))),
    reason="needs the reference tree, flang and the built library")


@pytest.fixture(scope="module")
def programs(tmp_path_factory):
    d = tmp_path_factory.mktemp("stub")
    ref, gpu = str(d / "bins_ref"), str(d / "bins_gpu")
    r = subprocess.run(["sh", SCRIPT, ref, gpu], capture_output=True, text=True, timeout=600)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    assert "compiled and linked" in r.stdout
    return ref, gpu


def run(prog, form):
    return subprocess.run([prog, str(form)], capture_output=True, text=True, timeout=60)


def test_stub_builds_in_reference_module(programs):
    pass   # the fixture builds the stub module and links the driver


@pytest.mark.parametrize("form", range(1, 9))
def test_create_bins_forms_match_reference(programs, form):
    ref, gpu = programs
    a, b = run(ref, form), run(gpu, form)
    assert a.returncode == 0 and b.returncode == 0, (a.stderr, b.stderr)
    va = [float(v) for v in a.stdout.split()]
    vb = [float(v) for v in b.stdout.split()]
    n = int(va[0])
    assert int(vb[0]) == n and len(va) == len(vb) == 3 + (n + 1) + n
    np.testing.assert_allclose(vb[1:], va[1:], rtol=2e-6)


def test_create_bins_ill_defined_form_refused_by_both(programs):
    ref, gpu = programs
    a, b = run(ref, 9), run(gpu, 9)
    assert a.returncode != 0 and b.returncode != 0
    assert "Binning is ill-defined." in a.stdout + a.stderr
    assert "Binning is ill-defined." in b.stdout + b.stderr
