"""The Fortran host side (module tpcf_gpu, 2pcf_amd/fortran/tpcf_gpu.f90)
without a GPU: its ISO_C_BINDING interfaces name exactly the C-ABI's
correlation, binning and context entry points with matching argument counts,
and the Fortran driver stops with the library's status instead of computing
anything on the CPU."""
import os
import re
import subprocess

import numpy as np
import pytest

import fortran_io as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FSRC = os.path.join(ROOT, "2pcf_amd", "fortran", "tpcf_gpu.f90")
HEADER = os.path.join(ROOT, "include", "tpcf_gpu.h")


def c_prototypes():
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for m in re.finditer(r"\b(tpcf_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", text):
        args = [a for a in m.group(2).split(",") if a.strip() and a.strip() != "void"]
        out[m.group(1)] = len(args)
    return out


def fortran_interfaces():
    text = open(FSRC).read()
    out = {}
    for m in re.finditer(r"function\s+(tpcf_[a-z0-9_]+)\s*\(([^)]*)\)\s*(?:&\s*)?(?:bind\(C\))?", text, flags=re.S):
        args = [a for a in re.sub(r"&\s*\n\s*", "", m.group(2)).split(",") if a.strip()]
        out[m.group(1)] = len(args)
    return out


def test_fortran_interfaces_match_header():
    c, f = c_prototypes(), fortran_interfaces()
    assert set(f) <= set(c), set(f) - set(c)
    for name in ("tpcf_shear_f32", "tpcf_shear_f64", "tpcf_kappa_f32", "tpcf_kappa_f64", "tpcf_create_bins",
                 "tpcf_ctx_create", "tpcf_ctx_destroy"):
        assert name in f, name
        assert f[name] == c[name], (name, f[name], c[name])


def test_fortran_bins_struct_matches_header():
    # bind(C) mirror of tpcf_bins: same members, same order
    text = open(FSRC).read()
    block = re.search(r"type, bind\(C\) :: tpcf_bins(.*?)end type", text, flags=re.S).group(1)
    names = re.findall(r"::\s*([\w ,()+A-Z_]+)", block)
    flat = [n.split("(")[0].strip() for part in names for n in part.split(",")]
    assert flat == ["x_min", "x_max", "bin_size", "n_bin", "spacing", "ihs_theta", "edges", "bin_centers"]


@pytest.mark.skipif(not os.path.exists(F.DRIVER), reason="Fortran driver not built (flang absent)")
def test_driver_stops_without_gpu(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = np.random.default_rng(1)
    lens = dict(x=r.uniform(0, 10, 50), y=r.uniform(0, 10, 50), val=r.normal(0, 1, 50), w=np.ones(50))
    src = dict(x=r.uniform(0, 10, 60), y=r.uniform(0, 10, 60), kappa_E=r.normal(0, 1, 60),
               kappa_B=r.normal(0, 1, 60), w=np.ones(60))
    F.write_input(tmp_path / "in", 1, 1, 5, 0.5, 5.0, 1.0, lens, src)
    p = subprocess.run([F.DRIVER, str(tmp_path / "in"), str(tmp_path / "out")], capture_output=True, text=True)
    assert p.returncode != 0
    assert "status" in p.stdout and " 7" in p.stdout      # TPCF_ERR_NODEVICE
    assert not (tmp_path / "out").exists()
