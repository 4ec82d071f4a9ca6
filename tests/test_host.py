"""CPU-only checks of the host side: synthetic workloads, the C-ABI library's exports,
and the Fortran interface module."""
import ctypes
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from superlite_amd import synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_header_and_library_exports():
    from superlite_amd import slgpu
    if not os.path.exists(slgpu.LIB):
        pytest.skip("libslgpu.so not built")
    L = ctypes.CDLL(slgpu.LIB)
    hdr = open(os.path.join(ROOT, "include", "slgpu.h")).read()
    declared = set(re.findall(r"\b(slgpu_[a-z_0-9]+)\s*\(", hdr))
    assert declared == set(slgpu.EXPORTS)
    for name in declared:
        assert hasattr(L, name), name


@pytest.mark.parametrize("name", ["c1", "c2", "c5"])
def test_workload_state(name):
    w = synth.WORKLOADS[name]
    st = synth.load_state(name)
    x, v, m, T, mf = synth.ejecta(w)
    np.testing.assert_array_equal(st["xarr"], x)
    assert st["mass"].shape == (w.nx,)
    if w.csm is None:
        assert st["mass"].sum() == pytest.approx(1.4 * synth.PC_MSUN, rel=1e-12)
    # homologous flow: divv = 3/t (grid_setup.f:48-51)
    np.testing.assert_allclose(st["divv"], 3.0 / synth.T_SNAP, rtol=1e-12)
    # 4 pi/3 (x+^3 - x-^3) (grid_volume.f:82-87)
    np.testing.assert_allclose(st["vol"], 4 * np.pi / 3 * np.diff(x ** 3), rtol=1e-12)
    assert np.all(st["oc"] >= 0) and st["oc"].shape == (w.nx, len(st["ion_iz"]))
    assert np.all(np.abs(st["natomfr"].sum(0) - 1.0) < 1e-12)


def test_line_list_properties():
    w = synth.WORKLOADS["c1"]
    st = synth.load_state("c1")
    wl0, gxs, chilw, iz, ii = synth.line_list(w, st["ion_iz"], st["ion_ii"])
    assert len(wl0) == w.nline
    assert np.all(np.diff(wl0) >= 0)
    assert wl0.dtype == np.float32 and iz.dtype == np.int16
    assert np.all((chilw >= 0) & (chilw <= 10 * 8065.55))
    assert set(zip(iz.tolist(), ii.tolist())) <= set(zip(st["ion_iz"].tolist(), st["ion_ii"].tolist()))
    # deterministic from the seed
    again = synth.line_list(w, st["ion_iz"], st["ion_ii"])
    assert np.array_equal(again[0], wl0) and np.array_equal(again[1], gxs)


def test_fortran_interface_compiles(tmp_path):
    fc = shutil.which("flang") or shutil.which("amdflang")
    if not fc:
        pytest.skip("no Fortran compiler")
    src = os.path.join(ROOT, "superlite_amd", "fortran", "slgpu_iface.f90")
    r = subprocess.run([fc, "-c", src, "-module-dir", str(tmp_path), "-o", str(tmp_path / "i.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_fortran_interface_binds_every_entry_point():
    """every function declared in include/slgpu.h has a bind(C) interface of that name"""
    import re
    hdr = open(os.path.join(ROOT, "include", "slgpu.h")).read()
    declared = set(re.findall(r"\b(slgpu_[a-z_0-9]+)\s*\(", hdr))
    f90 = open(os.path.join(ROOT, "superlite_amd", "fortran", "slgpu_iface.f90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(slgpu_[a-z_0-9]+)"\)', f90))
    assert declared, "no declarations found"
    assert declared <= bound, sorted(declared - bound)
