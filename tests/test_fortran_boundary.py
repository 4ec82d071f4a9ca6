"""The Fortran boundary end to end: a Fortran host (tests/fortran/slgpu_e2e.f90) compiled with
flang against superlite_amd/fortran/slgpu_iface.f90 and linked to libslgpu.so runs one whole
iteration of workload c1 (opacity build, DDMC preparation, sources, transport, flux tally)
through the C ABI; its outputs must equal those of the same calls made through the Python
binding.  Marked gpu; skipped where no flang is installed."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from superlite_amd import synth

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _flang():
    return shutil.which("flang") or shutil.which("amdflang")


def _write_input(path, w, st, wl, lines, tables, nstot, it, seed):
    nx, ng = w.nx, w.ng
    nion = len(st["ion_iz"])
    hdr = np.array([nx, ng, int(st["nep"]), int(st["nepg"]), nion, len(lines[0]), nstot, it, seed,
                    0, 0, 1], np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([w.tauddmc, w.taulump, w.lbol], np.float64).tobytes())
        for a, dt in ((st["xarr"], np.float64), (st["vxarr"], np.float64), (st["divv"], np.float64),
                      (st["vol"], np.float64), (st["rcell"], np.float64), (wl, np.float64),
                      (st["ion_iz"], np.int32), (st["ion_ii"], np.int32),
                      (tables[0], np.float32), (tables[1], np.float32), (tables[2], np.float64),
                      (lines[0], np.float32), (lines[1], np.float32), (lines[2], np.float32),
                      (lines[3], np.int16), (lines[4], np.int16),
                      (st["temp"], np.float64), (st["natom"], np.float64), (st["vol"], np.float64),
                      (st["mass"], np.float64), (st["nelec"], np.float64),
                      (np.asarray(st["natomfr"]).T, np.float64), (st["oc"], np.float64),
                      (st["ginv"], np.float64), (st["temp"], np.float64)):
            f.write(np.ascontiguousarray(a, dt).tobytes())


def _read_output(path, w, st):
    nx, ng, nep = w.nx, w.ng, int(st["nep"])
    buf = open(path, "rb").read()
    pos = [0]

    def take(n, dt):
        a = np.frombuffer(buf, dt, n, pos[0])
        pos[0] += a.nbytes
        return a.copy()
    out = {}
    out["cap"] = take(ng * nx, np.float32).reshape(nx, ng)
    out["emiss"] = take(ng * nx, np.float32).reshape(nx, ng)
    for k in ("sig", "capgrey", "capross"):
        out[k] = take(nx, np.float64)
    out["opaclump"] = take(11 * nx, np.float64).reshape(nx, 11)
    out["emitprob"] = take(nep * nx, np.float64).reshape(nx, nep)
    out["emit"] = take(nx, np.float64)
    out["sthermal"] = take(1, np.float64)[0]
    out["nvol"] = take(nx, np.int32)
    out["nnew"] = int(take(1, np.int64)[0])
    out["tally"] = take(nx, np.float64)
    out["jrad"] = take(ng * nx, np.float64).reshape(nx, ng)
    out["eamp"] = take(nx, np.float64)
    out["totals"] = take(7, np.float64)
    out["counters"] = take(10, np.int64)
    out["luminos"] = take(ng, np.float64)
    out["lumdev"] = take(ng, np.float64)
    out["lumnum"] = take(ng, np.int32)
    assert pos[0] == len(buf)
    return out


def test_fortran_host_iteration_c1(tmp_path, tables):
    fc = _flang()
    if not fc:
        pytest.skip("no Fortran compiler (flang) on this machine")
    from superlite_amd import slgpu
    say("case: Fortran host, one c1 iteration through the C ABI")
    libdir = os.path.join(ROOT, "superlite_amd")
    exe = str(tmp_path / "slgpu_e2e")
    r = subprocess.run([fc, "-O1", "-J", str(tmp_path), "-c",
                        os.path.join(ROOT, "superlite_amd", "fortran", "slgpu_iface.f90"),
                        "-o", str(tmp_path / "iface.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([fc, "-O1", "-J", str(tmp_path),
                        os.path.join(ROOT, "tests", "fortran", "slgpu_e2e.f90"),
                        str(tmp_path / "iface.o"), "-L" + libdir, "-lslgpu",
                        "-Wl,-rpath," + libdir, "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    w, st, wl, lines = synth.build("c1")
    nstot, it, seed = 1 << 12, 3, 11
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    _write_input(fin, w, st, wl, lines, tables, nstot, it, seed)
    say("running the Fortran host")
    r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=240)
    say(r.stdout.strip(), r.stderr.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    f = _read_output(fout, w, st)

    # the same iteration through the Python binding
    ctx = slgpu.setup_workload(w, st, wl, lines, tables, seed=seed, npartmax=nstot)
    op = ctx.opacity_build(st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"],
                           st["natomfr"], st["oc"], st["ginv"])
    prep = ctx.ddmc_prep(st["temp"], w.lbol)
    nvol, nnew = ctx.sources(nstot, it)
    tr = ctx.transport(it)
    fl = ctx.fluxtally()
    ctx.close()
    # deterministic outputs: bit-identical
    for k in ("cap", "emiss", "sig", "capgrey", "capross"):
        assert np.array_equal(f[k], op[k]), k
    for k in ("opaclump", "emitprob", "emit"):
        assert np.array_equal(f[k], prep[k]), k
    assert f["sthermal"] == prep["sthermal"]
    assert np.array_equal(f["nvol"], nvol) and f["nnew"] == nnew == nstot
    ct = [tr["nactive"], tr["nstepimc"], tr["nstepddmc"], tr["nstepmax"], tr["nmethswap"],
          tr["nflux"], tr["ndead"], tr["nerr"], tr["nwarn"]]
    assert list(f["counters"][:9]) == ct
    assert np.array_equal(f["lumnum"], fl["lumnum"])
    # FP64 atomic tallies: equal up to the order of the atomic additions
    for k, v in (("tally", tr["tally"]), ("jrad", tr["jrad"]), ("luminos", fl["luminos"]),
                 ("lumdev", fl["lumdev"])):
        np.testing.assert_allclose(f[k], v, rtol=1e-12, atol=0, err_msg=k)
    tot = [tr["eout"], tr["evelo"], tr["sflux"], tr["estart"], tr["edead"], tr["esrc"], tr["eampsum"]]
    np.testing.assert_allclose(f["totals"], tot, rtol=1e-12)
    say("Fortran host matches the Python binding:", f["counters"][1] + f["counters"][2], "particle-steps")
