"""SuperLite's on-disk formats (superlite_amd/io.py) against the reference's own readers and
writers: its Testsuite/Input decks (tests/golden/decks/) parsed by inputparmod.f and
inputstrmod.f, and the files OUTPUT/output_flux.f and OUTPUT/output_grid.f write for two
iterations of given values (tests/golden/ref_outputs.npz).  CPU only."""
import os

import numpy as np
import pytest

from superlite_amd import io

HERE = os.path.dirname(os.path.abspath(__file__))
DECKS = os.path.join(HERE, "golden", "decks")
NAMES = ["sn2017hcc-like", "s18", "sn1999em-like", "w7", "a4-10d-lte"]


@pytest.mark.parametrize("name", NAMES)
def test_read_input_par(golden, name):
    g = golden("ref_decks.npz")
    p = io.read_input_par(os.path.join(DECKS, name, "input.par"))
    assert list(p["in_ndim"]) == list(g[f"{name}__par_ndim"])
    assert list(p["in_flx_ndim"]) == list(g[f"{name}__par_flx_ndim"])
    iv = [p[k] for k in ("in_grp_ng", "in_src_n2s", "in_src_ns", "in_niter", "in_nomp", "in_grd_igeom")]
    assert iv == list(g[f"{name}__par_ivals"])
    rv = [p[k] for k in ("in_grp_wlmin", "in_grp_wlmax", "in_flx_wlmin", "in_flx_wlmax", "in_L_bol",
                         "in_trn_tauddmc", "in_taulump", "in_srcepwr", "in_L_inner")]
    assert rv == list(g[f"{name}__par_rvals"])
    lv = [int(p[k]) for k in ("in_puretran", "in_trn_errorfatal", "in_trn_noamp", "in_thm_aniso",
                              "in_nobbopac", "in_nobfopac", "in_noffopac", "in_nothmson", "in_nlte",
                              "in_noeos")]
    assert lv == list(g[f"{name}__par_lvals"])
    c = io.config_from_pars(p)
    assert c["nx"] == p["in_ndim"][0] and c["nstot"] == 2 ** p["in_src_n2s"]
    assert c["ampfact"] is False and c["flags"] == 0


@pytest.mark.parametrize("name", NAMES)
def test_read_input_str(golden, name):
    """every array the reference's reader fills, exactly; where the reference stops on an
    inconsistent deck (hazard H25), the same message"""
    g = golden("ref_decks.npz")
    nx = int(g[f"{name}__par_ndim"][0])
    path = os.path.join(DECKS, name, "input.str")
    if f"{name}__str_error" in g:
        with pytest.raises(ValueError, match=str(g[f"{name}__str_error"])):
            io.read_input_str(path, (nx, 1, 1))
        return
    s = io.read_input_str(path, (nx, 1, 1))
    for k in ("xleft", "vxleft", "mass", "temp"):
        assert np.array_equal(s[k], g[f"{name}__str_{k}"]), k
    assert np.array_equal(s["massfr"], g[f"{name}__str_massfr"].T)
    assert np.array_equal(s["iabund"], g[f"{name}__str_iabund"])
    flags = g[f"{name}__str_flags"]
    for k, f in zip(("vol", "radtemp", "lum", "ye"), flags):
        assert (s[k] is not None) == bool(f), k
        if f:
            assert np.array_equal(s[k], g[f"{name}__str_{k}"]), k
    mf = io.massfr30(s)
    assert mf.shape == (30, nx) and np.all(mf >= 0)


def test_read_input_str_errors(tmp_path):
    """the reference's stop conditions of read_inputstr on hand-made structure files"""
    good = "# 1D spherical\n#  2 1 1 5 1\n# x_left x_right vx_left vx_right mass temp h\n"
    def write(text):
        p = tmp_path / "input.str"
        p.write_text(text)
        return str(p)
    rows = "0 1 0 1 5 100 1\n1 2 1 2 5 100 1\n"
    # 5 columns declared: the label read takes 5 names (the rest of the record is dropped),
    # so there is no mass column
    with pytest.raises(ValueError, match="mass label not found"):
        io.read_input_str(write(good + rows), (2, 1, 1))
    hdr = "# 1D spherical\n#  2 1 1 7 1\n# x_left x_right vx_left vx_right mass temp h\n"
    s = io.read_input_str(write(hdr + rows), (2, 1, 1))
    assert list(s["xleft"]) == [0, 1, 2] and list(s["iabund"]) == [1]
    with pytest.raises(ValueError, match="incompatible nx"):
        io.read_input_str(write(hdr + rows), (3, 1, 1))
    with pytest.raises(ValueError, match="body too long"):
        io.read_input_str(write(hdr + rows + "9\n"), (2, 1, 1))
    # extra items on the last body record are dropped with it, as the reference's read does
    s2 = io.read_input_str(write(hdr + rows.rstrip("\n") + " 9 9\n\n"), (2, 1, 1))
    assert np.array_equal(s2["mass"], s["mass"])
    with pytest.raises(ValueError, match="format err: body"):
        io.read_input_str(write(hdr + "0 1 0 1 5 100 1\n1 2 1 2 5\n"), (2, 1, 1))
    with pytest.raises(ValueError, match="x grid not increasing"):
        io.read_input_str(write(hdr + "0 1 0 1 5 100 1\n1 1 1 2 5 100 1\n"), (2, 1, 1))
    with pytest.raises(ValueError, match="mass<0"):
        io.read_input_str(write(hdr + "0 1 0 1 -5 100 1\n1 2 1 2 5 100 1\n"), (2, 1, 1))
    with pytest.raises(ValueError, match="no such element"):
        io.read_input_str(write(hdr.replace(" h\n", " qq\n") + rows), (2, 1, 1))
    bad = tmp_path / "input.par"
    bad.write_text("&inputpars\n in_ndim = 4, 1, 1\n in_nosuchpar = 3\n/\n")
    with pytest.raises(ValueError, match="ivalid parameters"):
        io.read_input_par(str(bad))


def test_writers_byte_for_byte(golden, tmp_path):
    """two iterations through OutputWriter reproduce every file of the reference's writers"""
    g = golden("ref_outputs.npz")
    w = io.OutputWriter(str(tmp_path), g["xarr"], g["wl"], opacdump=True)
    for it in range(2):
        a = {k[len(f"in{it}_"):]: g[k] for k in g if k.startswith(f"in{it}_")}
        t = a["tot"]
        w.iteration(a["luminos"], a["lumnum"], a["lumdev"], a["nvol"], 1.0 / a["tempinv"],
                    a["radtemp"], a["capgrey"], a["capross"], a["sig"], a["tally"], a["vol"],
                    t[0], t[1], t[2], t[3])
    files = sorted(os.listdir(tmp_path))
    assert files == list(g["files"])
    for f in files:
        assert open(tmp_path / f).read() == str(g["text_" + f]), f


@pytest.mark.parametrize("x,w,e,expect", [
    (1.23456e41, 12, None, "  1.2346E+41"), (-3.3e41, 12, None, " -3.3000E+41"),
    (0.0, 12, None, "  0.0000E+00"), (3e-120, 12, None, "  3.0000-120"),
    (9.6817e-2, 12, 3, " 9.6817E-002"), (1.0e100, 12, 3, " 1.0000E+100")])
def test_fortran_e_format(x, w, e, expect):
    assert io.fmt_e(x, w, 4, e) == expect
