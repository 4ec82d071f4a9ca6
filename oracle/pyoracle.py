"""ctypes access to oracle/liboracle.so, the CPU FP64 restatement of SuperLite's hot path.

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker / reported CPU baseline, never as the product.
Array layouts follow oracle.h (reference memory order, 0-based storage).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_d = ctypes.c_double
_i = ctypes.c_int
_vp = ctypes.c_void_p
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so not built (make -C oracle)")
        _lib = ctypes.CDLL(LIB)
        _lib.or_binsrch.restype = _i
        _lib.or_binsrch.argtypes = [_d, _vp, _i, _i]
        _lib.or_bfxs.restype = _d
        _lib.or_bfxs.argtypes = [_i, _i, _d, _vp, _vp]
        _lib.or_specint0.restype = _d
        _lib.or_specint0.argtypes = [_d, _vp, _i]
        _lib.or_dopspeccalc.restype = _d
        _lib.or_dopspeccalc.argtypes = [_d, _vp, _i]
        _lib.or_specintv.argtypes = [_d, _vp, _i, _i, _vp]
        _lib.or_spectintv.argtypes = [_d, _vp, _i, _vp]
        _lib.or_emitgroup.restype = _i
        _lib.or_emitgroup.argtypes = [_vp, _d, _i]
        _lib.or_sourcenumbers.restype = ctypes.c_int64
        _lib.or_sourcenumbers.argtypes = [_i, _vp, ctypes.c_int64, _vp]
        _lib.or_sourceenergy.argtypes = [_i, _vp, _vp, _d, _vp, _vp]
        _lib.or_nlte_rad_est.argtypes = [_i, _i, _vp, _vp, _vp, _vp, _vp]
        _lib.or_iteration.restype = _i
        _lib.or_iteration.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32, _vp]
        _lib.or_sources.argtypes = [_vp, ctypes.c_uint32, ctypes.c_uint32] + [_vp] * 6
        _lib.or_opacity.argtypes = [_vp] + [_vp] * 6
        _lib.or_leakage.argtypes = [_vp, _vp]
        _lib.or_emission_probability.argtypes = [_vp, _vp]
        _lib.or_opacity_nlte.restype = _i
        _lib.or_opacity_nlte.argtypes = [_vp, _vp] + [_vp] * 9
        _lib.or_nlte_rate_matrix.restype = _i
        _lib.or_nlte_rate_matrix.argtypes = [_vp, _vp, _i, _vp, _vp]
        _lib.or_lu_solve.restype = _i
        _lib.or_lu_solve.argtypes = [_i, _vp, _vp, _vp]
        _lib.or_eos.restype = _i
        _lib.or_eos.argtypes = [_i, _vp] + [_vp] * 4 + [_d] + [_vp] * 3 + [_i] + [_vp] * 5
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class _OpacIn(ctypes.Structure):
    _fields_ = [("nc", _i), ("ng", _i), ("wl", _vp), ("temp", _vp), ("natom", _vp),
                ("vol", _vp), ("mass", _vp), ("nelec", _vp), ("natomfr", _vp),
                ("nion", _i), ("ion_iz", _vp), ("ion_ii", _vp), ("oc", _vp), ("ginv", _vp),
                ("nline", _i), ("wl0", _vp), ("gxs", _vp), ("chilw", _vp), ("liz", _vp),
                ("lii", _vp), ("bf_ph1", _vp), ("bf_ph2", _vp), ("ff_gff", _vp), ("flags", _i)]


class _Model(ctypes.Structure):
    _fields_ = [("nx", _i), ("xarr", _vp), ("vxarr", _vp), ("divv", _vp), ("vol", _vp),
                ("ng", _i), ("wl", _vp), ("wlinv", _vp), ("nep", _i), ("nepg", _i),
                ("tauddmc", _d), ("taulump", _d), ("puretran", _i), ("max_steps", _i),
                ("ampfact", _i), ("thm_aniso", _i),
                ("cap", _vp), ("sig", _vp), ("capgrey", _vp), ("tempinv", _vp),
                ("opaclump", _vp), ("emitprob", _vp), ("emit", _vp), ("radtemp", _vp),
                ("nvol", _vp), ("capemit", _vp), ("capemitgrey", _vp)]


class _Result(ctypes.Structure):
    _fields_ = [("tally", _vp), ("jrad", _vp), ("eamp", _vp), ("luminos", _vp), ("lumdev", _vp),
                ("lumnum", _vp), ("tot", _d * 7), ("ct", ctypes.c_int64 * 10),
                ("pstat", _vp), ("px", _vp), ("pmu", _vp), ("pwl", _vp), ("pe", _vp),
                ("pig", _vp), ("pix", _vp), ("pnstep", _vp), ("perr", _vp)]


def binsrch(x, arr, widerange):
    a = _c(arr, np.float64)
    return lib().or_binsrch(float(x), _p(a), len(a), int(widerange))


def bfxs(nz, ne, e, ph1, ph2):
    return lib().or_bfxs(nz, ne, float(e), _p(_c(ph1, np.float32)), _p(_c(ph2, np.float32)))


def specintv(tempinv, wlinv, n=None, offset=1):
    w = _c(wlinv, np.float64)
    n = len(w) - 1 if n is None else n
    ss = np.zeros(n)
    lib().or_specintv(float(tempinv), _p(w), n, offset, _p(ss))
    return ss


def spectintv(tempinv, wlinv):
    w = _c(wlinv, np.float64)
    ss = np.zeros(len(w) - 1)
    lib().or_spectintv(float(tempinv), _p(w), len(w) - 1, _p(ss))
    return ss


def specint0(tempinv, wlinv, ig):
    return lib().or_specint0(float(tempinv), _p(_c(wlinv, np.float64)), ig)


def dopspeccalc(tempinv, wlinv, ig):
    return lib().or_dopspeccalc(float(tempinv), _p(_c(wlinv, np.float64)), ig)


def opacity(wl, temp, natom, vol, mass, nelec, natomfr, ion_iz, ion_ii, oc, ginv,
            lines, tables, flags=0):
    """physical_opacity + means.  natomfr (30, nc); oc/ginv (nc, nion); lines =
    (wl0, gxs, chilw, iz, ii); tables = (ph1, ph2, gff).  Returns arrays with shape
    (nc, ng) for the group-resolved outputs (group-fastest memory, as gas_cap)."""
    nc = len(temp)
    ng = len(wl) - 1
    keep = dict(wl=_c(wl, np.float64), temp=_c(temp, np.float64), natom=_c(natom, np.float64),
                vol=_c(vol, np.float64), mass=_c(mass, np.float64), nelec=_c(nelec, np.float64),
                natomfr=_c(np.asarray(natomfr).T, np.float64), ion_iz=_c(ion_iz, np.int32),
                ion_ii=_c(ion_ii, np.int32), oc=_c(oc, np.float64), ginv=_c(ginv, np.float64),
                wl0=_c(lines[0], np.float32), gxs=_c(lines[1], np.float32),
                chilw=_c(lines[2], np.float32), liz=_c(lines[3], np.int16),
                lii=_c(lines[4], np.int16), ph1=_c(tables[0], np.float32),
                ph2=_c(tables[1], np.float32), gff=_c(tables[2], np.float64))
    s = _OpacIn(nc, ng, _p(keep["wl"]), _p(keep["temp"]), _p(keep["natom"]), _p(keep["vol"]),
                _p(keep["mass"]), _p(keep["nelec"]), _p(keep["natomfr"]), len(keep["ion_iz"]),
                _p(keep["ion_iz"]), _p(keep["ion_ii"]), _p(keep["oc"]), _p(keep["ginv"]),
                len(keep["wl0"]), _p(keep["wl0"]), _p(keep["gxs"]), _p(keep["chilw"]),
                _p(keep["liz"]), _p(keep["lii"]), _p(keep["ph1"]), _p(keep["ph2"]),
                _p(keep["gff"]), flags)
    cap64 = np.zeros((nc, ng))
    cap32 = np.zeros((nc, ng), np.float32)
    emiss = np.zeros((nc, ng), np.float32)
    sig, capgrey, capross = np.zeros(nc), np.zeros(nc), np.zeros(nc)
    lib().or_opacity(ctypes.byref(s), _p(cap64), _p(cap32), _p(sig), _p(capgrey),
                     _p(capross), _p(emiss))
    return dict(cap64=cap64, cap=cap32, sig=sig, capgrey=capgrey, capross=capross, emiss=emiss)


class _NlteIn(ctypes.Structure):
    _fields_ = [("nlev", _i), ("nlin", _i), ("ncoll", _i), ("glev", _vp), ("llw", _vp),
                ("lup", _vp), ("wl0", _vp), ("f", _vp), ("gl", _vp), ("gu", _vp), ("aul", _vp),
                ("bul", _vp), ("blu", _vp), ("llw_coll", _vp), ("lup_coll", _vp),
                ("coll_c", _vp), ("dele", _vp), ("lev_pi", _vp), ("pi_c", _vp), ("rr_c", _vp),
                ("dele_pi", _vp), ("radtemp", _vp), ("ndens", _vp)]


def _opac_struct(wl, temp, natom, vol, mass, nelec, natomfr, ion_iz, ion_ii, oc, ginv,
                 lines, tables, flags):
    nc, ng = len(temp), len(wl) - 1
    keep = dict(wl=_c(wl, np.float64), temp=_c(temp, np.float64), natom=_c(natom, np.float64),
                vol=_c(vol, np.float64), mass=_c(mass, np.float64), nelec=_c(nelec, np.float64),
                natomfr=_c(np.asarray(natomfr).T, np.float64), ion_iz=_c(ion_iz, np.int32),
                ion_ii=_c(ion_ii, np.int32), oc=_c(oc, np.float64), ginv=_c(ginv, np.float64),
                wl0=_c(lines[0], np.float32), gxs=_c(lines[1], np.float32),
                chilw=_c(lines[2], np.float32), liz=_c(lines[3], np.int16),
                lii=_c(lines[4], np.int16), ph1=_c(tables[0], np.float32),
                ph2=_c(tables[1], np.float32), gff=_c(tables[2], np.float64))
    s = _OpacIn(nc, ng, _p(keep["wl"]), _p(keep["temp"]), _p(keep["natom"]), _p(keep["vol"]),
                _p(keep["mass"]), _p(keep["nelec"]), _p(keep["natomfr"]), len(keep["ion_iz"]),
                _p(keep["ion_iz"]), _p(keep["ion_ii"]), _p(keep["oc"]), _p(keep["ginv"]),
                len(keep["wl0"]), _p(keep["wl0"]), _p(keep["gxs"]), _p(keep["chilw"]),
                _p(keep["liz"]), _p(keep["lii"]), _p(keep["ph1"]), _p(keep["ph2"]),
                _p(keep["gff"]), flags)
    return s, keep


def _nlte_struct(sp, radtemp, ndens):
    """sp: superlite_amd.io.nlte_species(...); ndens (nc, 2) H I and H II densities"""
    keep = dict(glev=_c(sp["glev"], np.float64), llw=_c(sp["llw"], np.int32), lup=_c(sp["lup"], np.int32),
                wl0=_c(sp["wl0"], np.float64), f=_c(sp["f"], np.float64), gl=_c(sp["gl"], np.float64),
                gu=_c(sp["gu"], np.float64), aul=_c(sp["aul"], np.float64), bul=_c(sp["bul"], np.float64),
                blu=_c(sp["blu"], np.float64), llw_coll=_c(sp["llw_coll"], np.int32),
                lup_coll=_c(sp["lup_coll"], np.int32), coll_c=_c(sp["coll_c"], np.float64),
                dele=_c(sp["dele"], np.float64), lev_pi=_c(sp["lev_pi"], np.int32),
                pi_c=_c(sp["pi_c"], np.float64), rr_c=_c(sp["rr_c"], np.float64),
                dele_pi=_c(sp["dele_pi"], np.float64), radtemp=_c(radtemp, np.float64),
                ndens=_c(ndens, np.float64))
    n = _NlteIn(int(sp["nlev"]), len(keep["llw"]), len(keep["llw_coll"]),
                *[_p(keep[k]) for k in ("glev", "llw", "lup", "wl0", "f", "gl", "gu", "aul", "bul",
                                        "blu", "llw_coll", "lup_coll", "coll_c", "dele", "lev_pi",
                                        "pi_c", "rr_c", "dele_pi", "radtemp", "ndens")])
    return n, keep


def opacity_nlte(sp, radtemp, ndens, wl, temp, natom, vol, mass, nelec, natomfr, ion_iz,
                 ion_ii, oc, ginv, lines, tables, flags=0):
    """physical_opacity_nlte (H I in NLTE) + means.  sp: io.nlte_species; ndens (nc, 2);
    lines: the LTE line list without H I lines.  Returns cap64, cap, capemit, emiss (nc, ng),
    sig, capgrey, capemitgrey, capross (nc), npop (nc, nlev) and the return code."""
    nc, ng = len(temp), len(wl) - 1
    s, k1 = _opac_struct(wl, temp, natom, vol, mass, nelec, natomfr, ion_iz, ion_ii, oc, ginv,
                         lines, tables, flags)
    n, k2 = _nlte_struct(sp, radtemp, ndens)
    out = dict(cap64=np.zeros((nc, ng)), cap=np.zeros((nc, ng), np.float32),
               capemit=np.zeros((nc, ng), np.float32), emiss=np.zeros((nc, ng), np.float32),
               sig=np.zeros(nc), capgrey=np.zeros(nc), capemitgrey=np.zeros(nc),
               capross=np.zeros(nc), npop=np.zeros((nc, int(sp["nlev"]))))
    rc = lib().or_opacity_nlte(ctypes.byref(s), ctypes.byref(n), _p(out["cap64"]), _p(out["cap"]),
                               _p(out["capemit"]), _p(out["sig"]), _p(out["capgrey"]),
                               _p(out["capemitgrey"]), _p(out["capross"]), _p(out["emiss"]),
                               _p(out["npop"]))
    out["rc"] = rc
    return out


def nlte_rate_matrix(sp, radtemp, ndens, cell, wl, temp, natom, vol, mass, nelec, natomfr,
                     ion_iz, ion_ii, oc, ginv, lines, tables):
    """the rate matrix (nlev, nlev) and right-hand side of one cell (0-based)"""
    s, k1 = _opac_struct(wl, temp, natom, vol, mass, nelec, natomfr, ion_iz, ion_ii, oc, ginv,
                         lines, tables, 0)
    n, k2 = _nlte_struct(sp, radtemp, ndens)
    nl = int(sp["nlev"])
    a = np.zeros((nl, nl), order="F")
    b = np.zeros(nl)
    lib().or_nlte_rate_matrix(ctypes.byref(s), ctypes.byref(n), int(cell), _p(a), _p(b))
    return a, b


def lu_solve(a, b):
    """dgetf2 + dgetrs restatement; a (n, n), b (n).  Returns x and the info code."""
    n = len(b)
    af = np.asfortranarray(np.array(a, np.float64))
    x = np.array(b, np.float64, copy=True)
    ipiv = np.zeros(n, np.int32)
    info = lib().or_lu_solve(n, _p(af), _p(ipiv), _p(x))
    return x, info


class _EosData(ctypes.Structure):
    _fields_ = [("ni", _vp), ("e", _vp), ("q", _vp), ("g1", _vp), ("amass", _vp)]


def eos(ed, temp, mass, vol, massfr, ion_iz, ion_ii, nelec=None, acc=0.0):
    """massfr2natomfr + eos_update + ions_solve_eos (LTE) on every massive cell.
    ed: dict ni (30), e/q/g1 (30, 31), amass (30); massfr (30, nc).  nelec: start values
    (1.0, gasmod.f:101, if None).  Returns natom, natomfr (30, nc), nelec, oc/ginv (nc, nion),
    niter and the return code."""
    nc = len(temp)
    keep = dict(ni=_c(ed["ni"], np.int32), e=_c(ed["e"], np.float64), q=_c(ed["q"], np.float64),
                g1=_c(ed["g1"], np.float64), amass=_c(ed["amass"], np.float64))
    d = _EosData(*[_p(keep[k]) for k in ("ni", "e", "q", "g1", "amass")])
    ne = np.ones(nc) if nelec is None else np.array(nelec, np.float64, copy=True)
    nion = len(ion_iz)
    natom, natomfr = np.zeros(nc), np.zeros((nc, 30))
    oc, ginv, niter = np.zeros((nc, nion)), np.zeros((nc, nion)), np.zeros(nc, np.int32)
    args = [_c(temp, np.float64), _c(mass, np.float64), _c(vol, np.float64),
            _c(np.asarray(massfr).T, np.float64)]
    iz, ii = _c(ion_iz, np.int32), _c(ion_ii, np.int32)
    rc = lib().or_eos(nc, ctypes.byref(d), *[_p(a) for a in args], float(acc), _p(ne), _p(natom),
                      _p(natomfr), nion, _p(iz), _p(ii), _p(oc), _p(ginv), _p(niter))
    return dict(natom=natom, natomfr=natomfr.T.copy(), nelec=ne, oc=oc, ginv=ginv, niter=niter, rc=rc)


class Model:
    """Per-iteration transport state (grd_* arrays after bcast_nonpermanent)."""

    def __init__(self, xarr, vxarr, divv, vol, wl, nep, nepg, cap, sig, capgrey, tempinv,
                 radtemp, tauddmc=5.0, taulump=10.0, puretran=False, max_steps=0,
                 opaclump=None, emitprob=None, emit=None, nvol=None, ampfact=False,
                 thm_aniso=False, capemit=None, capemitgrey=None):
        """capemit (nc, ng) FP32 and capemitgrey (nc): the NLTE emission opacity (in_nlte);
        None for LTE."""
        self.a = dict(xarr=_c(xarr, np.float64), vxarr=_c(vxarr, np.float64),
                      divv=_c(divv, np.float64), vol=_c(vol, np.float64),
                      wl=_c(wl, np.float64), wlinv=_c(1.0 / np.asarray(wl, np.float64), np.float64),
                      cap=_c(cap, np.float32), sig=_c(sig, np.float64),
                      capgrey=_c(capgrey, np.float64), tempinv=_c(tempinv, np.float64),
                      radtemp=_c(radtemp, np.float64))
        self.nx = len(self.a["divv"])
        self.ng = len(self.a["wl"]) - 1
        self.nep, self.nepg = nep, nepg
        self.params = (tauddmc, taulump, int(puretran), int(max_steps), int(ampfact),
                       int(thm_aniso))
        for k, v in (("opaclump", opaclump), ("emitprob", emitprob), ("emit", emit)):
            self.a[k] = None if v is None else _c(v, np.float64)
        self.a["nvol"] = None if nvol is None else _c(nvol, np.int32)
        if (capemit is None) != (capemitgrey is None):
            raise ValueError("capemit and capemitgrey go together")
        self.a["capemit"] = None if capemit is None else _c(capemit, np.float32)
        self.a["capemitgrey"] = None if capemitgrey is None else _c(capemitgrey, np.float64)

    def set(self, **kw):
        for k, v in kw.items():
            dt = np.int32 if k == "nvol" else (np.float32 if k in ("cap", "capemit") else np.float64)
            self.a[k] = _c(v, dt)

    def struct(self):
        a = self.a
        t = self.params
        return _Model(self.nx, _p(a["xarr"]), _p(a["vxarr"]), _p(a["divv"]), _p(a["vol"]),
                      self.ng, _p(a["wl"]), _p(a["wlinv"]), self.nep, self.nepg, t[0], t[1],
                      t[2], t[3], t[4], t[5], _p(a["cap"]), _p(a["sig"]), _p(a["capgrey"]),
                      _p(a["tempinv"]), _p(a["opaclump"]), _p(a["emitprob"]), _p(a["emit"]),
                      _p(a["radtemp"]), _p(a["nvol"]), _p(a["capemit"]), _p(a["capemitgrey"]))

    def leakage(self):
        out = np.zeros((self.nx, 11))
        s = self.struct()
        lib().or_leakage(ctypes.byref(s), _p(out))
        return out

    def emission_probability(self):
        out = np.zeros((self.nx, self.nep))
        s = self.struct()
        lib().or_emission_probability(ctypes.byref(s), _p(out))
        return out

    def emitgroup(self, r, ic):
        s = self.struct()
        return lib().or_emitgroup(ctypes.byref(s), float(r), int(ic))

    def sources(self, seed, iteration):
        n = int(self.a["nvol"].sum())
        out = [np.zeros(n) for _ in range(6)]
        s = self.struct()
        lib().or_sources(ctypes.byref(s), seed, iteration, *[_p(o) for o in out])
        return dict(zip(("x", "mu", "om", "e", "e0", "wl"), out))

    def iteration(self, seed, iteration, per_particle=False):
        nc, ng = self.nx, self.ng
        n = int(self.a["nvol"].sum())
        out = dict(tally=np.zeros(nc), jrad=np.zeros((nc, ng)), eamp=np.zeros(nc), luminos=np.zeros(ng),
                   lumdev=np.zeros(ng), lumnum=np.zeros(ng, np.int32))
        pp = {}
        if per_particle:
            pp = dict(stat=np.zeros(n, np.uint8), x=np.zeros(n), mu=np.zeros(n), wl=np.zeros(n),
                      e=np.zeros(n), ig=np.zeros(n, np.int32), ix=np.zeros(n, np.int32),
                      nstep=np.zeros(n, np.int64), err=np.zeros(n, np.int32))
        r = _Result(_p(out["tally"]), _p(out["jrad"]), _p(out["eamp"]), _p(out["luminos"]),
                    _p(out["lumdev"]), _p(out["lumnum"]), (_d * 7)(), (ctypes.c_int64 * 10)(),
                    *[_p(pp.get(k)) for k in ("stat", "x", "mu", "wl", "e", "ig", "ix", "nstep", "err")])
        s = self.struct()
        rc = lib().or_iteration(ctypes.byref(s), seed, iteration, ctypes.byref(r))
        tot = list(r.tot)
        ct = list(r.ct)
        out.update(eout=tot[0], evelo=tot[1], sflux=tot[2], estart=tot[3], edead=tot[4],
                   esrc=tot[5], eampsum=tot[6], nactive=ct[0], nstepimc=ct[1], nstepddmc=ct[2], nstepmax=ct[3],
                   nmethswap=ct[4], nflux=ct[5], ndead=ct[6], nerr=ct[7], nwarn=ct[8],
                   nnewton=ct[9], rc=rc)
        if per_particle:
            out["particles"] = pp
        return out


def nlte_rad_est(wl, vol, tally, radtemp):
    """nltemod.f:231-260: the updated radtemp (nc) and jrad (nc, ng)"""
    wl, vol, tally = _c(wl, np.float64), _c(vol, np.float64), _c(tally, np.float64)
    rt = np.array(radtemp, np.float64)
    nc, ng = len(vol), len(wl) - 1
    jrad = np.zeros((nc, ng))
    lib().or_nlte_rad_est(nc, ng, _p(wl), _p(vol), _p(tally), _p(rt), _p(jrad))
    return rt, jrad


def sourceenergy(rcell, radtemp, lbol):
    rc = _c(rcell, np.float64)
    rt = _c(radtemp, np.float64)
    emit = np.zeros(len(rc))
    st = _d()
    lib().or_sourceenergy(len(rc), _p(rc), _p(rt), float(lbol), _p(emit), ctypes.byref(st))
    return emit, st.value


def sourcenumbers(emit, nstot):
    e = _c(emit, np.float64)
    nvol = np.zeros(len(e), np.int32)
    lib().or_sourcenumbers(len(e), _p(e), int(nstot), _p(nvol))
    return nvol
