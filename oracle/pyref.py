"""ctypes access to oracle/_ref/libslref.so: SuperLite's own Fortran hot path compiled
from the reference sources by oracle/Makefile (serial MPI stub), driven through
oracle/ref_harness.f90.

Test infrastructure: imported only by tests/ and tests/golden/make_golden.py, to pin
the C restatement (oracle/oracle.c) against the reference itself.

Fortran module state is process-global and allocated once, so every RefSession loads
its own private copy of the shared object (distinct path => distinct static data).
Arrays follow Fortran column-major order; a Fortran array a(n1, n2) is a numpy array of
shape (n2, n1) in C order here.
"""
from __future__ import annotations

import ctypes
import os
import shutil
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "libslref.so")
REF_DATA = "/root/reference/Data"

_d = ctypes.c_double
_i = ctypes.c_int


def available() -> bool:
    return os.path.exists(LIB)


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class RefSession:
    """One independent instance of the reference's module state."""

    def __init__(self, datadir: str = REF_DATA):
        if not available():
            raise RuntimeError("oracle/_ref/libslref.so not built (make -C oracle ref)")
        self._tmp = tempfile.mkdtemp(prefix="slref_")
        path = os.path.join(self._tmp, "libslref.so")
        shutil.copy(LIB, path)
        self.lib = ctypes.CDLL(path, mode=os.RTLD_LOCAL)
        self.lib.slref_emitgroup.restype = _i
        self.lib.slref_emitgroup.argtypes = [_d, _i]
        self.lib.slref_binsrch.restype = _i
        self.lib.slref_binsrch.argtypes = [_d, ctypes.c_void_p, _i, _i]
        self.lib.slref_bfxs.restype = _d
        self.lib.slref_bfxs.argtypes = [_i, _i, _d]
        self.lib.slref_specintv.argtypes = [_d, _i, ctypes.c_void_p, _i]
        self.lib.slref_spectintv.argtypes = [_d, _i, ctypes.c_void_p]
        cwd = os.getcwd()
        try:
            os.chdir(datadir)   # the reference's readers open data files by relative name
            self.lib.slref_read_tables()
        finally:
            os.chdir(cwd)
        self.nc = self.ng = self.nep = 0

    def close(self):
        shutil.rmtree(self._tmp, ignore_errors=True)

    # -- tables -----------------------------------------------------------
    def tables(self):
        ph1 = np.zeros((30, 30, 7, 6), np.float32)
        ph2 = np.zeros((30, 30, 7), np.float32)
        gff = np.zeros((41, 81), np.float64)
        self.lib.slref_get_bf(_p(ph1), _p(ph2))
        self.lib.slref_get_ff(_p(gff))
        return ph1, ph2, gff

    def ni(self):
        ni = np.zeros(30, np.int32)
        self.lib.slref_get_ni(_p(ni))
        return ni

    # -- setup ------------------------------------------------------------
    def setup(self, nx, ng, wlmin, wlmax, xleft, vxleft, n2s=10, nomp=1,
              tauddmc=5.0, taulump=10.0, puretran=False, errorfatal=True, lbol=1e43,
              ampfact=False, thm_aniso=False):
        xl = np.ascontiguousarray(xleft, np.float64)
        vl = np.ascontiguousarray(vxleft, np.float64)
        self.lib.slref_setup(_i(nx), _i(ng), _d(wlmin), _d(wlmax), _p(xl), _p(vl),
                             _i(n2s), _i(nomp), _d(tauddmc), _d(taulump),
                             _i(int(puretran)), _i(int(errorfatal)), _d(lbol),
                             _i(int(ampfact)), _i(int(thm_aniso)))
        self.nc, self.ng = nx, ng
        vol = np.zeros(nx)
        rcell = np.zeros(nx)
        divv = np.zeros(nx)
        nep = _i()
        nepg = _i()
        self.lib.slref_get_grid(_p(vol), _p(rcell), _p(divv), ctypes.byref(nep),
                                ctypes.byref(nepg))
        self.nep = nep.value
        return dict(vol=vol, rcell=rcell, divv=divv, nep=nep.value, nepg=nepg.value)

    def put_groups(self, wl):
        wl = np.ascontiguousarray(wl, np.float64)
        self.lib.slref_put_groups(_i(len(wl) - 1), _p(wl))

    def set_gas(self, mass, temp, massfr):
        m = np.ascontiguousarray(mass, np.float64)
        t = np.ascontiguousarray(temp, np.float64)
        f = np.ascontiguousarray(massfr.T, np.float64)  # (nc, 30) C == (30, nc) Fortran
        self.lib.slref_set_gas(_i(self.nc), _p(m), _p(t), _p(f))

    def eos(self):
        self.lib.slref_eos()

    def get_gas(self):
        nc = self.nc
        natom = np.zeros(nc)
        natomfr = np.zeros((nc, 30))
        nelec = np.zeros(nc)
        vol = np.zeros(nc)
        self.lib.slref_get_gas(_i(nc), _p(natom), _p(natomfr), _p(nelec), _p(vol))
        return natom, natomfr.T.copy(), nelec, vol

    def read_par(self, deckdir):
        """the reference's namelist reader on deckdir/input.par: the parameters the hot path uses"""
        ndim, fl = np.zeros(3, np.int32), np.zeros(3, np.int32)
        iv, rv, lv = np.zeros(6, np.int32), np.zeros(9), np.zeros(10, np.int32)
        cwd = os.getcwd()
        try:
            os.chdir(deckdir)
            self.lib.slref_read_par(_p(ndim), _p(fl), _p(iv), _p(rv), _p(lv))
        finally:
            os.chdir(cwd)
        return dict(ndim=ndim, flx_ndim=fl, ivals=iv, rvals=rv, lvals=lv)

    def read_str(self, deckdir, nx):
        """the reference's input.str reader (1D spherical) on deckdir/input.str"""
        nab = _i()
        flags = np.zeros(4, np.int32)
        cwd = os.getcwd()
        try:
            os.chdir(deckdir)
            self.lib.slref_read_str(_i(nx), ctypes.byref(nab), _p(flags))
        finally:
            os.chdir(cwd)
        na = nab.value
        out = dict(xleft=np.zeros(nx + 1), vxleft=np.zeros(nx + 1), mass=np.zeros(nx),
                   temp=np.zeros(nx), massfr=np.zeros((nx, na)), iabund=np.zeros(na, np.int32),
                   vol=np.zeros(nx), radtemp=np.zeros(nx), lum=np.zeros(nx), ye=np.zeros(nx))
        self.lib.slref_get_str(_i(nx), _i(na), *[_p(out[k]) for k in
                                                ("xleft", "vxleft", "mass", "temp", "massfr", "iabund",
                                                 "vol", "radtemp", "lum", "ye")])
        out["flags"] = flags
        return out

    def write_outputs(self, outdir, a, opacdump=True):
        """the reference's output_flux + output_grid writers, one iteration, into outdir"""
        nc, ng = self.nc, self.ng
        keys = ("luminos", "lumdev", "lumnum", "nvol", "tempinv", "radtemp", "capgrey", "capross",
                "sig", "tally", "vol", "tot")
        arrs = [np.ascontiguousarray(a[k], np.int32 if k in ("lumnum", "nvol") else np.float64)
                for k in keys]
        cwd = os.getcwd()
        try:
            os.chdir(outdir)
            self.lib.slref_write_outputs(_i(nc), _i(ng), *[_p(x) for x in arrs], _i(int(opacdump)))
        finally:
            os.chdir(cwd)

    def eos_data(self):
        """ni(30), e/q/g1 (30, 31) per element and stage, element masses (30) [amu]"""
        ni = np.zeros(30, np.int32)
        e, q, g1 = np.zeros((30, 31)), np.zeros((30, 31)), np.zeros((30, 31))
        am = np.zeros(30)
        self.lib.slref_get_eos_data(_p(ni), _p(e), _p(q), _p(g1), _p(am))
        return dict(ni=ni, e=e, q=q, g1=g1, amass=am)

    def get_pops(self, ion_iz, ion_ii):
        iz = np.ascontiguousarray(ion_iz, np.int32)
        ii = np.ascontiguousarray(ion_ii, np.int32)
        nion = len(iz)
        oc = np.zeros((self.nc, nion))
        ginv = np.zeros((self.nc, nion))
        self.lib.slref_get_pops(_i(self.nc), _i(nion), _p(iz), _p(ii), _p(oc), _p(ginv))
        return oc, ginv

    def put_gas(self, mass, temp, natom, natomfr, nelec, ion_iz, ion_ii, oc, ginv):
        a = [np.ascontiguousarray(v, np.float64) for v in (mass, temp, natom)]
        fr = np.ascontiguousarray(np.asarray(natomfr).T, np.float64)
        ne = np.ascontiguousarray(nelec, np.float64)
        iz = np.ascontiguousarray(ion_iz, np.int32)
        ii = np.ascontiguousarray(ion_ii, np.int32)
        o = np.ascontiguousarray(oc, np.float64)
        g = np.ascontiguousarray(ginv, np.float64)
        self.lib.slref_put_gas(_i(self.nc), _p(a[0]), _p(a[1]), _p(a[2]), _p(fr), _p(ne),
                               _i(len(iz)), _p(iz), _p(ii), _p(o), _p(g))

    def set_lines(self, wl0, gxs, chilw, iz, ii):
        arrs = [np.ascontiguousarray(wl0, np.float32), np.ascontiguousarray(gxs, np.float32),
                np.ascontiguousarray(chilw, np.float32), np.ascontiguousarray(iz, np.int16),
                np.ascontiguousarray(ii, np.int16)]
        self.lib.slref_set_lines(_i(len(arrs[0])), *[_p(a) for a in arrs])

    # -- hot path ---------------------------------------------------------
    def opacity(self, flags=0):
        self.lib.slref_opacity(_i(flags))
        nc, ng = self.nc, self.ng
        cap = np.zeros((nc, ng), np.float32)
        emiss = np.zeros((nc, ng), np.float32)
        sig, capgrey, capross = np.zeros(nc), np.zeros(nc), np.zeros(nc)
        self.lib.slref_get_opacity(_i(nc), _i(ng), _p(cap), _p(sig), _p(capgrey),
                                   _p(capross), _p(emiss))
        return dict(cap=cap, sig=sig, capgrey=capgrey, capross=capross, emiss=emiss)

    def put_opacity(self, cap, sig, capgrey, capross):
        c = np.ascontiguousarray(cap, np.float32)
        s, g, r = (np.ascontiguousarray(v, np.float64) for v in (sig, capgrey, capross))
        self.lib.slref_put_opacity(_i(self.nc), _i(self.ng), _p(c), _p(s), _p(g), _p(r))

    def prep(self):
        self.lib.slref_prep()
        opaclump = np.zeros((self.nc, 11))
        emitprob = np.zeros((self.nc, self.nep))
        emit = np.zeros(self.nc)
        self.lib.slref_get_prep(_i(self.nc), _i(self.nep), _p(opaclump), _p(emitprob), _p(emit))
        return dict(opaclump=opaclump, emitprob=emitprob, emit=emit)

    def set_nlte(self, capemit=None, capemitgrey=None):
        """in_nlte with the emission opacity capemit (nc, ng) FP32 and capemitgrey (nc) for the
        following prep and transport; no arguments: back to LTE"""
        on = capemit is not None
        c = np.ascontiguousarray(capemit if on else np.zeros((self.nc, self.ng)), np.float32)
        g = np.ascontiguousarray(capemitgrey if on else np.zeros(self.nc), np.float64)
        self.lib.slref_set_nlte(_i(int(on)), _i(self.nc), _i(self.ng), _p(c), _p(g))

    def nlte_rad_est(self, tally, radtemp):
        t = np.ascontiguousarray(tally, np.float64)
        rt = np.array(radtemp, np.float64)
        jrad = np.zeros((self.nc, self.ng))
        self.lib.slref_nlte_rad_est(_i(self.nc), _i(self.ng), _p(t), _p(rt), _p(jrad))
        return rt, jrad

    def emitgroup(self, r, ic):
        return self.lib.slref_emitgroup(_d(r), _i(ic))

    def sourcenumbers(self, emit, ns):
        """grd_nvol and src_nnew of the reference's sourcenumbers for grd_emit = emit"""
        e = np.ascontiguousarray(emit, np.float64)
        nvol = np.zeros(self.nc, np.int32)
        nnew = _i()
        self.lib.slref_sourcenumbers(_i(self.nc), _p(e), _i(int(ns)), _p(nvol), ctypes.byref(nnew))
        return nvol, nnew.value

    def binsrch(self, x, arr, widerange):
        a = np.ascontiguousarray(arr, np.float64)
        return self.lib.slref_binsrch(_d(x), _p(a), _i(len(a)), _i(int(widerange)))

    def bfxs(self, nz, ne, e):
        return self.lib.slref_bfxs(_i(nz), _i(ne), _d(e))

    def specintv(self, tempinv, n, offset=1):
        ss = np.zeros(n)
        self.lib.slref_specintv(_d(tempinv), _i(n), _p(ss), _i(offset))
        return ss

    def spectintv(self, tempinv, n):
        ss = np.zeros(n)
        self.lib.slref_spectintv(_d(tempinv), _i(n), _p(ss))
        return ss

    def iteration(self):
        nc, ng = self.nc, self.ng
        tally = np.zeros(nc)
        jrad = np.zeros((nc, ng))
        eamp = np.zeros(nc)
        lum = np.zeros(ng)
        dev = np.zeros(ng)
        num = np.zeros(ng, np.int32)
        tot = np.zeros(4)
        ct = np.zeros(8, np.int64)
        self.lib.slref_iteration(_i(nc), _i(ng), _p(tally), _p(jrad), _p(eamp), _p(lum),
                                 _p(dev), _p(num), _p(tot), _p(ct))
        return dict(tally=tally, jrad=jrad, eamp=eamp, luminos=lum, lumdev=dev, lumnum=num,
                    eout=tot[0], evelo=tot[1], sflux=tot[2], sthermal=tot[3],
                    nnew=int(ct[0]), nactive=int(ct[1]), nstepimc=int(ct[2]),
                    nstepddmc=int(ct[3]), nstepmax=int(ct[4]), nmethswap=int(ct[5]),
                    nflux=int(ct[6]))
