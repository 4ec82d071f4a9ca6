"""ctypes binding of libslgpu.so (include/slgpu.h) and a host mirror of SuperLite's
per-iteration call sequence (superlite.f90:148-219).

The product path is the HIP library; this module only marshals host arrays.  Loading
fails loudly if libslgpu.so has not been built (python superlite_amd/build.py).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libslgpu.so")

SLGPU_FP64, SLGPU_MIXED = 0, 1
NOBB, NOBF, NOFF, NOTHOMSON = 1, 2, 4, 8
SLGPU_EINVAL_CODE, E_TRANSPORT = 1, 4

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_i64 = ctypes.c_int64
_u32 = ctypes.c_uint32

EXPORTS = ("slgpu_create", "slgpu_destroy", "slgpu_last_error_message", "slgpu_set_grid",
           "slgpu_set_groups", "slgpu_set_lines", "slgpu_set_bf_tables", "slgpu_set_ff_table",
           "slgpu_set_ions", "slgpu_opacity_build", "slgpu_ddmc_prep", "slgpu_sources",
           "slgpu_transport", "slgpu_fluxtally", "slgpu_get_particles", "slgpu_sample_groups",
           "slgpu_opacity_vjp", "slgpu_timers", "slgpu_stream", "slgpu_peak_rates",
           "slgpu_opacity_build_noalloc", "slgpu_set_eos_data", "slgpu_eos", "slgpu_eos_vjp",
           "slgpu_set_nlte_data", "slgpu_opacity_build_nlte", "slgpu_nlte_rad_est",
           "slgpu_eos_nlte_ndens", "slgpu_opacity_outputs_dev", "slgpu_opacity_vjp_dev",
           "slgpu_set_unstaged")


class SlgpuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"slgpu error {code}: {msg}")
        self.code = code


class Config(ctypes.Structure):
    _fields_ = [("nx", _i), ("ng", _i), ("nep", _i), ("nepg", _i), ("tauddmc", _d),
                ("taulump", _d), ("puretran", _i), ("errorfatal", _i), ("precision", _i),
                ("seed", _u32), ("npartmax", _i64), ("max_steps", _i64), ("ampfact", _i),
                ("thm_aniso", _i)]


class Gas(ctypes.Structure):
    _fields_ = [("temp", _vp), ("natom", _vp), ("vol", _vp), ("mass", _vp), ("nelec", _vp),
                ("natomfr", _vp), ("oc", _vp), ("ginv", _vp)]


_lib = None


def _one_hip_runtime():
    """torch (ROCm) ships its own libamdhip64 under another file name; loaded after the
    system one that libslgpu.so names, it would be a second HIP runtime in the process and
    would find no GPU.  Importing torch first lets libslgpu.so bind to torch's runtime (same
    soname), so device tensors and the library share one runtime."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load libslgpu.so (raises if it is missing: there is no fallback path)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built; run python superlite_amd/build.py")
        _one_hip_runtime()
        L = ctypes.CDLL(LIB)
        L.slgpu_last_error_message.restype = ctypes.c_char_p
        L.slgpu_create.argtypes = [ctypes.POINTER(Config), ctypes.POINTER(_vp)]
        L.slgpu_destroy.argtypes = [_vp]
        L.slgpu_set_grid.argtypes = [_vp] + [_vp] * 5
        L.slgpu_set_groups.argtypes = [_vp, _vp]
        L.slgpu_set_lines.argtypes = [_vp, _i64] + [_vp] * 5
        L.slgpu_set_bf_tables.argtypes = [_vp, _vp, _vp]
        L.slgpu_set_ff_table.argtypes = [_vp, _vp]
        L.slgpu_set_ions.argtypes = [_vp, _i, _vp, _vp]
        L.slgpu_opacity_build.argtypes = [_vp, ctypes.POINTER(Gas), _i] + [_vp] * 6
        L.slgpu_opacity_build_noalloc.argtypes = [_vp, ctypes.POINTER(Gas), _i] + [_vp] * 6
        L.slgpu_set_nlte_data.argtypes = [_vp, _i, _i, _i] + [_vp] * 18
        L.slgpu_opacity_build_nlte.argtypes = [_vp, ctypes.POINTER(Gas), _vp, _vp, _i] + [_vp] * 9
        L.slgpu_nlte_rad_est.argtypes = [_vp, _vp, _vp]
        L.slgpu_eos_nlte_ndens.argtypes = [_vp, _vp]
        L.slgpu_opacity_outputs_dev.argtypes = [_vp] * 4
        L.slgpu_opacity_vjp_dev.argtypes = [_vp] * 9
        L.slgpu_ddmc_prep.argtypes = [_vp, _vp, _d, _vp, _vp, _vp, _vp]
        L.slgpu_sources.argtypes = [_vp, _i64, _u32, _vp, _vp]
        L.slgpu_transport.argtypes = [_vp, _u32, _vp, _vp, _vp, _vp, _vp]
        L.slgpu_fluxtally.argtypes = [_vp, _vp, _vp, _vp]
        L.slgpu_get_particles.argtypes = [_vp, _i64] + [_vp] * 8
        L.slgpu_sample_groups.argtypes = [_vp, _i, _i, _i64, _vp, _vp, _vp]
        L.slgpu_set_unstaged.argtypes = [_vp, _i]
        L.slgpu_opacity_vjp.argtypes = [_vp] + [_vp] * 8
        L.slgpu_timers.argtypes = [_vp, _vp, _i, _i]
        L.slgpu_stream.restype = _vp
        L.slgpu_stream.argtypes = [_vp]
        L.slgpu_peak_rates.argtypes = [_vp, _vp]
        L.slgpu_set_eos_data.argtypes = [_vp] * 6
        L.slgpu_eos.argtypes = [_vp, _vp, _vp, _vp, _d] + [_vp] * 6
        L.slgpu_eos_vjp.argtypes = [_vp] * 8
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(_vp)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class Context:
    """One slgpu context: device-resident mesh, groups, lines, opacities, particles."""

    def __init__(self, nx, ng, nep, nepg, tauddmc=5.0, taulump=10.0, puretran=False,
                 errorfatal=True, precision=SLGPU_FP64, seed=1, npartmax=1 << 20, max_steps=0,
                 ampfact=False, thm_aniso=False):
        self.L = lib()
        self.nx, self.ng, self.nep, self.nepg = nx, ng, nep, nepg
        cfg = Config(nx, ng, nep, nepg, tauddmc, taulump, int(puretran), int(errorfatal),
                     precision, seed, npartmax, max_steps, int(ampfact), int(thm_aniso))
        h = _vp()
        self._check(self.L.slgpu_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.npart = 0
        self.builds = 0          # counts opacity builds and EOS solves (state the backward uses)
        self.nlev = 0            # levels of the NLTE atom (slgpu_set_nlte_data)

    def _check(self, rc):
        if rc != 0:
            raise SlgpuError(rc, self.L.slgpu_last_error_message().decode())

    def close(self):
        if self.h:
            self.L.slgpu_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- static inputs ----------------------------------------------------
    def set_grid(self, xarr, vxarr, divv, vol, rcell):
        a = [_c(v, np.float64) for v in (xarr, vxarr, divv, vol, rcell)]
        self._check(self.L.slgpu_set_grid(self.h, *[_p(x) for x in a]))

    def set_groups(self, wl):
        self._check(self.L.slgpu_set_groups(self.h, _p(_c(wl, np.float64))))

    def set_lines(self, wl0, gxs, chilw, iz, ii):
        a = [_c(wl0, np.float32), _c(gxs, np.float32), _c(chilw, np.float32),
             _c(iz, np.int16), _c(ii, np.int16)]
        self._check(self.L.slgpu_set_lines(self.h, len(a[0]), *[_p(x) for x in a]))

    def set_tables(self, ph1, ph2, gff):
        self._check(self.L.slgpu_set_bf_tables(self.h, _p(_c(ph1, np.float32)),
                                               _p(_c(ph2, np.float32))))
        self._check(self.L.slgpu_set_ff_table(self.h, _p(_c(gff, np.float64))))

    def set_ions(self, ion_iz, ion_ii):
        iz, ii = _c(ion_iz, np.int32), _c(ion_ii, np.int32)
        self._check(self.L.slgpu_set_ions(self.h, len(iz), _p(iz), _p(ii)))
        self.nion = len(iz)

    # -- per iteration ----------------------------------------------------
    def opacity_build(self, temp, natom, vol, mass, nelec, natomfr, oc, ginv, flags=0,
                      want=("cap64", "cap", "sig", "capgrey", "capross", "emiss"), noalloc=False):
        """physical_opacity + opacity_planckmean + opacity_rossmean.
        natomfr (30, nc); oc, ginv (nc, nion).  Group-resolved outputs have shape (nc, ng).
        noalloc: through slgpu_opacity_build_noalloc (static scratch, no device allocation)."""
        nc, ng = self.nx, self.ng
        keep = [_c(temp, np.float64), _c(natom, np.float64), _c(vol, np.float64),
                _c(mass, np.float64), _c(nelec, np.float64), _c(np.asarray(natomfr).T, np.float64),
                _c(oc, np.float64), _c(ginv, np.float64)]
        gas = Gas(*[_p(x) for x in keep])
        out = {}
        if "cap64" in want: out["cap64"] = np.zeros((nc, ng))
        if "cap" in want: out["cap"] = np.zeros((nc, ng), np.float32)
        if "emiss" in want: out["emiss"] = np.zeros((nc, ng), np.float32)
        for k in ("sig", "capgrey", "capross"):
            if k in want: out[k] = np.zeros(nc)
        self.builds += 1
        fn = self.L.slgpu_opacity_build_noalloc if noalloc else self.L.slgpu_opacity_build
        self._check(fn(self.h, ctypes.byref(gas), flags, _p(out.get("cap64")), _p(out.get("cap")),
                       _p(out.get("sig")), _p(out.get("capgrey")), _p(out.get("capross")),
                       _p(out.get("emiss"))))
        return out

    def set_nlte_data(self, sp):
        """the H I model atom: a dict as superlite_amd.io.nlte_species returns"""
        keep = dict(glev=_c(sp["glev"], np.float64), llw=_c(sp["llw"], np.int32),
                    lup=_c(sp["lup"], np.int32), wl0=_c(sp["wl0"], np.float64),
                    f=_c(sp["f"], np.float64), gl=_c(sp["gl"], np.float64), gu=_c(sp["gu"], np.float64),
                    aul=_c(sp["aul"], np.float64), bul=_c(sp["bul"], np.float64),
                    blu=_c(sp["blu"], np.float64), cl=_c(sp["llw_coll"], np.int32),
                    cu=_c(sp["lup_coll"], np.int32), cc=_c(sp["coll_c"], np.float64),
                    dele=_c(sp["dele"], np.float64), levpi=_c(sp["lev_pi"], np.int32),
                    pic=_c(sp["pi_c"], np.float64), rrc=_c(sp["rr_c"], np.float64),
                    delepi=_c(sp["dele_pi"], np.float64))
        self.nlev = int(sp["nlev"])
        self._check(self.L.slgpu_set_nlte_data(
            self.h, self.nlev, len(keep["llw"]), len(keep["cl"]),
            *[_p(keep[k]) for k in ("glev", "llw", "lup", "wl0", "f", "gl", "gu", "aul", "bul", "blu",
                                    "cl", "cu", "cc", "dele", "levpi", "pic", "rrc", "delepi")]))

    def opacity_build_nlte(self, temp, natom, vol, mass, nelec, natomfr, oc, ginv, radtemp, ndens,
                           flags=0):
        """physical_opacity_nlte + means: H I in NLTE.  ndens (nc, 2) H I and H II densities.
        Returns cap64, cap, capemit, emiss (nc, ng), sig, capgrey, capemitgrey, capross (nc)
        and npop (nc, nlev)."""
        nc, ng = self.nx, self.ng
        keep = [_c(temp, np.float64), _c(natom, np.float64), _c(vol, np.float64),
                _c(mass, np.float64), _c(nelec, np.float64), _c(np.asarray(natomfr).T, np.float64),
                _c(oc, np.float64), _c(ginv, np.float64)]
        gas = Gas(*[_p(x) for x in keep])
        rt, nd = _c(radtemp, np.float64), _c(ndens, np.float64)
        out = dict(cap64=np.zeros((nc, ng)), cap=np.zeros((nc, ng), np.float32),
                   capemit=np.zeros((nc, ng), np.float32), emiss=np.zeros((nc, ng), np.float32),
                   sig=np.zeros(nc), capgrey=np.zeros(nc), capemitgrey=np.zeros(nc),
                   capross=np.zeros(nc), npop=np.zeros((nc, max(self.nlev, 1))))
        self.builds += 1
        self._check(self.L.slgpu_opacity_build_nlte(
            self.h, ctypes.byref(gas), _p(rt), _p(nd), flags, _p(out["cap64"]), _p(out["cap"]),
            _p(out["capemit"]), _p(out["sig"]), _p(out["capgrey"]), _p(out["capemitgrey"]),
            _p(out["capross"]), _p(out["emiss"]), _p(out["npop"])))
        return out

    def nlte_rad_est(self, radtemp, want_jrad=True):
        """nlte_rad_est after the transport: the updated radtemp (nc) and the group-centre
        intensity jrad (nc, ng), or None without want_jrad."""
        rt = np.array(radtemp, np.float64)
        jrad = np.zeros((self.nx, self.ng)) if want_jrad else None
        self._check(self.L.slgpu_nlte_rad_est(self.h, _p(rt), _p(jrad)))
        return rt, jrad

    def ddmc_prep(self, radtemp, lbol, want=True):
        nc = self.nx
        out = dict(opaclump=np.zeros((nc, 11)), emitprob=np.zeros((nc, self.nep)),
                   emit=np.zeros(nc)) if want else {}
        st = _d()
        self._check(self.L.slgpu_ddmc_prep(self.h, _p(_c(radtemp, np.float64)), float(lbol),
                                           _p(out.get("opaclump")), _p(out.get("emitprob")),
                                           _p(out.get("emit")), ctypes.byref(st)))
        out["sthermal"] = st.value
        return out

    def sources(self, nstot, iteration):
        nvol = np.zeros(self.nx, np.int32)
        nnew = _i64()
        self._check(self.L.slgpu_sources(self.h, int(nstot), iteration, _p(nvol), ctypes.byref(nnew)))
        self.npart = nnew.value
        return nvol, nnew.value

    def transport(self, iteration, want=True, check=True):
        nc, ng = self.nx, self.ng
        tally = np.zeros(nc) if want else None
        jrad = np.zeros((nc, ng)) if want else None
        eamp = np.zeros(nc) if want else None
        tot = np.zeros(7)
        ct = np.zeros(10, np.int64)
        rc = self.L.slgpu_transport(self.h, iteration, _p(tally), _p(jrad), _p(eamp), _p(tot),
                                    _p(ct))
        out = dict(tally=tally, jrad=jrad, eamp=eamp, eout=tot[0], evelo=tot[1], sflux=tot[2],
                   estart=tot[3], edead=tot[4], esrc=tot[5], eampsum=tot[6], nactive=int(ct[0]), nstepimc=int(ct[1]),
                   nstepddmc=int(ct[2]), nstepmax=int(ct[3]), nmethswap=int(ct[4]),
                   nflux=int(ct[5]), ndead=int(ct[6]), nerr=int(ct[7]), nwarn=int(ct[8]), rc=rc)
        if check and rc != 0:
            self._check(rc)
        return out

    def fluxtally(self):
        ng = self.ng
        lum, dev, num = np.zeros(ng), np.zeros(ng), np.zeros(ng, np.int32)
        self._check(self.L.slgpu_fluxtally(self.h, _p(lum), _p(dev), _p(num)))
        return dict(luminos=lum, lumdev=dev, lumnum=num)

    def particles(self, n=None):
        n = self.npart if n is None else n
        pp = dict(stat=np.zeros(n, np.uint8), x=np.zeros(n), mu=np.zeros(n), wl=np.zeros(n),
                  e=np.zeros(n), ig=np.zeros(n, np.int32), nstep=np.zeros(n, np.int64),
                  err=np.zeros(n, np.int32))
        self._check(self.L.slgpu_get_particles(self.h, n, *[_p(pp[k]) for k in
                                                            ("stat", "x", "mu", "wl", "e", "ig", "nstep", "err")]))
        return pp

    def opacity_vjp(self, grad_cap64=None, grad_capgrey=None, grad_capross=None):
        """dL/d(T, natom, n_e, natomfr, oc) of the last opacity_build for the given
        dL/d(cap64 (nc, ng), capgrey (nc), capross (nc)); natomfr (30, nc), oc (nc, nion)."""
        nc, ng = self.nx, self.ng
        keep = [None if g is None else _c(g, np.float64) for g in (grad_cap64, grad_capgrey, grad_capross)]
        out = dict(temp=np.zeros(nc), natom=np.zeros(nc), nelec=np.zeros(nc),
                   natomfr=np.zeros((nc, 30)), oc=np.zeros((nc, max(self.nion, 1))))
        self._check(self.L.slgpu_opacity_vjp(self.h, *[_p(k) for k in keep],
                                             *[_p(out[k]) for k in ("temp", "natom", "nelec", "natomfr", "oc")]))
        out["natomfr"] = out["natomfr"].T.copy()
        out["oc"] = out["oc"][:, : self.nion]
        return out

    def opacity_vjp_dev(self, p_grad_cap64=0, p_grad_capgrey=0, p_grad_capross=0):
        """opacity_vjp with the input gradients given as device addresses on the context's GPU
        (0 = none; e.g. a contiguous float64 torch tensor's data_ptr(), written and
        synchronised by the caller)"""
        nc = self.nx
        out = dict(temp=np.zeros(nc), natom=np.zeros(nc), nelec=np.zeros(nc),
                   natomfr=np.zeros((nc, 30)), oc=np.zeros((nc, max(self.nion, 1))))
        self._check(self.L.slgpu_opacity_vjp_dev(
            self.h, _vp(p_grad_cap64 or None), _vp(p_grad_capgrey or None), _vp(p_grad_capross or None),
            *[_p(out[k]) for k in ("temp", "natom", "nelec", "natomfr", "oc")]))
        out["natomfr"] = out["natomfr"].T.copy()
        out["oc"] = out["oc"][:, : self.nion]
        return out

    def opacity_outputs_dev(self, p_cap64=0, p_capgrey=0, p_capross=0):
        """copy the last build's cap64 (nc, ng), capgrey and capross (nc) to device addresses"""
        self._check(self.L.slgpu_opacity_outputs_dev(self.h, _vp(p_cap64 or None), _vp(p_capgrey or None),
                                                     _vp(p_capross or None)))

    def set_eos_data(self, ed):
        """the host's LTE EOS data: ni (30), e/q/g1 (30, 31), amass (30)"""
        a = [_c(ed["ni"], np.int32), _c(ed["e"], np.float64), _c(ed["q"], np.float64),
             _c(ed["g1"], np.float64), _c(ed["amass"], np.float64)]
        self._check(self.L.slgpu_set_eos_data(self.h, *[_p(x) for x in a]))

    def eos(self, temp, mass, massfr, nelec=None, acc=0.0):
        """massfr2natomfr + eos_update (LTE) on the device.  massfr (30, nc); nelec: start
        values (1.0 if None).  Returns natom, natomfr (30, nc), nelec, oc/ginv (nc, nion), niter."""
        nc = self.nx
        ne = np.ones(nc) if nelec is None else np.array(nelec, np.float64, copy=True)
        out = dict(natom=np.zeros(nc), natomfr=np.zeros((nc, 30)),
                   oc=np.zeros((nc, max(self.nion, 1))), ginv=np.zeros((nc, max(self.nion, 1))),
                   niter=np.zeros(nc, np.int32))
        self.builds += 1
        keep = [_c(temp, np.float64), _c(mass, np.float64), _c(np.asarray(massfr).T, np.float64)]
        self._check(self.L.slgpu_eos(self.h, *[_p(k) for k in keep], float(acc), _p(ne),
                                     *[_p(out[k]) for k in ("natom", "natomfr", "oc", "ginv", "niter")]))
        out["nelec"] = ne
        out["natomfr"] = out["natomfr"].T.copy()
        out["oc"], out["ginv"] = out["oc"][:, : self.nion], out["ginv"][:, : self.nion]
        return out

    def eos_nlte_ndens(self):
        """nlte_ndens of the last eos(): (nc, 2) H I and H II number densities"""
        out = np.zeros((self.nx, 2))
        self._check(self.L.slgpu_eos_nlte_ndens(self.h, _p(out)))
        return out

    def eos_vjp(self, grad_natom=None, grad_natomfr=None, grad_nelec=None, grad_oc=None):
        """dL/d(temp, mass, massfr (30, nc)) of the last eos() for the given dL/d(natom,
        natomfr (30, nc), nelec, oc (nc, nion)); temp is the EOS's part only"""
        nc = self.nx
        g = [None if grad_natom is None else _c(grad_natom, np.float64),
             None if grad_natomfr is None else _c(np.asarray(grad_natomfr).T, np.float64),
             None if grad_nelec is None else _c(grad_nelec, np.float64),
             None if grad_oc is None else _c(grad_oc, np.float64)]
        out = dict(temp=np.zeros(nc), mass=np.zeros(nc), massfr=np.zeros((nc, 30)))
        self._check(self.L.slgpu_eos_vjp(self.h, *[_p(x) for x in g],
                                         *[_p(out[k]) for k in ("temp", "mass", "massfr")]))
        out["massfr"] = out["massfr"].T.copy()
        return out

    def sample_groups(self, table, r, cell, direct=False):
        """group sampled by table (0-3 lumped L/R/D/S, 4 emitgroup) for draws r in cells"""
        r = np.ascontiguousarray(r, np.float64)
        cell = np.ascontiguousarray(cell, np.int32)
        out = np.zeros(len(r), np.int32)
        self._check(self.L.slgpu_sample_groups(self.h, int(table), int(direct), len(r), _p(r),
                                               _p(cell), _p(out)))
        return out

    def set_unstaged(self, on=True):
        """verification hook: take the opacity build's and backward's unstaged kernels"""
        self._check(self.L.slgpu_set_unstaged(self.h, int(bool(on))))

    def timers(self, reset=False):
        ms = np.zeros(12)
        self._check(self.L.slgpu_timers(self.h, _p(ms), 12, int(reset)))
        return dict(opacity=ms[0], prep=ms[1], sources=ms[2], transport=ms[3], fluxtally=ms[4],
                    opacity_kernel=ms[5], transport_kernel=ms[6], opacity_vjp=ms[7], eos=ms[8],
                    eos_vjp=ms[9], opacity_vjp_kernels=ms[10], opacity_kernels=ms[11])


def peak_rates():
    """Measured exponential throughput of the device: (FP32 v_exp_f32 /s, FP64 exp() /s)."""
    L = lib()
    a, b = _d(), _d()
    rc = L.slgpu_peak_rates(ctypes.byref(a), ctypes.byref(b))
    if rc != 0:
        raise SlgpuError(rc, "slgpu_peak_rates failed")
    return a.value, b.value


def setup_workload(w, st, wl, lines, tables, precision=SLGPU_FP64, seed=1, npartmax=None,
                   puretran=False, errorfatal=True, max_steps=0, ampfact=False, thm_aniso=False):
    """Create a context holding one workload's static inputs (superlite.f90:87-131)."""
    npm = npartmax if npartmax is not None else (1 << w.n2s)
    ctx = Context(w.nx, w.ng, int(st["nep"]), int(st["nepg"]), w.tauddmc, w.taulump,
                  puretran=puretran, errorfatal=errorfatal, precision=precision, seed=seed,
                  npartmax=npm, max_steps=max_steps, ampfact=ampfact, thm_aniso=thm_aniso)
    ctx.set_grid(st["xarr"], st["vxarr"], st["divv"], st["vol"], st["rcell"])
    ctx.set_groups(wl)
    ctx.set_ions(st["ion_iz"], st["ion_ii"])
    ctx.set_tables(*tables)
    ctx.set_lines(*lines)
    return ctx


def iteration(ctx, st, nstot, it, lbol, want=True, flags=0):
    """One SuperLite iteration on the device, in the order of superlite.f90:157-219:
    gas_update (opacity + means), sourceenergy/leakage/emission_probability, sources,
    particle_advance, fluxtally.  Returns the tallies."""
    op = ctx.opacity_build(st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"],
                           st["natomfr"], st["oc"], st["ginv"], flags=flags,
                           want=("cap", "sig", "capgrey", "capross") if want else ())
    prep = ctx.ddmc_prep(st["temp"], lbol, want=want)
    nvol, nnew = ctx.sources(nstot, it)
    tr = ctx.transport(it, want=want)
    fl = ctx.fluxtally() if want else ctx.fluxtally()
    return dict(opacity=op, prep=prep, nvol=nvol, nnew=nnew, transport=tr, flux=fl)


def iteration_nlte(ctx, st, nstot, it, lbol, radtemp, ndens, flags=0):
    """One iteration with H I in NLTE (superlite.f90:157-208 with in_nlte): the NLTE opacity
    build, the DDMC prep and samplers on gas_capemit, sources, transport, flux tally, and
    nlte_rad_est.  Returns the tallies and the updated radtemp."""
    op = ctx.opacity_build_nlte(st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"],
                                st["natomfr"], st["oc"], st["ginv"], radtemp, ndens, flags=flags)
    prep = ctx.ddmc_prep(radtemp, lbol)
    nvol, nnew = ctx.sources(nstot, it)
    tr = ctx.transport(it)
    fl = ctx.fluxtally()
    rt, jrad = ctx.nlte_rad_est(radtemp)
    return dict(opacity=op, prep=prep, nvol=nvol, nnew=nnew, transport=tr, flux=fl,
                radtemp=rt, jrad=jrad)
