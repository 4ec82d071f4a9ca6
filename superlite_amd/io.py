"""SuperLite's on-disk formats on the host side of the library.

Readers
  read_input_par   the `&inputpars` namelist of input.par (inputparmod.f:105-139 names and
                   defaults, read by read_inputpars :216-231)
  read_input_str   the 1D-spherical structure file input.str (inputstrmod.f:85-274
                   read_inputstr, elnam2elcode :361-397)
  config_from_pars the slgpu_config fields and opacity flags those parameters select
Writers (one call per iteration, appending like the reference after its first iteration)
  OutputWriter     output.flx_grid/_luminos/_lumnum/_lumdev (OUTPUT/output_flux.f:3-66),
                   output.grd_grid, output.tot_energy and the output.grd_* arrays
                   (OUTPUT/output_grid.f:3-170)

Number formats reproduce the reference's Fortran edit descriptors: `(1p,e12.4)` gives
d.ddddE+xx in 12 columns and drops the E for 3-digit exponents, `(1p,es12.4e3)` keeps three
exponent digits, `i12` right-justifies, and list-directed headers print blank-separated
tokens.  The readers raise ValueError with the reference's stop message where it stops.
"""
from __future__ import annotations

import math
import os
import re

import numpy as np

# inputparmod.f:1-104: the namelist's parameters with their types and defaults
_PAR_DEFAULTS = {
    "in_name": "spn", "in_comment": "", "in_nomp": 1, "in_niter": 1, "in_grd_igeom": 11,
    "in_ndim": [1, 1, 1], "in_grp_ng": -1, "in_grp_ngs": 1, "in_grp_wlmin": 100e-8,
    "in_grp_wlmax": 32000e-8, "in_flx_ndim": [0, 1, 1], "in_flx_wlmin": 1000e-8,
    "in_flx_wlmax": 32000e-8, "in_src_ns": 0, "in_src_n2s": -1, "in_srcepwr": 1.0,
    "in_L_bol": 1.0, "in_L_inner": 0.0, "in_surfsrcmu": "isot", "in_srcmax": 0.0,
    "in_trn_errorfatal": True, "in_trn_tauddmc": 5.0, "in_taulump": 10.0,
    "in_puretran": False, "in_trn_noamp": True, "in_thm_aniso": False, "in_noeos": False,
    "in_opacmixrossel": 0.0, "in_nobbopac": False, "in_nobfopac": False,
    "in_noffopac": False, "in_nothmson": False, "in_nebular": False, "in_R_phot": -1.0,
    "in_nlte": False, "in_nlte_nelem": 1, "in_opacanaltype": "none", "in_ldisp1": 1.0,
    "in_ldisp2": 1.0, "in_gas_sigcoef": 0.0, "in_gas_sigtpwr": 0.0, "in_gas_sigrpwr": 0.0,
    "in_gas_capcoef": 0.0, "in_gas_captpwr": 0.0, "in_gas_caprpwr": 0.0,
    "in_io_grabstdout": False, "in_io_dogrdtally": False, "in_io_nogriddump": False,
    "in_io_opacdump": "off", "in_io_pdensdump": "off", "in_io_profdump": False,
    "in_io_debugging": False,
}
_PAR_KEYS = {k.lower(): k for k in _PAR_DEFAULTS}

# chemical element symbols by atomic number (elemdatamod.f elem_data%sym, Z = 1..111)
ELEMENT_SYMBOLS = (
    "h he li be b c n o f ne na mg al si p s cl ar k ca sc ti v cr mn fe co ni cu zn ga ge as "
    "se br kr rb sr y zr nb mo tc ru rh pd ag cd in sn sb te i xe cs ba la ce pr nd pm sm eu "
    "gd tb dy ho er tm yb lu hf ta w re os ir pt au hg tl pb bi po at rn fr ra ac th pa u np "
    "pu am cm bk cf es fm md no lr rf db sg bh hs mt ds rg").split()


# ---------------------------------------------------------------------------- readers
def _parse_value(tok: str, like):
    t = tok.strip()
    if isinstance(like, bool):
        tl = t.lower().strip(".")
        if tl in ("t", "true"):
            return True
        if tl in ("f", "false"):
            return False
        raise ValueError(f"read_inputpars: ivalid parameters or values in namelist ({tok!r})")
    if isinstance(like, str):
        if len(t) >= 2 and t[0] == t[-1] and t[0] in "'\"":
            return t[1:-1]
        return t
    if isinstance(like, int):
        return int(t)
    return float(t.lower().replace("d", "e"))


def read_input_par(path: str) -> dict:
    """The `&inputpars` namelist of input.par: every parameter with the reference's default
    unless the file sets it; unknown names raise ValueError as read_inputpars stops."""
    text = open(path).read()
    m = re.search(r"&inputpars(.*?)(^\s*/|\n\s*/)", text, re.S | re.M | re.I)
    if not m:
        raise ValueError("read_inputpars: namelist missing or bad in input.par")
    body = "\n".join(line.split("!")[0] for line in m.group(1).splitlines())
    pars = {k: (list(v) if isinstance(v, list) else v) for k, v in _PAR_DEFAULTS.items()}
    # name = value[, value ...] until the next name
    items = re.findall(r"([A-Za-z_][A-Za-z_0-9]*)\s*=\s*(.*?)(?=(?:[A-Za-z_][A-Za-z_0-9]*\s*=)|\Z)",
                       body, re.S)
    for name, raw in items:
        key = _PAR_KEYS.get(name.lower())
        if key is None:
            raise ValueError(f"read_inputpars: ivalid parameters or values in namelist ({name})")
        like = _PAR_DEFAULTS[key]
        if isinstance(like, str):
            toks = [raw.strip().rstrip(",").strip()]
        else:
            toks = [t for t in re.split(r"[,\s]+", raw.strip()) if t]
        if isinstance(like, list):
            vals = [_parse_value(t, like[0]) for t in toks]
            pars[key][: len(vals)] = vals
        else:
            if len(toks) != 1:
                raise ValueError(f"read_inputpars: ivalid parameters or values in namelist ({name})")
            pars[key] = _parse_value(toks[0], like)
    return pars


class _Records:
    """list-directed input over records: each read starts at the next record, takes as many
    blank/comma/tab-separated items as it needs from as many records as it needs, and drops
    the rest of its last record"""

    def __init__(self, lines):
        self.lines = lines
        self.pos = 0

    def skip(self):
        self.pos += 1

    def read(self, n):
        out = []
        while len(out) < n:
            if self.pos >= len(self.lines):
                return None
            out += [t for t in re.split(r"[\s,]+", self.lines[self.pos]) if t]
            self.pos += 1
        return out[:n]


def read_input_str(path: str, ndim) -> dict:
    """The 1D-spherical input.str (read_inputstr with igeom = 11): cell edges xleft and
    vxleft (nx+1), mass, temp (nx), mass fractions (nabund, nx) with their atomic numbers
    iabund, and vol / rad_temp / lum / ye where the file has them."""
    nx, ny, nz = (int(v) for v in ndim)
    with open(path) as f:
        rec = _Records(f.read().splitlines())
    rec.skip()
    hdr = rec.read(6)
    try:
        nx_r, ny_r, nz_r, ncol, nabund = (int(v) for v in hdr[1:6])
    except (TypeError, ValueError):
        raise ValueError("read_inputstr: input.str fmt err: dimensions")
    if nx_r != nx:
        raise ValueError("read_inputstr: incompatible nx dimension")
    if ny_r != ny:
        raise ValueError("read_inputstr: incompatible ny dimension")
    if nz_r != nz:
        raise ValueError("read_inputstr: incompatible nz dimension")
    nvar = ncol - nabund
    words = rec.read(1 + ncol)
    if words is None:
        raise ValueError("read_inputstr: input.str fmt err: col labels")
    labels = [w.lower() for w in words[1:1 + nvar]]
    abund = [w.lower() for w in words[1 + nvar:]]
    ncell = nx * ny * nz
    body = rec.read(ncol * ncell)
    if body is None:
        raise ValueError("read_inputstr: input.str format err: body")
    if rec.read(1) is not None:
        raise ValueError("read_inputstr: input.str body too long")
    try:
        raw = np.array([float(v.lower().replace("d", "e")) for v in body]).reshape(ncell, ncol)
    except ValueError:
        raise ValueError("read_inputstr: input.str format err: body")
    if np.any(raw != raw):
        raise ValueError("read_inputstr: nan in input")
    col = {name: labels.index(name) for name in
           ("x_left", "vx_left", "mass", "vol", "ye", "temp", "rad_temp", "lum") if name in labels}
    if "mass" not in col:
        raise ValueError("read_inputstr: mass label not found")
    if "temp" not in col:
        raise ValueError("read_inputstr: temp label not found")
    xleft, vxleft = np.zeros(nx + 1), np.zeros(nx + 1)
    if "x_left" not in col and "vx_left" not in col and raw[0, 0] != 0.0:
        # only right bounds: the inner boundary is at 0
        xleft[1:], vxleft[1:] = raw[:nx, 0], raw[:nx, 1]
    else:
        if "vx_left" not in col:
            raise ValueError("read_inputstr: input.str fmt err: vx_left column missing")
        jv = col["vx_left"]
        xleft[0], xleft[1:] = raw[0, 0], raw[:nx, 1]
        vxleft[0], vxleft[1:] = raw[0, jv], raw[:nx, jv + 1]
    if np.any(np.diff(xleft) <= 0.0):
        raise ValueError("read_inputstr: x grid not increasing")
    out = dict(xleft=xleft, vxleft=vxleft, mass=raw[:nx, col["mass"]].copy(),
               temp=raw[:nx, col["temp"]].copy())
    if np.any(out["mass"] < 0.0):
        raise ValueError("read_inputstr: mass<0")
    for name, key in (("vol", "vol"), ("rad_temp", "radtemp"), ("lum", "lum"), ("ye", "ye")):
        out[key] = raw[:nx, col[name]].copy() if name in col else None
    massfr = raw[:nx, nvar:].T.copy()
    if np.any(massfr < 0.0):
        raise ValueError("read_inputstr: massfr<0")
    iabund = []
    for lab in abund:
        if lab not in ELEMENT_SYMBOLS:
            raise ValueError("elnam2elcode: no such element found in elemdata")
        iabund.append(ELEMENT_SYMBOLS.index(lab) + 1)
    if len(set(iabund)) != len(iabund):
        raise ValueError("read_inputstr: duplicate abund columns")
    out.update(massfr=massfr, iabund=np.array(iabund, np.int32), labels=abund)
    return out


def massfr30(strc: dict) -> np.ndarray:
    """the structure's mass fractions as the (30, nx) array of elements Z = 1..30 the
    opacity and EOS take (heavier columns are not part of the LTE opacity's element set)"""
    nx = len(strc["mass"])
    mf = np.zeros((30, nx))
    for k, z in enumerate(strc["iabund"]):
        if z <= 30:
            mf[z - 1] = strc["massfr"][k]
    return mf


def config_from_pars(pars: dict) -> dict:
    """slgpu_config fields and opacity build flags the namelist selects"""
    flags = (1 if pars["in_nobbopac"] else 0) | (2 if pars["in_nobfopac"] else 0) | \
        (4 if pars["in_noffopac"] else 0) | (8 if pars["in_nothmson"] else 0)
    ns = pars["in_src_ns"] if pars["in_src_n2s"] < 0 else 2 ** pars["in_src_n2s"]
    return dict(nx=pars["in_ndim"][0], ng=pars["in_grp_ng"], wlmin=pars["in_grp_wlmin"],
                wlmax=pars["in_grp_wlmax"], tauddmc=pars["in_trn_tauddmc"],
                taulump=pars["in_taulump"], puretran=pars["in_puretran"],
                errorfatal=pars["in_trn_errorfatal"], ampfact=not pars["in_trn_noamp"],
                thm_aniso=pars["in_thm_aniso"], flags=flags, nstot=ns, lbol=pars["in_L_bol"],
                niter=pars["in_niter"])


# ---------------------------------------------------------------------------- writers
def fmt_e(x: float, width: int = 12, digits: int = 4, expdigits: int | None = None) -> str:
    """Fortran (1p,Ew.d) / (1p,Ew.dEe): one digit before the point; without Ee a 3-digit
    exponent replaces the E (d.dddd+100), with Ee the exponent has e digits"""
    x = float(x)
    if x != x:
        return "NaN".rjust(width)
    if math.isinf(x):
        return ("-Infinity" if x < 0 else "Infinity").rjust(width)
    s = f"{x:.{digits}E}"
    mant, ex = s.split("E")
    e = int(ex)
    sign = "-" if e < 0 else "+"
    if expdigits is None:
        body = f"{mant}E{sign}{abs(e):02d}" if abs(e) <= 99 else f"{mant}{sign}{abs(e):03d}"
    else:
        body = f"{mant}E{sign}{abs(e):0{expdigits}d}"
    return body.rjust(width)


def _line_e(a, **kw):
    return "".join(fmt_e(v, **kw) for v in a) + "\n"


def _line_i(a):
    return "".join(f"{int(v):12d}" for v in a) + "\n"


def _list_directed(*tokens):
    return " " + " ".join(str(t) for t in tokens) + "\n"


class OutputWriter:
    """The reference's per-iteration output files in `outdir` (output_flux, output_grid):
    the first call writes the grid files and headers and replaces old files, later calls
    append one iteration."""

    def __init__(self, outdir: str, xarr, flx_wl, opacdump: bool = True, nogriddump: bool = False):
        self.outdir = outdir
        self.xarr = np.asarray(xarr, float)
        self.flx_wl = np.asarray(flx_wl, float)
        self.opacdump = opacdump
        self.nogriddump = nogriddump
        self.first = True

    def _write(self, name, text):
        with open(os.path.join(self.outdir, name), "w" if self.first else "a") as f:
            f.write(text)

    def iteration(self, luminos, lumnum, lumdev, nvol, temp, radtemp, capgrey, capross, sig,
                  tally, vol, eout, evelo, sflux, sthermal, capemitgrey=None):
        ng, nx = len(self.flx_wl) - 1, len(self.xarr) - 1
        mask = lambda a: np.where(np.asarray(a, float) > 1e-99, a, 0.0)   # output_flux.f:37-38
        if self.first:
            with open(os.path.join(self.outdir, "output.flx_grid"), "w") as f:
                f.write(_list_directed("#", ng, 1, 1))
                f.write(_line_e(self.flx_wl))
                f.write(_line_e([-1.0, 1.0]))
                f.write(_line_e([0.0, 2.0 * 3.1415926535897932385]))
            with open(os.path.join(self.outdir, "output.grd_grid"), "w") as f:
                f.write(_list_directed("#", 11))
                f.write(_list_directed("#", nx, 1, 1))
                f.write(_list_directed("#", nx, 1, nx))
                f.write(_line_e(self.xarr))
                f.write(_line_e([-1.0, 1.0]))
                f.write(_line_e([0.0, 2.0 * 3.1415926535897932385]))
                f.write(_line_i(np.arange(1, nx + 1)))
        self._write("output.flx_luminos", _line_e(mask(luminos)))
        self._write("output.flx_lumnum", _line_i(lumnum))
        self._write("output.flx_lumdev", _line_e(mask(lumdev)))
        head = ""
        if self.first:
            head = "#    4\n#" + "".join(s.rjust(12) for s in ("eout", "evelo", "sflux", "sthermal")) + "\n"
        self._write("output.tot_energy", head + " " + "".join(fmt_e(v) for v in (eout, evelo, sflux, sthermal)) + "\n")
        if not self.nogriddump:
            self._write("output.grd_nvol", _line_i(nvol))
            self._write("output.grd_temp", _line_e(temp))
            self._write("output.grd_radtemp", _line_e(radtemp))
            if self.opacdump:
                ce = np.zeros(nx) if capemitgrey is None else capemitgrey
                self._write("output.grd_capgrey", _line_e(capgrey, expdigits=3))
                self._write("output.grd_capemitgrey", _line_e(ce, expdigits=3))
                self._write("output.grd_capross", _line_e(capross, expdigits=3))
                self._write("output.grd_sig", _line_e(sig, expdigits=3))
            self._write("output.grd_eraddens", _line_e(np.asarray(tally, float) / np.asarray(vol, float)))
        self.first = False


# ---------------------------------------------------------------------------
# NLTE model atom (nltemod.f:88-166 read_atom_nlte, read_nlte_data.f:148-233)

def _fixed(line: str, widths) -> list:
    """Split one record by Fortran field widths (a TAB is one character, as a formatted
    read counts it); negative widths are skipped columns (nX)."""
    out, pos = [], 0
    for w in widths:
        if w < 0:
            pos += -w
            continue
        out.append(line[pos:pos + w])
        pos += w
    return out


def _num(s: str, kind=float):
    s = s.strip()
    if not s:
        return kind(0)
    return kind(float(s)) if kind is float else kind(s)


def read_atom_nlte(path: str) -> dict:
    """The raw records of a data.nlte.* file: levels '(i5,9x,i5,f11.3,i5)', radiative lines
    '(i4,1x,i4,3x,es11.5,3x,es11.5)', electron-impact lines
    '(i3,1x,i3,4(1x,es13.6),1x,f5.3,1x,es13.6)' and one PI/RR record per level
    '(i3,2(4(1x,es13.6),1x,f5.3),1x,es13.6)' (RR fit first, then PI).  Stops as the
    reference does on a bad header or on data left after the last record."""
    with open(path) as f:
        lines = f.read().split("\n")
    if lines and lines[-1] == "":
        lines.pop()
    head = lines[3].split()
    nlev, nlin, ncoll = int(head[0]), int(head[1]), int(head[2])
    if nlev <= 0 or nlin <= 0:
        raise ValueError("read_atom_nlte: bad level or line count")
    p = 4
    lev = [_fixed(lines[p + k], (5, -9, 5, 11, 5)) for k in range(nlev)]
    p += nlev
    lin = [_fixed(lines[p + k], (4, -1, 4, -3, 11, -3, 11)) for k in range(nlin)]
    p += nlin
    col = [_fixed(lines[p + k], (3, -1, 3, -1, 13, -1, 13, -1, 13, -1, 13, -1, 5, -1, 13))
           for k in range(ncoll)]
    p += ncoll
    pir = [_fixed(lines[p + k], (3,) + (-1, 13, -1, 13, -1, 13, -1, 13, -1, 5) * 2 + (-1, 13))
           for k in range(nlev)]
    p += nlev
    if any(l.strip() for l in lines[p:]):
        raise ValueError("read_atom_nlte: data remaining on input file")
    r = {
        "lev_id": np.array([_num(x[0], int) for x in lev]),
        "lev_g": np.array([_num(x[1], int) for x in lev]),
        "lev_chi": np.array([_num(x[2]) for x in lev]),
        "lev_n": np.array([_num(x[3], int) for x in lev]),
        "lin_lev1": np.array([_num(x[0], int) for x in lin]),
        "lin_lev2": np.array([_num(x[1], int) for x in lin]),
        "lin_f": np.array([_num(x[2]) for x in lin]),
        "lin_wl0": np.array([_num(x[3]) for x in lin]),
        "col_lev1": np.array([_num(x[0], int) for x in col]),
        "col_lev2": np.array([_num(x[1], int) for x in col]),
        "col_c": np.array([[_num(x[k]) for k in range(2, 6)] for x in col]).reshape(-1, 4),
        "col_n": np.array([np.float32(_num(x[6])) for x in col], np.float32),
        "col_dele": np.array([_num(x[7]) for x in col]),
        "pi_lev": np.array([_num(x[0], int) for x in pir]),
        "rr_c": np.array([[_num(x[k]) for k in range(1, 5)] for x in pir]),
        "pi_c": np.array([[_num(x[k]) for k in range(6, 10)] for x in pir]),
        "pi_dele": np.array([_num(x[11]) for x in pir]),
    }
    return r


# physconstmod.f constants used by read_nlte_data.f
_PC_PI = 3.1415926535897932385
_PC_E = 4.80325e-10
_PC_ME = 9.1093826e-28
_PC_C = 2.997924562e10
_PC_H = 6.6260693e-27
_PC_ANG = 1e-8


def nlte_species(raw: dict) -> dict:
    """read_nlte_data.f:148-233 for one species: the permanent arrays the NLTE opacity uses.
    Lines and collision lines are stored lower level first; a radiative line's wavelength
    comes from its levels' energies (1e8/|chi_u|-|chi_l| in Angstrom) unless that is not
    positive, then from the file; the Einstein coefficients use the file's wavelength and
    Econst = 4 pi (pi e^2/(m_e c)) with the single-precision pi e^2/(m_e c) of the
    reference."""
    fconst = float(np.float32(_PC_PI * _PC_E ** 2 / (_PC_ME * _PC_C)))
    econst = 4.0 * _PC_PI * fconst
    chi, g = raw["lev_chi"], raw["lev_g"].astype(float)
    lo = np.minimum(raw["lin_lev1"], raw["lin_lev2"])
    hi = np.maximum(raw["lin_lev1"], raw["lin_lev2"])
    n = len(lo)
    wl0 = np.zeros(n)
    for k in range(n):
        dchi = abs(chi[hi[k] - 1]) - abs(chi[lo[k] - 1])
        if dchi != 0.0:
            wl0[k] = 1e8 / dchi
        if wl0[k] <= 0.0:
            wl0[k] = raw["lin_wl0"][k]
    f = raw["lin_f"]
    wf = raw["lin_wl0"] * _PC_ANG
    gl, gu = g[lo - 1], g[hi - 1]
    aul = (2.0 * econst * gl * f) / (gu * wf ** 2)
    bul = (econst * wf * gl * f) / (_PC_H * _PC_C * gu)
    blu = (econst * wf * f) / (_PC_H * _PC_C)
    clo = np.minimum(raw["col_lev1"], raw["col_lev2"])
    chi_ = np.maximum(raw["col_lev1"], raw["col_lev2"])
    return {
        "nlev": len(chi), "glev": g, "chilev": chi,
        "llw": lo, "lup": hi, "wl0": wl0, "f": f, "gl": gl, "gu": gu,
        "aul": aul, "bul": bul, "blu": blu,
        "llw_coll": clo, "lup_coll": chi_, "coll_c": raw["col_c"], "dele": raw["col_dele"],
        "lev_pi": raw["pi_lev"], "pi_c": raw["pi_c"], "rr_c": raw["rr_c"], "dele_pi": raw["pi_dele"],
    }
