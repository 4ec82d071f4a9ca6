"""GPU parity: the HIP library (through its C ABI) against the CPU oracle and against
SuperLite's own outputs in tests/golden/.  Marked gpu; run on an MI355X."""
import os
import sys

import numpy as np
import pytest
from scipy import stats

from oracle import pyoracle as po
from superlite_amd import synth

pytestmark = pytest.mark.gpu


def ulp_diff(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def S():
    from superlite_amd import slgpu
    return slgpu


def make_ctx(S, name, tables, **kw):
    w, st, wl, lines = synth.build(name)
    ctx = S.setup_workload(w, st, wl, lines, tables, **kw)
    return ctx, w, st, wl, lines


def gpu_opacity(ctx, st, flags=0):
    return ctx.opacity_build(st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"],
                             st["natomfr"], st["oc"], st["ginv"], flags=flags)


def test_library_exports(S):
    L = S.lib()
    for name in S.EXPORTS:
        assert hasattr(L, name)


def test_opacity_c1(S, tables, golden):
    say("case: opacity c1")
    ctx, w, st, wl, lines = make_ctx(S, "c1", tables, npartmax=1024)
    g = gpu_opacity(ctx, st)
    o = po.opacity(wl, st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"], st["natomfr"],
                   st["ion_iz"], st["ion_ii"], st["oc"], st["ginv"], lines, tables)
    rel = np.abs(g["cap64"] - o["cap64"]) / np.abs(o["cap64"])
    say("c1 cap64 max rel vs oracle", rel.max())
    assert rel.max() < 1e-12
    ref = golden("ref_c1_opacity.npz")
    assert ulp_diff(g["cap"], ref["cap"]).max() <= 1
    assert ulp_diff(g["emiss"], ref["emiss"]).max() <= 1
    np.testing.assert_allclose(g["sig"], ref["sig"], rtol=1e-14)
    np.testing.assert_allclose(g["capgrey"], ref["capgrey"], rtol=1e-12)
    np.testing.assert_allclose(g["capross"], ref["capross"], rtol=1e-12)
    ctx.close()


@pytest.mark.parametrize("flags", [0, 6, 5, 3, 8])
def test_opacity_edge(S, tables, golden, flags):
    say("case: opacity edge flags", flags)
    e = golden("ref_edge_opacity.npz")
    ctx = S.Context(3, 5, 2, 2, npartmax=16)
    ctx.set_grid(e["xarr"], e["vxarr"], e["divv"], e["vol"], 0.5 * (e["xarr"][1:] + e["xarr"][:-1]))
    ctx.set_groups(e["wl"])
    ctx.set_ions(e["ion_iz"], e["ion_ii"])
    ctx.set_tables(*tables)
    ctx.set_lines(e["wl0"], e["gxs"], e["chilw"], e["liz"], e["lii"])
    g = ctx.opacity_build(e["temp"], e["natom"], e["vol"], e["mass"], e["nelec"], e["natomfr"],
                          e["oc"], e["ginv"], flags=flags)
    o = po.opacity(e["wl"], e["temp"], e["natom"], e["vol"], e["mass"], e["nelec"], e["natomfr"],
                   e["ion_iz"], e["ion_ii"], e["oc"], e["ginv"],
                   (e["wl0"], e["gxs"], e["chilw"], e["liz"], e["lii"]), tables, flags=flags)
    nz = o["cap64"] != 0
    assert np.array_equal(g["cap64"] != 0, nz)
    np.testing.assert_allclose(g["cap64"][nz], o["cap64"][nz], rtol=1e-12)
    if f"cap_{flags}" in e:
        assert ulp_diff(g["cap"], e[f"cap_{flags}"]).max() <= 1
    np.testing.assert_allclose(g["sig"], o["sig"], rtol=1e-14)
    ctx.close()


def test_opacity_c2_full_size(S, tables):
    """production size: 400 shells x 10^4 groups x 2e6 lines; oracle on 16 of the shells"""
    say("case: opacity c2")
    ctx, w, st, wl, lines = make_ctx(S, "c2", tables, npartmax=1024)
    g = gpu_opacity(ctx, st)
    cells = np.arange(0, w.nx, 25)
    sub = {k: st[k][cells] for k in ("temp", "natom", "vol", "mass", "nelec", "oc", "ginv")}
    o = po.opacity(wl, sub["temp"], sub["natom"], sub["vol"], sub["mass"], sub["nelec"],
                   st["natomfr"][:, cells], st["ion_iz"], st["ion_ii"], sub["oc"], sub["ginv"],
                   lines, tables)
    rel = np.abs(g["cap64"][cells] - o["cap64"]) / np.abs(o["cap64"])
    say("c2 cap64 max rel vs oracle", rel.max())
    assert rel.max() < 1e-12
    assert ulp_diff(g["cap"][cells], o["cap"]).max() <= 1
    np.testing.assert_allclose(g["capgrey"][cells], o["capgrey"], rtol=1e-12)
    np.testing.assert_allclose(g["capross"][cells], o["capross"], rtol=1e-12)
    ctx.close()


def test_opacity_c5_full_size(S, tables):
    """1000 shells with the dense CSM shell (8000-9000 km/s, rho x1e3, T to 5e4 K), 10^4
    groups, 4e6 lines; the oracle on 22 shells including the CSM ones"""
    say("case: opacity c5")
    ctx, w, st, wl, lines = make_ctx(S, "c5", tables, npartmax=1024)
    g = gpu_opacity(ctx, st)
    cells = np.unique(np.concatenate([np.arange(0, w.nx, 50), [272, 285, 298]]))
    sub = {k: st[k][cells] for k in ("temp", "natom", "vol", "mass", "nelec", "oc", "ginv")}
    o = po.opacity(wl, sub["temp"], sub["natom"], sub["vol"], sub["mass"], sub["nelec"],
                   st["natomfr"][:, cells], st["ion_iz"], st["ion_ii"], sub["oc"], sub["ginv"],
                   lines, tables)
    m = o["cap64"] != 0
    rel = np.abs(g["cap64"][cells][m] - o["cap64"][m]) / np.abs(o["cap64"][m])
    say("c5 cap64 max rel vs oracle", rel.max(), "T range", sub["temp"].min(), sub["temp"].max())
    assert np.array_equal(g["cap64"][cells] != 0, m)
    assert rel.max() < 1e-12
    assert ulp_diff(g["cap"][cells], o["cap"]).max() <= 1
    np.testing.assert_allclose(g["capgrey"][cells], o["capgrey"], rtol=1e-12)
    np.testing.assert_allclose(g["capross"][cells], o["capross"], rtol=1e-12)
    ctx.close()


def oracle_model(st, w, wl, g, prep=None, nvol=None, **kw):
    m = po.Model(st["xarr"], st["vxarr"], st["divv"], st["vol"], wl, int(st["nep"]),
                 int(st["nepg"]), g["cap"], g["sig"], g["capgrey"], 1.0 / st["temp"], st["temp"],
                 tauddmc=w.tauddmc, taulump=w.taulump, **kw)
    if prep is not None:
        m.set(opaclump=prep["opaclump"], emitprob=prep["emitprob"], emit=prep["emit"])
    if nvol is not None:
        m.set(nvol=nvol)
    return m


def test_ddmc_prep_c1(S, tables):
    say("case: ddmc prep c1")
    ctx, w, st, wl, lines = make_ctx(S, "c1", tables, npartmax=1024)
    g = gpu_opacity(ctx, st)
    p = ctx.ddmc_prep(st["temp"], w.lbol)
    m = oracle_model(st, w, wl, g)
    ol = m.leakage()
    for k in (0, 1, 6, 7, 9, 10):
        np.testing.assert_allclose(p["opaclump"][:, k], ol[:, k], rtol=1e-12, atol=0,
                                   err_msg=f"opaclump({k + 1})")
    np.testing.assert_allclose(p["emitprob"], m.emission_probability(), rtol=1e-12)
    emit, sth = po.sourceenergy(st["rcell"], st["temp"], w.lbol)
    np.testing.assert_allclose(p["emit"], emit, rtol=1e-15)
    assert p["sthermal"] == pytest.approx(w.lbol, rel=1e-12)
    ctx.close()


@pytest.mark.parametrize("name", ["c1", "c2", "c5"])
def test_sampler_tables_exact(S, tables, name):
    """The bucket maps the transport samples groups with return exactly what the plain
    searches return, for random draws and for draws on and next to every kind of bucket
    edge; emitgroup also matches the oracle's emitgroup.f restatement."""
    say(f"case: sampler tables {name}")
    ctx, w, st, wl, lines = make_ctx(S, name, tables, npartmax=1024)
    g = gpu_opacity(ctx, st)
    p = ctx.ddmc_prep(st["temp"], w.lbol)
    K = 1 << 14
    rng = np.random.default_rng(11)
    b = rng.integers(0, K, 2048).astype(np.float64)
    edges = np.concatenate([b / K, np.nextafter((b + 1) / K, 0.0), np.nextafter(b / K, 1.0),
                            np.nextafter(b / K, 0.0), [1.0, np.nextafter(1.0, 0.0), 2.0 ** -54]])
    r = np.concatenate([rng.random(8192), edges])
    r = r[(r > 0.0) & (r <= 1.0)]
    cells = np.arange(1, w.nx + 1, max(1, w.nx // 24))
    rr = np.tile(r, len(cells))
    cc = np.repeat(cells, len(r)).astype(np.int32)
    for t in range(5):
        a = ctx.sample_groups(t, rr, cc)
        ref = ctx.sample_groups(t, rr, cc, direct=True)
        bad = np.flatnonzero(a != ref)
        say(f"table {t}: {len(rr)} draws, {len(bad)} mismatches")
        assert len(bad) == 0, (t, rr[bad[:5]], cc[bad[:5]], a[bad[:5]], ref[bad[:5]])
        assert (a[(t == 4) | (a != 0)] >= 1).all() and (a <= w.ng).all()
    m = oracle_model(st, w, wl, g, prep=p)
    sel = rng.choice(len(rr), 20000, replace=False)
    e = ctx.sample_groups(4, rr[sel], cc[sel])
    o = np.array([m.emitgroup(rr[k], cc[k]) for k in sel])
    bad = np.flatnonzero(e != o)
    say("emitgroup vs oracle: mismatches", len(bad), "of", len(sel))
    # every mismatch is a draw within rounding of an edge of the oracle's emitgroup CDF
    for k in bad:
        ic = int(cc[sel[k]])
        r1 = rr[sel[k]] * g["capgrey"][ic - 1]
        edges = emitgroup_edges(st, w, wl, g, p, ic)
        gap = np.min(np.abs(edges - r1)) / r1
        say(f"  r={rr[sel[k]]!r} cell {ic}: gpu {e[k]} oracle {o[k]}, nearest CDF edge {gap:.1e}")
        assert gap < 1e-12 and abs(int(e[k]) - int(o[k])) == 1
    ctx.close()


def emitgroup_edges(st, w, wl, g, p, ic):
    """every value r*capgrey is compared against in emitgroup.f:23-51 for cell ic: the binned
    emission probabilities and the in-bin running sums of specintv*cap from each bin's start"""
    spec = po.specintv(1.0 / st["temp"][ic - 1], 1.0 / wl)
    cap = g["cap"][ic - 1].astype(np.float64)
    ep = p["emitprob"][ic - 1]
    nep, nepg = int(st["nep"]), int(st["nepg"])
    edges = [ep]
    for iep in range(nep):
        g0, g1 = iep * nepg, min(iep * nepg + nepg - 1, w.ng)
        acc = 0.0 if iep == 0 else ep[iep - 1]
        run = []
        for ig in range(g0, g1 - 1 + 1):
            if ig >= w.ng:
                break
            acc = acc + spec[ig] * cap[ig]
            run.append(acc)
        edges.append(np.array(run))
    return np.concatenate(edges)


def test_precision_form_c1(S, tables):
    """FP32-mixed against FP64 on the same inputs and seeds: opacities within 1e-5, the
    per-shell deposition and the escaped energy within 1%, the spectrum difference far
    inside the Monte Carlo noise, and energy still conserved to 1e-12."""
    say("case: precision form c1")
    n2s = 20
    out = {}
    for prec in (S.SLGPU_FP64, S.SLGPU_MIXED):
        ctx, w, st, wl, lines = make_ctx(S, "c1", tables, npartmax=1 << n2s, seed=9,
                                         precision=prec)
        op = gpu_opacity(ctx, st)
        ctx.ddmc_prep(st["temp"], w.lbol, want=False)
        ctx.sources(1 << n2s, 1)
        tr = ctx.transport(1)
        fl = ctx.fluxtally()
        out[prec] = (op, tr, fl)
        ctx.close()
    (o0, t0, f0), (o1, t1, f1) = out[S.SLGPU_FP64], out[S.SLGPU_MIXED]
    m = o0["cap64"] > 0
    rel = np.abs(o1["cap64"][m] - o0["cap64"][m]) / o0["cap64"][m]
    say("cap64 max rel", rel.max())
    assert rel.max() < 1e-5
    for k in ("capgrey", "capross"):
        assert np.max(np.abs(o1[k] - o0[k]) / o0[k]) < 1e-5, k
    tally = np.abs(t1["tally"] - t0["tally"]) / np.abs(t0["tally"])
    say("shell tally max rel", tally.max(), "eout rel", abs(t1["eout"] - t0["eout"]) / t0["eout"])
    assert tally.max() < 1e-2
    assert abs(t1["eout"] - t0["eout"]) / t0["eout"] < 1e-2
    sel = f0["lumnum"] > 100
    z = (f1["luminos"][sel] - f0["luminos"][sel]) / np.sqrt(f0["lumdev"][sel] + f1["lumdev"][sel])
    say("spectrum z rms", np.sqrt(np.mean(z ** 2)))
    assert np.sqrt(np.mean(z ** 2)) < 0.5
    for t in (t0, t1):
        assert abs(t["estart"] - (t["eout"] + t["evelo"] + t["edead"])) < 1e-12 * t["estart"]


def _perturbed(st, kind, idx, h):
    s = dict(st)
    a = np.array(st[kind], dtype=np.float64, copy=True)
    a[idx] *= 1.0 + h
    s[kind] = a
    return s, st[kind][idx] * h


def _fd_check(ctx, st, flags, probes, Wc, h=1e-4):
    """central differences of sum(Wc*cap64), taken element by element, against the VJP"""
    base = gpu_opacity(ctx, st, flags)
    g = ctx.opacity_vjp(grad_cap64=Wc)
    worst = 0.0
    for kind, idx in probes:
        sp, dx = _perturbed(st, kind, idx, h)
        sm, _ = _perturbed(st, kind, idx, -h)
        fd = float(np.sum(Wc * (gpu_opacity(ctx, sp, flags)["cap64"] -
                                gpu_opacity(ctx, sm, flags)["cap64"]))) / (2 * dx)
        an = g[kind][idx]
        err = abs(fd - an) / max(abs(fd), abs(an), 1e-300)
        worst = max(worst, err)
        say(f"flags {flags} {kind:8s} {str(idx):12s} fd {fd: .6e} vjp {an: .6e} rel {err:.2e}")
    return worst, g, base


def test_opacity_vjp_central_differences(S, tables):
    """Backward of the opacity build against central differences (h = 1e-4 relative) of
    the GPU FP64 forward and of the oracle's FP64 forward.  Each input kind is differenced
    where its term is resolvable in FP64: the free-free inputs (natom, n_e, mass fractions)
    on ff-only builds, since ff is <= 1e-10 of the c1 opacity; the populations on ions
    holding at least 1e-3 of their cell's largest population; temperature with every term
    on, with bb+bf only and with ff only.  The means take the FP32-rounded gas_cap as the
    reference does, a step function that differences cannot resolve, so they are
    differenced on temperature, whose Planck-weight dependence is smooth (h = 1e-3), to
    1e-3: the staircase still shows in probes whose derivative nearly cancels."""
    say("case: opacity backward c1")
    ctx, w, st, wl, lines = make_ctx(S, "c1", tables, npartmax=1024)
    rng = np.random.default_rng(21)
    nc, ng = w.nx, w.ng
    temps = [("temp", (int(i),)) for i in rng.choice(nc, 6, replace=False)]
    oc = st["oc"]
    big = np.argwhere(oc >= 1e-3 * oc.max(axis=1, keepdims=True))
    ocs = [("oc", tuple(int(v) for v in big[k])) for k in rng.choice(len(big), 8, replace=False)]
    present = [z for z in range(30) if np.any(st["natomfr"][z] != 0)]
    ffs = [("natom", (int(i),)) for i in rng.choice(nc, 4, replace=False)]
    ffs += [("nelec", (int(i),)) for i in rng.choice(nc, 4, replace=False)]
    ffs += [("natomfr", (int(rng.choice(present)), int(i))) for i in rng.choice(nc, 6, replace=False)]
    checks = [(0, temps + ocs), (S.NOFF, temps[:3]), (S.NOBB | S.NOBF, temps[:3] + ffs)]
    worst = 0.0
    for flags, probes in checks:
        base = gpu_opacity(ctx, st, flags)
        Wc = rng.standard_normal((nc, ng)) / np.abs(base["cap64"][base["cap64"] > 0]).mean()
        wv, g_cap, _ = _fd_check(ctx, st, flags, probes, Wc)
        worst = max(worst, wv)
        if flags == 0:
            g0, W0 = g_cap, Wc
    say("worst rel error (group opacity)", worst)
    assert worst < 1e-4
    # the means, temperature probes
    base = gpu_opacity(ctx, st)
    Wp = rng.standard_normal(nc) / np.abs(base["capgrey"]).mean()
    Wr = rng.standard_normal(nc) / np.abs(base["capross"]).mean()
    g_mean = ctx.opacity_vjp(grad_capgrey=Wp, grad_capross=Wr)
    worst_m = 0.0
    for kind, idx in temps:
        sp, dx = _perturbed(st, kind, idx, 1e-3)
        sm, _ = _perturbed(st, kind, idx, -1e-3)
        op, om = gpu_opacity(ctx, sp), gpu_opacity(ctx, sm)
        fd = (float(np.sum(Wp * (op["capgrey"] - om["capgrey"]))) +
              float(np.sum(Wr * (op["capross"] - om["capross"])))) / (2 * dx)
        an = g_mean[kind][idx]
        err = abs(fd - an) / max(abs(fd), abs(an), 1e-300)
        worst_m = max(worst_m, err)
        say(f"means {kind} {idx} fd {fd: .6e} vjp {an: .6e} rel {err:.2e}")
    # the FP32 gas_cap staircase inside the means limits these differences to ~1e-4
    assert worst_m < 1e-3
    # the oracle's FP64 forward differenced on temperature and population probes
    for kind, idx in temps[:2] + ocs[:2]:
        caps = []
        dx = st[kind][idx] * 1e-4
        for sgn in (1, -1):
            sp, _ = _perturbed(st, kind, idx, sgn * 1e-4)
            o = po.opacity(wl, sp["temp"], sp["natom"], sp["vol"], sp["mass"], sp["nelec"],
                           sp["natomfr"], sp["ion_iz"], sp["ion_ii"], sp["oc"], sp["ginv"], lines, tables)
            caps.append(o["cap64"])
        fd = float(np.sum(W0 * (caps[0] - caps[1]))) / (2 * dx)
        err = abs(fd - g0[kind][idx]) / max(abs(fd), 1e-300)
        say(f"oracle {kind} {idx}: fd {fd:.6e} vjp {g0[kind][idx]:.6e} rel {err:.2e}")
        assert err < 1e-4
    Wc = W0
    g_cap = g0
    gpu_opacity(ctx, st)
    # the autograd.Function gives the same gradient
    import torch
    from superlite_amd import autograd as sag
    consts = dict(vol=st["vol"], mass=st["mass"], ginv=st["ginv"])
    ts = {k: torch.tensor(np.asarray(st[k], np.float64), requires_grad=True)
          for k in ("temp", "natom", "nelec", "natomfr", "oc")}
    cap, grey, ross = sag.opacity(ctx, consts, ts["temp"], ts["natom"], ts["nelec"], ts["natomfr"], ts["oc"])
    (cap * torch.from_numpy(Wc)).sum().backward()
    np.testing.assert_allclose(ts["temp"].grad.numpy(), g_cap["temp"], rtol=1e-12)
    np.testing.assert_allclose(ts["oc"].grad.numpy(), g_cap["oc"], rtol=1e-12, atol=0)
    ctx.close()


@pytest.mark.parametrize("part,flags", [("bf", 1 | 4 | 8), ("bb", 2 | 4 | 8)])
def test_opacity_vjp_populations_linear(S, tables, part, flags):
    """With only bf (or only bb) on, the group opacity is linear in the ion populations, so
    dL/doc(i, k) = sum_g G(i, g) cap_k(i, g) where cap_k is the build with oc = 1 for ion k
    and 0 for the others: the backward's per-ion gradient against that sum over all ions and
    cells of c1 (FP64, relative to the largest gradient), and in the precision form.  The bb
    sum skips lines of ions whose population is not positive (physical_opacity.f:146), so
    there (hydrogen at c1) the opacity is flat in oc and the gradient is 0."""
    say(f"case: population gradient of the {part} part against its linear reference, c1")
    rng = np.random.default_rng(13)
    for prec, tol in ((S.SLGPU_FP64, 1e-12), (S.SLGPU_MIXED, 1e-5)):
        ctx, w, st, wl, lines = make_ctx(S, "c1", tables, npartmax=1024, precision=prec)
        G = rng.standard_normal((w.nx, w.ng))
        args = [st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"], st["natomfr"], st["oc"], st["ginv"]]
        ctx.opacity_build(*args, flags=flags, want=("cap64",))
        g = ctx.opacity_vjp(G, None, None)["oc"]
        ref = np.zeros_like(st["oc"])
        for k in range(st["oc"].shape[1]):
            oc1 = np.zeros_like(st["oc"])
            oc1[:, k] = 1.0
            a = list(args)
            a[6] = oc1
            ref[:, k] = (G * ctx.opacity_build(*a, flags=flags, want=("cap64",))["cap64"]).sum(axis=1)
        ctx.close()
        live = st["oc"] > 0 if part == "bb" else np.ones(st["oc"].shape, bool)
        err = np.abs(g - ref)[live].max() / np.abs(ref[live]).max()
        say(f"{part} precision {prec}: max |goc - ref| / max|ref| = {err:.2e} "
            f"over {live.sum()} of {live.size} (cell, ion)")
        assert err < tol
        assert np.all(g[~live] == 0.0)


def test_opacity_vjp_precision_form(S, tables):
    """The precision form's backward (FP32 line terms) against the FP64 backward on
    the same inputs and output gradients: every input gradient within 1e-5 of the largest
    FP64 gradient of its kind."""
    say("case: opacity backward, precision form c1")
    rng = np.random.default_rng(5)
    res = {}
    for prec in (S.SLGPU_FP64, S.SLGPU_MIXED):
        ctx, w, st, wl, lines = make_ctx(S, "c1", tables, npartmax=1024, precision=prec)
        base = gpu_opacity(ctx, st)
        if prec == S.SLGPU_FP64:
            Wc = rng.standard_normal(base["cap64"].shape) / np.abs(base["cap64"]).mean()
            Wp = rng.standard_normal(w.nx) / np.abs(base["capgrey"]).mean()
            Wr = rng.standard_normal(w.nx) / np.abs(base["capross"]).mean()
        res[prec] = ctx.opacity_vjp(Wc, Wp, Wr)
        ctx.close()
    f, m = res[S.SLGPU_FP64], res[S.SLGPU_MIXED]
    for k in ("temp", "natom", "nelec", "natomfr", "oc"):
        scale = np.abs(f[k]).max()
        err = np.abs(m[k] - f[k]).max() / scale
        say(f"{k}: max |mixed - fp64| / max|fp64| = {err:.2e}")
        assert err < 1e-5, k


def test_opacity_unstaged_kernels_c1(S, tables):
    """The opacity build and its backward through the kernels they take when the ion
    populations do not fit in LDS (forced with slgpu_set_unstaged; no workload here has
    enough ions to select them by size), against the staged kernels on the same inputs and
    output gradients, in both precisions: FP64 within 1e-12, the precision form within 1e-5
    of the largest value of each kind (its staged line sums run in ion order, the unstaged
    ones in wavelength order)."""
    say("case: opacity build and backward, unstaged against staged kernels, c1")
    rng = np.random.default_rng(11)
    for prec, tol in ((S.SLGPU_FP64, 1e-12), (S.SLGPU_MIXED, 1e-5)):
        out = {}
        for unstaged in (False, True):
            ctx, w, st, wl, lines = make_ctx(S, "c1", tables, npartmax=1024, precision=prec)
            if unstaged:
                ctx.set_unstaged(True)
            o = gpu_opacity(ctx, st)
            if not unstaged:
                Wc = rng.standard_normal(o["cap64"].shape) / np.abs(o["cap64"]).mean()
                Wp = rng.standard_normal(w.nx) / np.abs(o["capgrey"]).mean()
                Wr = rng.standard_normal(w.nx) / np.abs(o["capross"]).mean()
            out[unstaged] = (o, ctx.opacity_vjp(Wc, Wp, Wr))
            ctx.close()
        (o0, g0), (o1, g1) = out[False], out[True]
        for k in ("cap64", "capgrey", "capross"):
            err = np.abs(o1[k] - o0[k]).max() / np.abs(o0[k]).max()
            say(f"precision {prec} {k}: max |unstaged - staged| / max = {err:.2e}")
            assert err < tol, k
        for k in ("temp", "natom", "nelec", "natomfr", "oc"):
            scale = np.abs(g0[k]).max()
            err = np.abs(g1[k] - g0[k]).max() / scale
            say(f"precision {prec} grad {k}: max |unstaged - staged| / max = {err:.2e}")
            assert err < tol, k


def test_sources_c1_per_particle(S, tables):
    say("case: sources c1")
    ctx, w, st, wl, lines = make_ctx(S, "c1", tables, npartmax=1 << 14, seed=5)
    g = gpu_opacity(ctx, st)
    p = ctx.ddmc_prep(st["temp"], w.lbol)
    nvol, nnew = ctx.sources(1 << 14, 3)
    assert nnew == 1 << 14
    assert np.array_equal(nvol, po.sourcenumbers(p["emit"], 1 << 14))
    m = oracle_model(st, w, wl, g, prep=p, nvol=nvol)
    o = m.sources(5, 3)
    # the GPU keeps the sampled state in its particle store; read it back before transport
    pp = ctx.particles(nnew)
    np.testing.assert_allclose(pp["x"], o["x"], rtol=1e-13)
    np.testing.assert_allclose(pp["mu"], o["mu"], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(pp["e"], o["e"], rtol=1e-14)
    same_wl = np.isclose(pp["wl"], o["wl"], rtol=1e-13)
    # a different wavelength means a different group; each such particle's group draw lies
    # within rounding of an edge of the oracle's cumulative source spectrum
    bad = np.flatnonzero(~same_wl)
    say("source wavelength mismatches", len(bad), "of", nnew)
    pstart = np.concatenate([[0], np.cumsum(nvol)])
    for q in bad:
        i = int(np.searchsorted(pstart, q, side="right"))         # owning cell, 1-based
        prob = po.specintv(1.0 / st["temp"][i - 1], 1.0 / wl)
        prob = prob / prob.sum()
        edges = np.cumsum(prob)
        r = philox_draw0(5, 3, int(q))
        gap = np.min(np.abs(edges - r)) / r
        gg = np.searchsorted(wl, [pp["wl"][q], o["wl"][q]])
        say(f"  particle {q} cell {i}: groups {gg}, draw {r!r}, nearest CDF edge {gap:.1e}")
        assert gap < 1e-12 and abs(int(gg[0]) - int(gg[1])) == 1
    ctx.close()


def philox_draw0(seed, iteration, p):
    """draw 0 of source particle p: Philox-4x32-10 block (p, 0, 0, 0) under key
    (seed, iteration), words 0 and 1 as a 53-bit uniform"""
    import ctypes

    class U4(ctypes.Structure):
        _fields_ = [("v", ctypes.c_uint32 * 4)]
    f = po.lib().or_philox_block
    f.restype = U4
    f.argtypes = [U4, ctypes.c_uint32, ctypes.c_uint32]
    out = f(U4((ctypes.c_uint32 * 4)(p, 0, 0, 0)), seed, iteration)
    a, b = out.v[0], out.v[1]
    return (float(((a >> 5) << 26) | (b >> 6)) + 0.5) * 2.0 ** -53


def run_both(S, tables, n2s, seed, it, scale=1.0, max_steps=0):
    """GPU iteration, and the oracle iteration on the GPU's own per-iteration state."""
    ctx, w, st, wl, lines = make_ctx(S, "c1", tables, npartmax=1 << n2s, seed=seed,
                                     max_steps=max_steps)
    st = dict(st)
    st["oc"] = st["oc"] * scale
    st["nelec"] = st["nelec"] * scale
    flags = 0 if scale == 1.0 else S.NOFF
    g = ctx.opacity_build(st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"],
                          st["natomfr"], st["oc"], st["ginv"], flags=flags)
    p = ctx.ddmc_prep(st["temp"], w.lbol)
    nvol, nnew = ctx.sources(1 << n2s, it)
    tr = ctx.transport(it)
    fl = ctx.fluxtally()
    pp = ctx.particles(nnew)
    m = oracle_model(st, w, wl, g, prep=p, nvol=nvol, max_steps=max_steps)
    o = m.iteration(seed, it, per_particle=True)
    ctx.close()
    return tr, fl, pp, o


@pytest.mark.parametrize("scale", [1e-5, 1.0])
def test_transport_c1_per_particle(S, tables, scale):
    say("case: transport c1 per particle, scale", scale)
    tr, fl, pp, o = run_both(S, tables, 12, 7, 2, scale=scale)
    op = o["particles"]
    assert tr["nerr"] == 0 and o["nerr"] == 0
    same = (pp["stat"] == op["stat"]) & (pp["ig"] == op["ig"]) & \
        np.isclose(pp["e"], op["e"], rtol=1e-9) & np.isclose(pp["wl"], op["wl"], rtol=1e-9)
    say("identical trajectories", same.mean(), "steps gpu", tr["nstepimc"], tr["nstepddmc"],
        "oracle", o["nstepimc"], o["nstepddmc"])
    # same discrete outcomes for every particle (exp/log differ by ulps between the device
    # and the host, which has not flipped a comparison on any tested particle)
    assert same.all() and np.array_equal(pp["nstep"], op["nstep"])
    assert (tr["nstepimc"], tr["nstepddmc"]) == (o["nstepimc"], o["nstepddmc"])
    # energy bookkeeping: every particle's start energy ends as escaped, adiabatic or dead
    assert tr["estart"] == pytest.approx(tr["eout"] + tr["evelo"] + tr["edead"], rel=1e-12)
    assert tr["esrc"] == pytest.approx(o["esrc"], rel=1e-12)
    assert tr["nactive"] == o["nactive"]
    # flux tally of the device equals the binning of the device's own final states
    flux = pp["stat"] == 1
    ig = np.clip(np.searchsorted(synth.group_grid(synth.WORKLOADS["c1"]), pp["wl"][flux],
                                 side="right"), 1, 1000)
    lum = np.bincount(ig - 1, weights=pp["e"][flux], minlength=1000)
    np.testing.assert_allclose(fl["luminos"], lum, rtol=1e-12, atol=1e-300)
    assert fl["lumnum"].sum() == flux.sum() == tr["nflux"]
    np.testing.assert_allclose(fl["luminos"].sum(), tr["eout"], rtol=1e-12)
    # tallies agree with the oracle's to the trajectory-divergence level
    np.testing.assert_allclose(tr["tally"], o["tally"], rtol=2e-2)
    assert abs(tr["eout"] - o["eout"]) / o["eout"] < 2e-2


def test_transport_modes_c1_per_particle(S, tables):
    """The amplification factor and anisotropic Thomson scattering switched on: GPU against
    the oracle particle by particle, the per-cell grd_eamp tally, and the energy identity
    e_start = e_out + e_velo + e_dead + sum(grd_eamp)."""
    say("case: transport c1 per particle, ampfact + thm_aniso")
    ctx, w, st, wl, lines = make_ctx(S, "c1", tables, npartmax=1 << 12, seed=19, ampfact=True,
                                     thm_aniso=True)
    g = gpu_opacity(ctx, st)
    p = ctx.ddmc_prep(st["temp"], w.lbol)
    nvol, nnew = ctx.sources(1 << 12, 4)
    tr = ctx.transport(4)
    pp = ctx.particles(nnew)
    ctx.close()
    m = oracle_model(st, w, wl, g, prep=p, nvol=nvol, ampfact=True, thm_aniso=True)
    o = m.iteration(19, 4, per_particle=True)
    op = o["particles"]
    assert tr["nerr"] == 0 and o["nerr"] == 0
    same = (pp["stat"] == op["stat"]) & (pp["ig"] == op["ig"]) & \
        np.isclose(pp["e"], op["e"], rtol=1e-9) & np.isclose(pp["wl"], op["wl"], rtol=1e-9)
    say("identical trajectories", same.mean(), "eamp gpu", tr["eampsum"], "oracle", o["eampsum"])
    assert same.all()
    assert tr["eampsum"] > 0
    assert tr["estart"] == pytest.approx(tr["eout"] + tr["evelo"] + tr["edead"] + tr["eampsum"],
                                         rel=1e-12)
    np.testing.assert_allclose(tr["eamp"].sum(), tr["eampsum"], rtol=1e-12)
    np.testing.assert_allclose(tr["eamp"], o["eamp"], rtol=2e-2, atol=1e-3 * o["eamp"].max())


def coarse(a, nb=20):
    a = np.asarray(a)
    n = a.shape[-1] - a.shape[-1] % nb
    return a[..., :n].reshape(a.shape[:-1] + (nb, -1)).sum(-1)


def welch_z(ref_rows, our_rows):
    """Per bin, the difference of the two samples' per-iteration means over its standard
    error, as the normal deviate of the same two-sided Welch t p-value (the variances are
    estimated from as few as 24 rows, so t is used, not the normal).  Bins where both
    samples are constant are dropped."""
    a, b = np.asarray(ref_rows, float), np.asarray(our_rows, float)
    na, nb = len(a), len(b)
    va, vb = a.var(0, ddof=1) / na, b.var(0, ddof=1) / nb
    s2 = va + vb
    ok = s2 > 0
    va, vb, s2 = va[ok], vb[ok], s2[ok]
    t = (a.mean(0)[ok] - b.mean(0)[ok]) / np.sqrt(s2)
    nu = s2 ** 2 / (va ** 2 / (na - 1) + vb ** 2 / (nb - 1))
    return np.sign(t) * stats.norm.isf(stats.t.sf(np.abs(t), nu))


def chi2_bound(dof, p=1e-3):
    """chi^2/dof above which a sum of dof squared unit normals lies with probability p"""
    return stats.chi2.isf(p, dof) / dof


def zmax_bound(n, p=1e-3):
    """|z| that the largest of n unit normals exceeds with probability p (Sidak)"""
    return stats.norm.isf(0.5 * (1.0 - (1.0 - p) ** (1.0 / n)))


@pytest.mark.parametrize("case", ["c1mixed", "c1", "c2"])
def test_transport_statistics_vs_reference(S, tables, golden, case):
    """MC tallies within 3 sigma of SuperLite's own Monte Carlo (golden per-iteration rows)."""
    say("case: transport statistics", case)
    if not os.path.exists(os.path.join(os.path.dirname(__file__), "golden", f"ref_{case}_mc.npz")):
        pytest.skip(f"ref_{case}_mc.npz not generated (tests/golden/make_golden.py --mc-c2)")
    g = golden(f"ref_{case}_mc.npz")
    scale = float(g["scale"])
    n2s = int(g["n2s"])
    workload = str(g["workload"]) if "workload" in g else "c1"
    w, st, wl, lines = synth.build(workload)
    st = dict(st)
    st["oc"] = st["oc"] * scale
    st["nelec"] = st["nelec"] * scale
    ctx = S.setup_workload(w, st, wl, lines, tables, npartmax=1 << n2s, seed=21)
    rows = []
    for it in range(24 if workload == "c1" else 16):
        r = S.iteration(ctx, st, 1 << n2s, it + 1, w.lbol, flags=int(g["flags"]))
        rows.append(dict(luminos=r["flux"]["luminos"], tally=r["transport"]["tally"],
                         eout=r["transport"]["eout"], evelo=r["transport"]["evelo"]))
        t = r["transport"]
        assert t["estart"] == pytest.approx(t["eout"] + t["evelo"] + t["edead"], rel=1e-12)
    ctx.close()

    def z(ref, ours):
        a, b = np.asarray(ref, float), np.asarray(ours, float)
        ma, sa = a.mean(0), a.std(0, ddof=1) / np.sqrt(len(a))
        mb, sb = b.mean(0), b.std(0, ddof=1) / np.sqrt(len(b))
        s = np.sqrt(sa ** 2 + sb ** 2)
        ok = s > 0
        return (ma[ok] - mb[ok]) / s[ok]
    for key, f in (("luminos", coarse), ("tally", lambda a: coarse(a, 10))):
        zz = z(f(g[key]), [f(r[key]) for r in rows])
        say(case, key, "coarse chi2/dof", np.mean(zz ** 2), "max|z|", np.max(np.abs(zz)))
        assert np.mean(zz ** 2) < 2.5
        assert np.max(np.abs(zz)) < 4.5
    # per group (groups with at least 100 reference escapes over its iterations) and per
    # shell: chi^2/dof against the chi^2(dof) quantile, the largest |z| against the bound
    # for that many bins, both at p = 1e-3
    keep = np.asarray(g["lumnum"]).sum(0) >= 100
    per = [("tally", np.asarray(g["tally"]), np.array([r["tally"] for r in rows]))]
    if keep.sum() >= 20:
        per.append(("luminos", np.asarray(g["luminos"])[:, keep],
                    np.array([r["luminos"] for r in rows])[:, keep]))
    for key, ref_rows, our_rows in per:
        zz = welch_z(ref_rows, our_rows)
        n = len(zz)
        say(case, key, "per bin: n", n, "chi2/dof", np.mean(zz ** 2), "bound", chi2_bound(n),
            "max|z|", np.max(np.abs(zz)), "bound", zmax_bound(n))
        assert np.mean(zz ** 2) < chi2_bound(n)
        assert np.max(np.abs(zz)) < zmax_bound(n)
    for key in ("eout", "evelo"):
        zz = z(g[key][:, None], [[r[key]] for r in rows])
        assert abs(zz[0]) < 4.0, (key, zz)


def test_transport_error_policy(S, tables):
    """step cap reached: fatal with errorfatal (particle_advance.f90:262-263), else 'dead'"""
    say("case: error policy")
    w, st, wl, lines = synth.build("c1")
    for fatal in (True, False):
        ctx = S.setup_workload(w, st, wl, lines, tables, npartmax=256, seed=3, max_steps=3,
                               errorfatal=fatal)
        gpu_opacity(ctx, st)
        ctx.ddmc_prep(st["temp"], w.lbol)
        ctx.sources(256, 1)
        tr = ctx.transport(1, check=False)
        assert tr["nerr"] > 0
        assert tr["ndead"] == tr["nerr"]
        assert (tr["rc"] == S.E_TRANSPORT) == fatal
        assert tr["estart"] == pytest.approx(tr["eout"] + tr["evelo"] + tr["edead"], rel=1e-12)
        pp = ctx.particles()
        assert set(np.unique(pp["err"][pp["stat"] == 2])) <= {200}
        ctx.close()


def test_call_order_errors(S, tables):
    w, st, wl, lines = synth.build("c1")
    ctx = S.Context(w.nx, w.ng, int(st["nep"]), int(st["nepg"]), npartmax=64)
    with pytest.raises(S.SlgpuError) as ei:
        ctx.sources(64, 1)
    assert ei.value.code == 2
    with pytest.raises(S.SlgpuError):
        ctx.set_lines(lines[0][::-1].copy(), *lines[1:])   # unsorted line list
    ctx.close()
