"""NLTE hydrogen opacity on the GPU (slgpu_opacity_build_nlte through the C ABI) against the
oracle's restatement of GAS/physical_opacity_nlte.f, at the c2 size (400 shells, 10^4
groups, 2e6 lines, the 100-level H I atom).  Marked gpu; run on an MI355X."""
import numpy as np
import pytest

from oracle import pyoracle as po
from superlite_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def ulp_diff(a, b):
    a = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    b = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(a - b)


@pytest.fixture(scope="module")
def c2():
    w, st, wl, lines = synth.build("c2")
    ed = synth.load_eos_data("c2")
    kH = [k for k in range(len(st["ion_iz"])) if st["ion_iz"][k] == 1 and st["ion_ii"][k] == 1][0]
    nh = st["natom"] * np.asarray(st["natomfr"])[0] / st["vol"]
    n1 = st["oc"][:, kH] * ed["q"][0][0] / ed["g1"][0][0]
    ndens = np.stack([n1, nh - n1], axis=1)
    # radiation temperature a little below the gas temperature, as NLTE iterations leave it
    radtemp = 0.9 * st["temp"]
    return w, st, wl, lines, ndens, radtemp


def test_nlte_opacity_c2_vs_oracle(c2):
    from superlite_amd import slgpu
    w, st, wl, lines, ndens, radtemp = c2
    tables = synth.load_tables()
    sp = synth.load_nlte_data()
    say("case: nlte opacity c2, 400 shells")
    ctx = slgpu.setup_workload(w, st, wl, lines, tables, npartmax=1024)
    with pytest.raises(slgpu.SlgpuError):      # the atom must be set first
        ctx.opacity_build_nlte(st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"],
                               st["natomfr"], st["oc"], st["ginv"], radtemp, ndens)
    ctx.set_nlte_data(sp)
    g = ctx.opacity_build_nlte(st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"],
                               st["natomfr"], st["oc"], st["ginv"], radtemp, ndens)
    ctx.close()
    cells = np.linspace(0, w.nx - 1, 10).astype(int)
    keep = ~((np.asarray(lines[3]) == 1) & (np.asarray(lines[4]) == 1))
    lte_lines = tuple(np.asarray(x)[keep] for x in lines)
    o = po.opacity_nlte(sp, radtemp[cells], ndens[cells], wl, st["temp"][cells], st["natom"][cells],
                        st["vol"][cells], st["mass"][cells], st["nelec"][cells],
                        np.asarray(st["natomfr"])[:, cells], st["ion_iz"], st["ion_ii"],
                        st["oc"][cells], st["ginv"][cells], lte_lines, tables)
    assert o["rc"] == 0
    # populations: the same factorisation; rates differ by the device's and the host's
    # exp/log/pow ulps, which the solve passes on scaled by the matrix's conditioning
    pg, po_ = g["npop"][cells], o["npop"]
    big = np.abs(po_) > 1e-6 * np.abs(po_).max(axis=1, keepdims=True)
    rel = np.max(np.abs(pg[big] - po_[big]) / np.abs(po_[big]))
    say("populations max rel", rel)
    assert rel < 1e-9
    np.testing.assert_allclose(g["npop"].sum(axis=1)[st["mass"] > 0], 1.0, rtol=1e-10)
    sel = o["cap64"] > 0
    r64 = np.max(np.abs(g["cap64"][cells][sel] - o["cap64"][sel]) / o["cap64"][sel])
    say("cap64 max rel", r64)
    assert r64 < 1e-8
    for k in ("cap", "capemit", "emiss"):
        u = ulp_diff(g[k][cells], o[k])
        say(k, "max ulp", u.max(), "frac > 1 ulp", np.mean(u > 1))
        assert u.max() <= 2 and np.mean(u > 1) < 1e-4, k
    for k in ("capgrey", "capemitgrey", "capross", "sig"):
        np.testing.assert_allclose(g[k][cells], o[k], rtol=1e-8, err_msg=k)
    # NLTE differs from LTE where the H lines are: the emission opacity is not the absorption
    assert not np.array_equal(g["capemit"], g["cap"])


# --- the transport after an NLTE build: the DDMC prep and the group samplers on gas_capemit
# (leakage_opacity11.f90:76-108, emission_probability.f90:28-60, emitgroup.f:23-50,
# diffusion11.f90:136-182,576-581) and nlte_rad_est (nltemod.f:231-260)
def _nlte_setup(S, c2, sp, **kw):
    w, st, wl, lines, ndens, radtemp = c2
    ctx = S.setup_workload(w, st, wl, lines, synth.load_tables(), **kw)
    ctx.set_nlte_data(sp)
    g = ctx.opacity_build_nlte(st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"],
                               st["natomfr"], st["oc"], st["ginv"], radtemp, ndens)
    return ctx, g


def _nlte_model(c2, g, prep=None, nvol=None, **kw):
    w, st, wl, lines, ndens, radtemp = c2
    m = po.Model(st["xarr"], st["vxarr"], st["divv"], st["vol"], wl, int(st["nep"]),
                 int(st["nepg"]), g["cap"], g["sig"], g["capgrey"], 1.0 / st["temp"], radtemp,
                 tauddmc=w.tauddmc, taulump=w.taulump, capemit=g["capemit"],
                 capemitgrey=g["capemitgrey"], **kw)
    if prep is not None:
        m.set(opaclump=prep["opaclump"], emitprob=prep["emitprob"], emit=prep["emit"])
    if nvol is not None:
        m.set(nvol=nvol)
    return m


def test_nlte_prep_and_samplers_c2(c2):
    """after the NLTE build at c2: grd_opaclump (with capemitlump and igemitmax) and
    grd_emitprob against the oracle's in_nlte prep on the same opacities; the bucket maps of
    the five group samplers equal their plain searches; emitgroup equals the oracle's"""
    from superlite_amd import slgpu as S
    w, st, wl, lines, ndens, radtemp = c2
    say("case: nlte ddmc prep and samplers c2")
    ctx, g = _nlte_setup(S, c2, synth.load_nlte_data(), npartmax=1024)
    p = ctx.ddmc_prep(radtemp, w.lbol)
    m = _nlte_model(c2, g)
    ol = m.leakage()
    for k in (0, 1, 6, 7, 8, 10):
        np.testing.assert_allclose(p["opaclump"][:, k], ol[:, k], rtol=1e-12, atol=0,
                                   err_msg=f"opaclump({k + 1})")
    assert np.array_equal(p["opaclump"][:, 9], ol[:, 9])
    lumped = p["opaclump"][:, 6] > 0
    assert lumped.any() and np.all(p["opaclump"][lumped, 8] > 0)
    np.testing.assert_allclose(p["emitprob"], m.emission_probability(), rtol=1e-12)
    K = 1 << 14
    rng = np.random.default_rng(5)
    b = rng.integers(0, K, 1024).astype(np.float64)
    r = np.concatenate([rng.random(4096), b / K, np.nextafter((b + 1) / K, 0.0),
                        np.nextafter(b / K, 1.0), [1.0, np.nextafter(1.0, 0.0)]])
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
    mp = _nlte_model(c2, g, prep=p)
    sel = rng.choice(len(rr), 8000, replace=False)
    e = ctx.sample_groups(4, rr[sel], cc[sel])
    o = np.array([mp.emitgroup(rr[k], cc[k]) for k in sel])
    bad = np.flatnonzero(e != o)
    say("emitgroup vs oracle: mismatches", len(bad), "of", len(sel))
    # a mismatch may only be a draw within rounding of an edge of emitgroup's CDF
    for k in bad:
        ic = int(cc[sel[k]])
        r1 = rr[sel[k]] * g["capemitgrey"][ic - 1]
        spec = po.specintv(1.0 / st["temp"][ic - 1], 1.0 / wl)
        run = np.cumsum(spec * g["capemit"][ic - 1].astype(np.float64))
        edges = np.concatenate([p["emitprob"][ic - 1], run])
        assert np.min(np.abs(edges - r1)) / r1 < 1e-11 and abs(int(e[k]) - int(o[k])) == 1
    # the LTE prep on the same context differs (emission follows capemit, not cap)
    g2 = ctx.opacity_build(st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"],
                           st["natomfr"], st["oc"], st["ginv"])
    p2 = ctx.ddmc_prep(radtemp, w.lbol)
    assert np.all(p2["opaclump"][:, 8] == 0)
    assert not np.allclose(p2["emitprob"], p["emitprob"], rtol=1e-3)
    ctx.close()


def test_nlte_transport_c2_per_particle(c2):
    """one NLTE iteration at c2 (2^10 particles, walks truncated at 3000 events): every
    particle's final state equals the oracle's in_nlte transport; then nlte_rad_est on the
    transport's tally against the oracle"""
    from superlite_amd import slgpu as S
    w, st, wl, lines, ndens, radtemp = c2
    say("case: nlte transport per particle c2 (truncated at 3000 events)")
    ctx, g = _nlte_setup(S, c2, synth.load_nlte_data(), precision=S.SLGPU_FP64,
                         npartmax=1 << 10, seed=23, max_steps=3000, errorfatal=False)
    prep = ctx.ddmc_prep(radtemp, w.lbol)
    nvol, nnew = ctx.sources(1 << 10, 1)
    tr = ctx.transport(1, check=False)
    pp = ctx.particles(nnew)
    rt, jrad = ctx.nlte_rad_est(radtemp)
    ctx.close()
    cons = abs(tr["estart"] - (tr["eout"] + tr["evelo"] + tr["edead"])) / tr["estart"]
    assert cons < 1e-12, cons
    o = _nlte_model(c2, g, prep=prep, nvol=nvol, max_steps=3000).iteration(23, 1, per_particle=True)
    op = o["particles"]
    same = (pp["stat"] == op["stat"]) & (pp["ig"] == op["ig"]) & (pp["nstep"] == op["nstep"]) & \
        np.isclose(pp["e"], op["e"], rtol=1e-9) & np.isclose(pp["wl"], op["wl"], rtol=1e-9) & \
        np.isclose(pp["x"], op["x"], rtol=1e-9)
    say(f"identical trajectories {same.sum()} of {len(same)}; steps gpu {tr['nstepimc']} "
        f"{tr['nstepddmc']} oracle {o['nstepimc']} {o['nstepddmc']}")
    assert set(np.unique(pp["err"])) <= {0, 200}
    assert same.all()
    rto, jo = po.nlte_rad_est(wl, st["vol"], tr["tally"], radtemp)
    np.testing.assert_allclose(rt, rto, rtol=1e-15, atol=0)
    np.testing.assert_allclose(jrad, jo, rtol=1e-13, atol=0)


def test_nlte_call_order(c2):
    """the backward covers the LTE build only; new NLTE data invalidates an NLTE build's
    opacity (the DDMC prep then needs a new build); nlte_rad_est needs a transport"""
    from superlite_amd import slgpu as S
    w, st, wl, lines, ndens, radtemp = c2
    say("case: nlte call order")
    sp = synth.load_nlte_data()
    ctx, g = _nlte_setup(S, c2, sp, npartmax=1024)
    with pytest.raises(S.SlgpuError):
        ctx.opacity_vjp(grad_capgrey=np.ones(w.nx))
    with pytest.raises(S.SlgpuError):
        ctx.nlte_rad_est(radtemp)
    ctx.set_nlte_data(sp)
    with pytest.raises(S.SlgpuError):
        ctx.ddmc_prep(radtemp, w.lbol)
    ctx.close()


def test_nlte_iteration_device_chain_c2(c2):
    """one NLTE iteration with every step on the device, in both precision forms: the EOS and
    its nlte_ndens, the NLTE opacity, the DDMC prep, sources, transport, flux tally and
    nlte_rad_est (superlite.f90:157-208 with in_nlte); the densities equal those derived from
    the reference's own EOS state, and the iteration conserves energy"""
    from superlite_amd import slgpu as S
    w, st, wl, lines, ndens, radtemp = c2
    sp = synth.load_nlte_data()
    for prec in (S.SLGPU_FP64, S.SLGPU_MIXED):
        say(f"case: nlte iteration on the device c2, precision {prec}")
        ctx = S.setup_workload(w, st, wl, lines, synth.load_tables(), precision=prec,
                               npartmax=1 << 16, seed=3)
        ctx.set_eos_data(synth.load_eos_data("c2"))
        ctx.set_nlte_data(sp)
        e = ctx.eos(st["temp"], st["mass"], st["massfr"])
        nd = ctx.eos_nlte_ndens()
        nh = e["natom"] * e["natomfr"][0] / st["vol"]
        np.testing.assert_allclose(nd[:, 0], ndens[:, 0], rtol=1e-12)
        assert np.all(np.abs(nd[:, 1] - ndens[:, 1]) <= 1e-13 * nh)
        s2 = dict(st)
        s2.update(natom=e["natom"], natomfr=e["natomfr"], nelec=e["nelec"], oc=e["oc"], ginv=e["ginv"])
        r = S.iteration_nlte(ctx, s2, 1 << 16, 1, w.lbol, radtemp, nd)
        ctx.close()
        tr = r["transport"]
        cons = abs(tr["estart"] - (tr["eout"] + tr["evelo"] + tr["edead"])) / tr["estart"]
        say(f"steps {tr['nstepimc']} {tr['nstepddmc']} flux {tr['nflux']} conservation {cons:.1e}")
        assert cons < 1e-12 and tr["nerr"] == 0
        assert r["flux"]["lumnum"].sum() == tr["nflux"]
        assert np.all(np.isfinite(r["radtemp"])) and np.all(r["radtemp"] > 0)
        np.testing.assert_allclose(r["opacity"]["npop"].sum(axis=1)[st["mass"] > 0], 1.0, rtol=1e-10)
