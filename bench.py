"""Benchmark of SuperLite's per-iteration hot path on one MI355X.

A step is one SuperLite iteration (superlite.f90:157-219) on the device: opacity build
(bb/bf/ff + Planck/Rosseland means), DDMC prep (leakage opacities, emission
probabilities, sampling tables), source sampling, IMC/DDMC transport of all particles
and the flux tally.  Inputs (mesh, groups, line list, atomic tables) are device resident
before the timed region; each step uploads only the per-cell gas state.

value = particle-steps (transport11 + diffusion11 events, particle_advance.f90:209,219)
per second of whole-step wall time, for the FP32-mixed precision form by default; the FP64
path is timed beside it on the same inputs and seeds, and the line carries the mixed-vs-FP64
differences of the opacities, shell tallies, escaped energy and spectrum.  The opacity rate
(cell-groups per second of the opacity-build phase) is reported beside it.

One iteration at the default workload (400 shells, 10^4 groups, 2e6 lines, 2^24 particles)
takes tens of seconds, so --steps is an upper bound: each precision times whole iterations
until its share of --budget-s would be exceeded (at least one), and `steps` reports the
iterations actually timed beside `steps_requested`.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0


def cpu_baseline(w, st, wl, lines, tables, target_s=10.0):
    """The oracle (CPU FP64 restatement, OpenMP) on a bounded sample of the same workload:
    the opacity over a subset of shells, then the transport of a particle sample sized from
    a first small run so the whole baseline stays near target_s of CPU time."""
    from oracle import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cells = np.arange(0, w.nx, max(1, w.nx // 8))
    t0 = time.perf_counter()
    op = po.opacity(wl, st["temp"][cells], st["natom"][cells], st["vol"][cells], st["mass"][cells],
                    st["nelec"][cells], st["natomfr"][:, cells], st["ion_iz"], st["ion_ii"],
                    st["oc"][cells], st["ginv"][cells], lines, tables)
    t_op = time.perf_counter() - t0
    opac_rate = len(cells) * w.ng / t_op
    full = po.opacity(wl, st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"],
                      st["natomfr"], st["ion_iz"], st["ion_ii"], st["oc"], st["ginv"], lines,
                      tables)
    m = po.Model(st["xarr"], st["vxarr"], st["divv"], st["vol"], wl, int(st["nep"]),
                 int(st["nepg"]), full["cap"], full["sig"], full["capgrey"], 1.0 / st["temp"],
                 st["temp"], tauddmc=w.tauddmc, taulump=w.taulump)
    m.set(opaclump=m.leakage(), emitprob=m.emission_probability())
    emit, _ = po.sourceenergy(st["rcell"], st["temp"], w.lbol)
    m.set(emit=emit)
    n, spent = 32, 0.0
    while True:
        m.set(nvol=po.sourcenumbers(emit, n))
        t0 = time.perf_counter()
        r = m.iteration(77, 1)
        dt = time.perf_counter() - t0
        spent += dt
        steps = r["nstepimc"] + r["nstepddmc"]
        rate = steps / dt
        sample = f"{n} particles of the workload ({steps} particle-steps, {dt:.1f} s)"
        left = target_s - spent
        # per-particle cost is heavy-tailed: grow by at most 4x, to about half the time left
        if left < 2.0 * dt or n >= (1 << 16):
            break
        n = min(1 << 16, n * min(4, 1 << int(np.log2(max(2.0, 0.5 * left / dt)))))
    return dict(value=rate, unit="particle-steps/s", cores=threads, kind="port",
                sample=sample, opacity_cell_groups_per_s=opac_rate,
                opacity_sample=f"{len(cells)} of {w.nx} cells x {w.ng} groups x {w.nline} lines",
                note="FP64 oracle restatement on the host cores; a baseline, not a target")


def _counter_record(kernel, workload):
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("workload") != workload:
        return None
    return rec


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of
    known size) from the committed rocprofv3 counter summary, if it covers this workload."""
    rec = _counter_record(kernel, workload)
    return None if rec is None else rec["fetch_bytes"] + rec["write_bytes"]


# VALU issue rate of MI355X: 256 CUs x 4 SIMDs x 2.4 GHz, one wave64 instruction per 4 cycles
VALU_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4


def valu_issue_ms(kernel, workload):
    """Time the committed SQ_INSTS_VALU count of one launch of `kernel` takes at the VALU
    issue rate, or None."""
    rec = _counter_record(kernel, workload)
    if rec is None or "valu_wave_insts" not in rec:
        return None
    return rec["valu_wave_insts"] / VALU_ISSUE_PER_S * 1e3


def calibration_copy(mib):
    """One elementwise kernel that reads and writes `mib` MiB: the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters of a --pmc run to bytes."""
    import torch
    x = torch.ones(mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
    y = x.mul(2.0)
    torch.cuda.synchronize()
    print(f"# calibration copy: {mib} MiB read, {mib} MiB written (elementwise mul)",
          file=sys.stderr, flush=True)
    del x, y


def run_precision(slgpu, w, st, wl, lines, tables, prec, args, budget_s):
    """Warm up, then time whole iterations at the workload's particle count: at most
    args.steps, and no more than fit in budget_s of wall time (at least one)."""
    nstot = 1 << w.n2s
    ctx = slgpu.setup_workload(w, st, wl, lines, tables, precision=prec, seed=2024,
                               npartmax=nstot, errorfatal=False)
    name = "mixed" if prec == slgpu.SLGPU_MIXED else "fp64"
    for k in range(args.warmup):
        r = slgpu.iteration(ctx, st, 1 << min(args.warmup_n2s, w.n2s), 1000 + k, w.lbol, want=False)
        print(f"# {name} warmup {k}: steps {r['transport']['nstepimc'] + r['transport']['nstepddmc']}",
              file=sys.stderr, flush=True)
    ctx.timers(reset=True)
    steps, nerr, first, done = 0, 0, None, 0
    t0 = time.perf_counter()
    while done < args.steps:
        r = slgpu.iteration(ctx, st, nstot, done + 1, w.lbol, want=True)
        tr = r["transport"]
        steps += tr["nstepimc"] + tr["nstepddmc"]
        nerr += tr["nerr"]
        if first is None:
            first = r
        done += 1
        el = time.perf_counter() - t0
        print(f"# {name} step {done - 1}: particle-steps {tr['nstepimc'] + tr['nstepddmc']} "
              f"(imc {tr['nstepimc']}, ddmc {tr['nstepddmc']}), max/particle {tr['nstepmax']}, "
              f"escaped {tr['nflux']}, errors {tr['nerr']}, elapsed {el:.1f} s",
              file=sys.stderr, flush=True)
        if el + el / done > budget_s:
            break
    dt = time.perf_counter() - t0
    tm = ctx.timers()
    ctx.close()
    return dict(steps=steps, iters=done, dt=dt, tm=tm, nerr=nerr, first=first)


def precision_errors(m, f):
    """FP32-mixed against FP64 on the same inputs and seeds (the first timed iteration of
    each, iteration number 1 in both)."""
    om, of = m["first"]["opacity"], f["first"]["opacity"]
    c0, c1 = of["cap"].astype(np.float64), om["cap"].astype(np.float64)
    sel = c0 > 0
    tm_, tf = m["first"]["transport"], f["first"]["transport"]
    fm, ff = m["first"]["flux"], f["first"]["flux"]
    g = ff["lumnum"] > 1000
    z = (fm["luminos"][g] - ff["luminos"][g]) / np.sqrt(fm["lumdev"][g] + ff["lumdev"][g])
    return {
        "gas_cap_rel_max": float(np.max(np.abs(c1[sel] - c0[sel]) / c0[sel])),
        "capgrey_rel_max": float(np.max(np.abs(om["capgrey"] - of["capgrey"]) / of["capgrey"])),
        "capross_rel_max": float(np.max(np.abs(om["capross"] - of["capross"]) / of["capross"])),
        "shell_tally_rel_max": float(np.max(np.abs(tm_["tally"] - tf["tally"]) / np.abs(tf["tally"]))),
        "escaped_energy_rel": float(abs(tm_["eout"] - tf["eout"]) / tf["eout"]),
        "spectrum_groups_compared": int(g.sum()),
        "spectrum_diff_over_mc_sigma_rms": float(np.sqrt(np.mean(z ** 2))),
        "spectrum_group_rel_median": float(np.median(np.abs(fm["luminos"][g] - ff["luminos"][g]) / ff["luminos"][g])),
    }


def opacity_grad(args):
    """Differentiation form at the workload's size: forward = device EOS (massfr2natomfr +
    LTE Saha solve) + opacity build + means, backward = opacity VJP + EOS VJP, giving
    dL/d(T, rho, X) for dL/d(cap, Planck mean, Rosseland mean), the output gradients resident
    on the device.  Timed with the library's HIP-event timers: whole phases (including the
    host transfers of the per-cell inputs and input gradients) and the kernels alone."""
    from superlite_amd import slgpu, synth
    w, st, wl, lines = synth.build(args.workload)
    tables = synth.load_tables()
    prec = slgpu.SLGPU_MIXED if args.precision == "mixed" else slgpu.SLGPU_FP64
    ctx = slgpu.setup_workload(w, st, wl, lines, tables, npartmax=1024, seed=2024, precision=prec)
    ctx.set_eos_data(synth.load_eos_data(args.workload))
    rng = np.random.default_rng(7)
    # the output gradients live on the device, as a torch caller's would: the backward reads
    # them in place (slgpu_opacity_vjp_dev), so neither phase moves an nc x ng array over PCIe
    import torch
    gcap = torch.from_numpy(rng.standard_normal((w.nx, w.ng))).to("cuda:0")
    ones = torch.ones(w.nx, dtype=torch.float64, device="cuda:0")
    torch.cuda.synchronize()
    nelec = st["nelec"].copy()

    def step():
        e = ctx.eos(st["temp"], st["mass"], st["massfr"], nelec=nelec)
        ctx.opacity_build(st["temp"], e["natom"], st["vol"], st["mass"], e["nelec"], e["natomfr"],
                          e["oc"], e["ginv"], want=())
        g = ctx.opacity_vjp_dev(gcap.data_ptr(), ones.data_ptr(), ones.data_ptr())
        ctx.eos_vjp(g["natom"], g["natomfr"], g["nelec"], g["oc"])
    for _ in range(args.warmup):
        step()
    ctx.timers(reset=True)
    t0 = time.perf_counter()
    K = min(args.steps, 200)
    for _ in range(K):
        step()
    dt = time.perf_counter() - t0
    tm = ctx.timers()
    ctx.close()
    cg = w.nx * w.ng
    fwd = (tm["eos"] + tm["opacity"]) / K
    bwd = (tm["opacity_vjp"] + tm["eos_vjp"]) / K
    fk, bk = tm["opacity_kernels"] / K, tm["opacity_vjp_kernels"] / K
    print(json.dumps({
        "metric": "opacity cell·groups/s, forward and backward (differentiation form)",
        "value": cg / (bwd * 1e-3), "unit": "cell·groups/s (backward)",
        "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": dt / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if prec == slgpu.SLGPU_MIXED else "f64", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w.nx} shells x {w.ng} groups x {w.nline} lines, "
                               "EOS + opacity forward, VJP w.r.t. T, rho and mass fractions"},
        "forward_ms": fwd, "backward_ms": bwd, "backward_over_forward": bwd / fwd,
        "opacity_kernels_forward_ms": fk, "opacity_kernels_backward_ms": bk,
        "kernels_backward_over_forward": bk / fk,
        "eos_ms": tm["eos"] / K, "eos_vjp_ms": tm["eos_vjp"] / K,
        "forward_cell_groups_per_s": cg / (fwd * 1e-3)}))


def nlte_iteration(args):
    """Iterations with hydrogen in NLTE (superlite.f90:157-208 with in_nlte), every step on the
    device: the EOS and its nlte_ndens, the NLTE opacity build (per shell the 100-level H I
    rate equations and their LU solve, the H lines' NLTE absorption and emission beside the
    other ions' LTE lines, bf and ff, the means), the DDMC prep and samplers on the emission
    opacity, sources, transport, flux tally and nlte_rad_est.  Every iteration starts from the
    same gas state and radiation temperature (0.9 T), so each times the same work; value =
    particle-steps per second of whole-iteration wall time, as in the LTE transport mode."""
    from superlite_amd import slgpu, synth
    w, st, wl, lines = synth.build(args.workload)
    tables = synth.load_tables()
    prec = slgpu.SLGPU_FP64 if args.precision == "fp64" else slgpu.SLGPU_MIXED
    nstot = 1 << w.n2s
    ctx = slgpu.setup_workload(w, st, wl, lines, tables, precision=prec, seed=2024,
                               npartmax=nstot, errorfatal=False)
    ctx.set_eos_data(synth.load_eos_data(args.workload))
    ctx.set_nlte_data(synth.load_nlte_data())
    radtemp = 0.9 * st["temp"]

    def step(n, it):
        e = ctx.eos(st["temp"], st["mass"], st["massfr"], nelec=st["nelec"])
        nd = ctx.eos_nlte_ndens()
        s2 = dict(st)
        s2.update(natom=e["natom"], natomfr=e["natomfr"], nelec=e["nelec"], oc=e["oc"], ginv=e["ginv"])
        return slgpu.iteration_nlte(ctx, s2, n, it, w.lbol, radtemp, nd)
    for k in range(args.warmup):
        step(1 << min(args.warmup_n2s, w.n2s), 1000 + k)
    ctx.timers(reset=True)
    steps, nerr, done = 0, 0, 0
    t0 = time.perf_counter()
    while done < args.steps:
        r = step(nstot, done + 1)
        tr = r["transport"]
        steps += tr["nstepimc"] + tr["nstepddmc"]
        nerr += tr["nerr"]
        done += 1
        el = time.perf_counter() - t0
        print(f"# nlte step {done - 1}: particle-steps {tr['nstepimc'] + tr['nstepddmc']}, escaped "
              f"{tr['nflux']}, errors {tr['nerr']}, elapsed {el:.1f} s", file=sys.stderr, flush=True)
        if el + el / done > args.budget_s:
            break
    dt = time.perf_counter() - t0
    tm = ctx.timers()
    ctx.close()
    K = done
    mixed = prec == slgpu.SLGPU_MIXED
    print(json.dumps({
        "metric": "particle-steps/s, one iteration with H I in NLTE",
        "value": steps / dt, "unit": "particle-steps/s",
        "n_gpus": 1, "steps": K, "steps_requested": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / K * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if mixed else "f64", "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w.nx} shells x {w.ng} groups x {w.nline} lines, "
                               f"2^{w.n2s} particles/iteration, H I in NLTE (100 levels, 900 lines)",
                   "precision": "mixed" if mixed else "fp64",
                   "step": "EOS + nlte_ndens, NLTE opacity, DDMC prep, sources, transport, "
                           "flux tally, nlte_rad_est"},
        "transport_errors": nerr,
        "nlte_opacity_kernels_ms": tm["opacity_kernels"] / K,
        "nlte_opacity_cell_groups_per_s": w.nx * w.ng * K / (tm["opacity_kernels"] * 1e-3),
        "phase_ms_per_step": {k: v / K for k, v in tm.items()}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="most iterations timed per precision (the time budget may stop earlier)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"],
                    help="both: value is the FP32-mixed form, timed beside FP64 on the same seeds")
    ap.add_argument("--budget-s", type=float, default=330.0,
                    help="wall-clock budget of the timed iterations, shared by the precisions "
                         "(55%% to the first); each times at least one iteration")
    ap.add_argument("--warmup-n2s", type=int, default=16,
                    help="log2 particles of the warmup iterations")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mode", default="transport", choices=["transport", "opacity-grad", "nlte"],
                    help="opacity-grad: time the opacity build and its backward instead; "
                         "nlte: whole iterations with H I in NLTE (precision mixed unless fp64)")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="run one elementwise copy of this size first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "superlite_amd runs on one GPU; the sharded path is designed "
                                   "but not built (run with --gpus 1)", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.calibration_copy_mib > 0:
        calibration_copy(args.calibration_copy_mib)
    if args.mode == "opacity-grad":
        opacity_grad(args)
        return
    if args.mode == "nlte":
        nlte_iteration(args)
        return
    from superlite_amd import slgpu, synth
    w, st, wl, lines = synth.build(args.workload)
    tables = synth.load_tables()
    nstot = 1 << w.n2s
    print(f"# workload {args.workload}: nx={w.nx} ng={w.ng} lines={w.nline} "
          f"particles=2^{w.n2s}", file=sys.stderr, flush=True)
    exp32, exp64 = slgpu.peak_rates()
    print(f"# measured exponential throughput: v_exp_f32 {exp32:.3e}/s, FP64 exp {exp64:.3e}/s",
          file=sys.stderr, flush=True)
    precs = {"both": [slgpu.SLGPU_MIXED, slgpu.SLGPU_FP64], "mixed": [slgpu.SLGPU_MIXED],
             "fp64": [slgpu.SLGPU_FP64]}[args.precision]
    shares = [0.55, 0.45] if len(precs) == 2 else [1.0]
    runs = {p: run_precision(slgpu, w, st, wl, lines, tables, p, args, args.budget_s * f)
            for p, f in zip(precs, shares)}
    main_p = precs[0]
    R = runs[main_p]
    K = R["iters"]
    value = R["steps"] / R["dt"]
    tm = R["tm"]
    mixed = main_p == slgpu.SLGPU_MIXED
    kname = "k_transport<true>" if mixed else "k_transport<false>"
    # dominant kernel: k_transport.  Algorithmic bytes: per particle-step the FP32 group
    # opacity gather (4 B) and the FP64 read-modify-write of the step's grd_tally and
    # grd_jrad entries (2 x 16 B) = 36 B; per particle its state read once and written once,
    # 2 x 56 B (FP32-mixed) or 2 x 80 B (FP64).  Lookup tables the implementation adds
    # (thresholds, bucket maps) are not algorithmic bytes: they show in the measured traffic.
    t_kernel = tm["transport_kernel"] * 1e-3 / K
    state_b = 56.0 if mixed else 80.0
    bytes_kernel = 36.0 * R["steps"] / K + 2.0 * state_b * nstot
    achieved = bytes_kernel / t_kernel / 1e9
    # opacity build: one exponential per (line, cell) in the line loop
    nline_used = int(np.count_nonzero((lines[0] * 1e-8 >= wl[0]) & (lines[0] * 1e-8 < wl[-1])))
    t_opk = tm["opacity_kernel"] * 1e-3 / K
    line_evals = float(nline_used) * w.nx
    exp_peak = exp32 if mixed else exp64
    opac_alg_bytes = w.nx * w.ng * (8.0 + 4.0 + 4.0) + 16.0 * nline_used
    out = {
        "metric": "particle-steps/s + opacity cell·groups/s, FP32-mixed vs FP64 at ≤1% tally err",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1, "steps": K, "steps_requested": args.steps, "warmup": args.warmup,
        "ms_per_step": R["dt"] / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w.nx} shells to {w.vmax / 1e5:.0f} km/s, "
                               f"{w.ng} groups {w.wlmin * 1e8:.0f}-{w.wlmax * 1e8:.0f} A, "
                               f"{w.nline} lines, 2^{w.n2s} particles/iteration",
                   "precision": "mixed" if mixed else "fp64",
                   "step": "one whole iteration: opacity build, DDMC prep, sources, transport, flux tally"},
        # the build on the device (cell prep, bb+bf+ff, transpose, means), inputs resident;
        # the phase rate adds the gas-state upload and the gas_cap download over PCIe
        "opacity_cell_groups_per_s": w.nx * w.ng * K / (tm["opacity_kernels"] * 1e-3),
        "opacity_phase_cell_groups_per_s": w.nx * w.ng * K / (tm["opacity"] * 1e-3),
        "opacity_kernel_cell_groups_per_s": w.nx * w.ng / t_opk,
        "transport_errors": R["nerr"],
        "phase_ms_per_step": {k: v / K for k, v in tm.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS,
                     "traffic": measured_traffic(kname, args.workload),
                     "kernel": kname, "kernel_ms": t_kernel * 1e3,
                     "algorithmic_bytes": f"36 B/particle-step + 2x{state_b:.0f} B/particle",
                     "limiter": "latency and divergence (lanes active, wave waits), not bandwidth"},
        "roofline_opacity": {"bound": "exp", "kernel": "k_opacity<true>" if mixed else "k_opacity<false>",
                             "achieved": line_evals / t_opk, "peak": exp_peak,
                             "unit": "line evaluations/s vs measured " +
                                     ("v_exp_f32/s" if mixed else "FP64 exp()/s"),
                             "frac": line_evals / t_opk / exp_peak, "kernel_ms": t_opk * 1e3,
                             "hbm_achieved_gbs": opac_alg_bytes / t_opk / 1e9,
                             "hbm_frac": opac_alg_bytes / t_opk / 1e9 / PEAK_HBM_GBS,
                             # the binding resource: its VALU instructions (rocprofv3
                             # SQ_INSTS_VALU of one launch) at the issue rate, over the
                             # measured time
                             "valu_issue_ms": valu_issue_ms("k_opacity<true>" if mixed else "-", args.workload),
                             "valu_issue_frac": (lambda v: None if v is None else v / (t_opk * 1e3))(
                                 valu_issue_ms("k_opacity<true>" if mixed else "-", args.workload))},
        "measured_peaks": {"exp32_per_s": exp32, "exp64_per_s": exp64},
    }
    if slgpu.SLGPU_FP64 in runs and mixed:
        F = runs[slgpu.SLGPU_FP64]
        KF = F["iters"]
        out["fp64"] = {"value": F["steps"] / F["dt"], "steps": KF, "ms_per_step": F["dt"] / KF * 1e3,
                       "opacity_cell_groups_per_s": w.nx * w.ng * KF / (F["tm"]["opacity_kernels"] * 1e-3),
                       "opacity_phase_cell_groups_per_s": w.nx * w.ng * KF / (F["tm"]["opacity"] * 1e-3),
                       "opacity_kernel_cell_groups_per_s": w.nx * w.ng * KF / (F["tm"]["opacity_kernel"] * 1e-3),
                       "transport_errors": F["nerr"],
                       "phase_ms_per_step": {k: v / KF for k, v in F["tm"].items()}}
        out["mixed_speedup_vs_fp64"] = value / out["fp64"]["value"]
        out["mixed_vs_fp64_errors"] = precision_errors(R, F)
    if not args.no_cpu_baseline:
        print("# cpu baseline (oracle) ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(w, st, wl, lines, tables)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
