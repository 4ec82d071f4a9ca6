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


def cpu_baseline(w, st, wl, lines, tables, target_s=15.0):
    """The oracle (CPU FP64 restatement, OpenMP) on a bounded sample of the same workload."""
    from oracle import pyoracle as po
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    # opacity: a subset of cells over the full line list and group grid
    cells = np.arange(0, w.nx, max(1, w.nx // 8))
    t0 = time.perf_counter()
    op = po.opacity(wl, st["temp"][cells], st["natom"][cells], st["vol"][cells], st["mass"][cells],
                    st["nelec"][cells], st["natomfr"][:, cells], st["ion_iz"], st["ion_ii"],
                    st["oc"][cells], st["ginv"][cells], lines, tables)
    t_op = time.perf_counter() - t0
    opac_rate = len(cells) * w.ng / t_op
    # transport: full opacity for the model, then a small particle sample
    full = po.opacity(wl, st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"],
                      st["natomfr"], st["ion_iz"], st["ion_ii"], st["oc"], st["ginv"], lines,
                      tables)
    m = po.Model(st["xarr"], st["vxarr"], st["divv"], st["vol"], wl, int(st["nep"]),
                 int(st["nepg"]), full["cap"], full["sig"], full["capgrey"], 1.0 / st["temp"],
                 st["temp"], tauddmc=w.tauddmc, taulump=w.taulump)
    m.set(opaclump=m.leakage(), emitprob=m.emission_probability())
    emit, _ = po.sourceenergy(st["rcell"], st["temp"], w.lbol)
    m.set(emit=emit)
    n = 64
    spent, steps = 0.0, 0
    while spent < target_s and n <= (1 << 16):
        m.set(nvol=po.sourcenumbers(emit, n))
        t0 = time.perf_counter()
        r = m.iteration(77, 1)
        dt = time.perf_counter() - t0
        spent += dt
        steps = r["nstepimc"] + r["nstepddmc"]
        rate = steps / dt
        sample = f"{n} particles of the workload ({steps} particle-steps)"
        n *= 4
    return dict(value=rate, unit="particle-steps/s", cores=threads, kind="port",
                sample=sample, opacity_cell_groups_per_s=opac_rate,
                opacity_sample=f"{len(cells)} of {w.nx} cells x {w.ng} groups x {w.nline} lines")


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by a copy kernel of
    known size) from the committed rocprofv3 counter summary, if it covers this workload."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            rec = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not rec or rec.get("workload") != workload:
        return None
    return rec["fetch_bytes"] + rec["write_bytes"]


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


def run_precision(slgpu, w, st, wl, lines, tables, prec, args):
    """Warm up, then time args.steps whole iterations at the workload's particle count."""
    nstot = 1 << w.n2s
    ctx = slgpu.setup_workload(w, st, wl, lines, tables, precision=prec, seed=2024,
                               npartmax=nstot, errorfatal=False)
    name = "mixed" if prec == slgpu.SLGPU_MIXED else "fp64"
    for k in range(args.warmup):
        r = slgpu.iteration(ctx, st, 1 << min(args.warmup_n2s, w.n2s), 1000 + k, w.lbol, want=False)
        print(f"# {name} warmup {k}: steps {r['transport']['nstepimc'] + r['transport']['nstepddmc']}",
              file=sys.stderr, flush=True)
    ctx.timers(reset=True)
    steps, nerr, last = 0, 0, None
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = slgpu.iteration(ctx, st, nstot, k + 1, w.lbol, want=True)
        tr = r["transport"]
        steps += tr["nstepimc"] + tr["nstepddmc"]
        nerr += tr["nerr"]
        last = r
        print(f"# {name} step {k}: particle-steps {tr['nstepimc'] + tr['nstepddmc']} "
              f"(imc {tr['nstepimc']}, ddmc {tr['nstepddmc']}), max/particle {tr['nstepmax']}, "
              f"escaped {tr['nflux']}, errors {tr['nerr']}", file=sys.stderr, flush=True)
    dt = time.perf_counter() - t0
    tm = ctx.timers()
    ctx.close()
    return dict(steps=steps, dt=dt, tm=tm, nerr=nerr, last=last)


def precision_errors(m, f):
    """FP32-mixed against FP64 on the same inputs and seeds (last timed iteration)."""
    om, of = m["last"]["opacity"], f["last"]["opacity"]
    c0, c1 = of["cap"].astype(np.float64), om["cap"].astype(np.float64)
    sel = c0 > 0
    tm_, tf = m["last"]["transport"], f["last"]["transport"]
    fm, ff = m["last"]["flux"], f["last"]["flux"]
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
    """Differentiation form: the opacity build and its backward (slgpu_opacity_vjp) at the
    workload's size, timed with the library's HIP-event phase timers."""
    from superlite_amd import slgpu, synth
    w, st, wl, lines = synth.build(args.workload)
    tables = synth.load_tables()
    ctx = slgpu.setup_workload(w, st, wl, lines, tables, npartmax=1024, seed=2024)
    rng = np.random.default_rng(7)
    gcap = rng.standard_normal((w.nx, w.ng))
    ones = np.ones(w.nx)
    gas = (st["temp"], st["natom"], st["vol"], st["mass"], st["nelec"], st["natomfr"], st["oc"], st["ginv"])
    for _ in range(args.warmup):
        ctx.opacity_build(*gas, want=())
        ctx.opacity_vjp(gcap, ones, ones)
    ctx.timers(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.opacity_build(*gas, want=())
        ctx.opacity_vjp(gcap, ones, ones)
    dt = time.perf_counter() - t0
    tm = ctx.timers()
    ctx.close()
    K = args.steps
    cg = w.nx * w.ng
    fwd, bwd = tm["opacity"] / K, tm["opacity_vjp"] / K
    print(json.dumps({
        "metric": "opacity cell·groups/s, forward and backward (differentiation form)",
        "value": cg / (bwd * 1e-3), "unit": "cell·groups/s (backward)",
        "n_gpus": 1, "steps": K, "warmup": args.warmup, "ms_per_step": dt / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w.nx} shells x {w.ng} groups x {w.nline} lines, "
                               "forward + VJP w.r.t. T, natom, n_e, mass fractions, populations"},
        "forward_ms": fwd, "backward_ms": bwd, "backward_over_forward": bwd / fwd,
        "forward_cell_groups_per_s": cg / (fwd * 1e-3)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c2")
    ap.add_argument("--precision", default="both", choices=["both", "fp64", "mixed"],
                    help="both: value is the FP32-mixed form, timed beside FP64 on the same seeds")
    ap.add_argument("--warmup-n2s", type=int, default=16,
                    help="log2 particles of the warmup iterations")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--mode", default="transport", choices=["transport", "opacity-grad"],
                    help="opacity-grad: time the opacity build and its backward instead")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="run one elementwise copy of this size first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "superlite_amd runs on one GPU; multi-GPU sharding is designed "
                                   "but not built (run with --gpus 1)"}))
        sys.exit(2)

    if args.mode == "opacity-grad":
        opacity_grad(args)
        return
    if args.calibration_copy_mib > 0:
        calibration_copy(args.calibration_copy_mib)
    from superlite_amd import slgpu, synth
    w, st, wl, lines = synth.build(args.workload)
    tables = synth.load_tables()
    nstot = 1 << w.n2s
    print(f"# workload {args.workload}: nx={w.nx} ng={w.ng} lines={w.nline} "
          f"particles=2^{w.n2s}", file=sys.stderr, flush=True)
    precs = {"both": [slgpu.SLGPU_MIXED, slgpu.SLGPU_FP64], "mixed": [slgpu.SLGPU_MIXED],
             "fp64": [slgpu.SLGPU_FP64]}[args.precision]
    runs = {p: run_precision(slgpu, w, st, wl, lines, tables, p, args) for p in precs}
    main_p = precs[0]
    R = runs[main_p]
    K = args.steps
    value = R["steps"] / R["dt"]
    tm = R["tm"]
    opac_rate = w.nx * w.ng * K / (tm["opacity"] * 1e-3)
    # dominant kernel: k_transport.  Algorithmic bytes per particle-step: the FP32 group
    # opacity (4 B), the step's threshold record (32 B), one sampler-table entry (4 B) and
    # the FP64 read-modify-write of the grd_jrad tally (16 B) = 56 B; per particle, its
    # state read once and written once (5 FP64 in, 5 FP64 + status + group + cell + step
    # count + error out = 40 + 40 + 1 + 4 + 4 + 8 + 4 B = 101 B)
    t_kernel = tm["transport_kernel"] * 1e-3 / K
    bytes_kernel = 56.0 * R["steps"] / K + 101.0 * nstot
    achieved = bytes_kernel / t_kernel / 1e9
    mixed = main_p == slgpu.SLGPU_MIXED
    kname = "k_transport<true>" if mixed else "k_transport<false>"
    out = {
        "metric": "particle-steps/s + opacity cell·groups/s, FP32-mixed vs FP64 at ≤1% tally err",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1, "steps": K, "warmup": args.warmup,
        "ms_per_step": R["dt"] / K * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if mixed else "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w.nx} shells to {w.vmax / 1e5:.0f} km/s, "
                               f"{w.ng} groups {w.wlmin * 1e8:.0f}-{w.wlmax * 1e8:.0f} A, "
                               f"{w.nline} lines, 2^{w.n2s} particles/iteration",
                   "precision": "mixed" if mixed else "fp64"},
        "opacity_cell_groups_per_s": opac_rate,
        "opacity_kernel_cell_groups_per_s": w.nx * w.ng * K / (tm["opacity_kernel"] * 1e-3),
        "transport_errors": R["nerr"],
        "phase_ms_per_step": {k: v / K for k, v in tm.items()},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": achieved / PEAK_HBM_GBS,
                     "traffic": measured_traffic(kname, args.workload),
                     "kernel": kname},
    }
    if slgpu.SLGPU_FP64 in runs and mixed:
        F = runs[slgpu.SLGPU_FP64]
        out["fp64"] = {"value": F["steps"] / F["dt"], "ms_per_step": F["dt"] / K * 1e3,
                       "opacity_cell_groups_per_s": w.nx * w.ng * K / (F["tm"]["opacity"] * 1e-3),
                       "transport_errors": F["nerr"],
                       "phase_ms_per_step": {k: v / K for k, v in F["tm"].items()}}
        out["mixed_speedup_vs_fp64"] = value / out["fp64"]["value"]
        out["mixed_vs_fp64_errors"] = precision_errors(R, F)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, st, wl, lines, tables)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
