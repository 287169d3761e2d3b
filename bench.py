"""Benchmark: Mcell-RK3-steps/s of the explicit RK3 step on one MI355X.

Default workload (the metric's): viscous Taylor-Green vortex, Re=1600,
M=0.1, Pr=0.72, Sutherland viscosity, on [0,2 pi]^3 periodic, 256^3 nodes on
a 5% sinusoidally perturbed curvilinear mesh (seeded phases).  --workload
shock_vortex (Mach-2 shock / vortex, 256x128x128, inviscid) and hit
(isotropic turbulence 512^3, viscous) measure the learned and combined forms'
workloads.  One step = global CFL dt reduction + three SSP-RK3 stages (face
fluxes, viscous terms, update, ghost fill), inputs resident in HBM.  Prints
one JSON line: `value` is the plain FP64 form; `forms` adds the FP32 form and
the FP32 + learned-hybrid form with the hybrid's error against FP64.

The roofline object reports the kernel with the largest device time in the
timed region (HIP events on the block's stream); cpu_baseline times the CPU
oracle (C + OpenMP restatement of the reference) on a bounded sample of the
same workload.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0

# Algorithmic work per unit.  Face flux: FP64 operations of one face of the
# reference's method (normal average, Roe basis, 2 x characteristic WENO5,
# LF flux) = 673 add/sub + 795 mul + 102 div + 10 sqrt, counted on the oracle
# with an operation-counting scalar (make -C oracle opcount && oracle/opcount).
FLOPS_PER_FACE = 1580.0
# Algorithmic HBM bytes of the memory-bound kernels (compulsory traffic, each
# value moved once): viscous fluxes read u,v,w,T and the 10 metrics and write
# Ev,Fv,Gv (12) per node of the viscous range (N+4)^3; a divergence pass
# (reference-decomposed calls only: the fused stage's marches take the
# divergence) reads the 12 fluxes and read-modify-writes the 4
# momentum/energy rhs components per updated node; dt reads 6 primitives and
# 9 metrics per node.
VISC_FLUX_BYTES_PER_NODE = (4 + 10 + 12) * 8
VISC_DIV_BYTES_PER_NODE = (12 + 2 * 4) * 8
DT_BYTES_PER_NODE = (6 + 9) * 8


def face_counts(p):
    """Faces per axis the fused stage computes, and updated nodes per axis."""
    n = (p.IM, p.JM, p.KM)
    per = [p.bc[a] == 1 for a in range(3)]
    upd = [n[a] if per[a] else n[a] - 2 for a in range(3)]
    faces = []
    for a in range(3):
        f = n[a] - 1 + (2 if per[a] else 0)
        for b in range(3):
            if b != a:
                f *= upd[b]
        faces.append(f)
    return faces, upd


def kernel_work(name, p):
    """(kind, units of work per launch) for the per-kernel roofline."""
    faces, upd = face_counts(p)
    nodes = upd[0] * upd[1] * upd[2]
    if name == "march_eta":  # eta-faces
        return "flop", faces[1] * FLOPS_PER_FACE
    if name == "xfaces":
        return "flop", faces[0] * FLOPS_PER_FACE
    if name == "march_zeta":  # xi- and zeta-faces (xi fused into the zeta march)
        return "flop", (faces[0] + faces[2]) * FLOPS_PER_FACE
    if name == "visc_flux":
        return "byte", (p.IM + 4) * (p.JM + 4) * (p.KM + 4) * VISC_FLUX_BYTES_PER_NODE
    if name == "visc_div":
        return "byte", nodes * VISC_DIV_BYTES_PER_NODE
    if name == "dt":
        if p.viscous:  # the final pass over the viscous flux pass's maxima
            n = [(p.IM + 4 + 59) // 60, (p.JM + 4 + 3) // 4, (p.KM + 4 + 7) // 8]
            return "byte", n[0] * n[1] * n[2] * 8
        return "byte", p.IM * p.JM * p.KM * DT_BYTES_PER_NODE
    if name == "indicator":  # per axis launch: 6-node stencil + 2 node states read once, 1 B out
        return "byte", sum(faces) // 3 * (5 + 8) * 8
    # ghost shell of 4 layers: Q (5) + the primitives the stage wrote, read
    # and written; averaged over a step (stages 1-2: u, v, w, T on viscous
    # blocks, none on inviscid ones; stage 3: all)
    fields = (2 * 9 + 11) / 3 if p.viscous else (2 * 5 + 10) / 3
    shell = (p.IM + 8) * (p.JM + 8) * (p.KM + 8) - p.IM * p.JM * p.KM
    return "byte", int(shell * fields * 16)


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_tgv256.json")


def measured_traffic(name, p, workload):
    """HBM bytes per launch of kernel `name` from the committed rocprofv3
    FETCH_SIZE + WRITE_SIZE summary of this workload (None if not profiled)."""
    if workload != "tgv" or p.IM != 256 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        t = json.load(f).get("bytes_per_launch", {})
    return t.get(name)


def roofline_of(name, ms, n, p, workload):
    kind, work = kernel_work(name, p)
    avg_s = ms / n / 1e3
    if kind == "flop":
        ach = work / avg_s / 1e12
        return {"bound": "valu_fp64", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / PEAK_FP64_TFLOPS, "traffic": measured_traffic(name, p, workload),
                "kernel": name, "avg_ms": ms / n,
                "note": "FP64 vector-ALU-bound face-flux sweep (no matrix-core work); "
                        "peak = MI355X FP64 vector"}
    ach = work / avg_s / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": ach / PEAK_HBM_GBS, "traffic": measured_traffic(name, p, workload),
            "kernel": name, "avg_ms": ms / n}


def step_hbm(kt, steps, ms_step, p, workload):
    """HBM traffic of the whole step: each kernel's committed rocprofv3
    FETCH_SIZE + WRITE_SIZE bytes per launch times its launches per step
    (profiled pass), over the timed median step time (None if not profiled)."""
    total = 0.0
    for k, (ms, n) in kt.items():
        if not n:
            continue
        t = measured_traffic(k, p, workload)
        if t is None:
            return None
        total += t * n / steps
    gbs = total / (ms_step / 1e3) / 1e9
    return {"bytes_per_step": total, "achieved": gbs, "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": gbs / PEAK_HBM_GBS,
            "note": "measured HBM bytes per step (the viscous flux pass at its bytes without "
                    "the time-step maxima, 6 % fewer than the first stage's) over ms_per_step"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cpu():
    """CPU model and physical core count of this host (lscpu)."""
    info = {"model": None, "physical_cores": None, "logical_cpus": os.cpu_count()}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        kv = dict(line.split(":", 1) for line in out.splitlines() if ":" in line)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name")
        info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except (OSError, subprocess.CalledProcessError, KeyError, ValueError):
        pass
    return info


def cpu_baseline(N=256, steps=3):
    """The oracle on the host cores (bounded sample: TGV N^3, `steps` steps).
    Its OpenMP threads are bound one per physical core (OMP_PROC_BIND=close,
    OMP_PLACES=cores; the runtime reads them when the oracle is loaded)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import pyoracle as po
    from helpers import oracle_block
    from cbns3d_cuda_amd import cases

    case = cases.taylor_green(N)
    o = oracle_block(po, case)
    o.bc_primitive()
    t = 0.0
    t, _ = o.rk3_step(1e9, t)  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        t, _ = o.rk3_step(1e9, t)
    el = time.perf_counter() - t0
    cpu = host_cpu()
    return {"value": N ** 3 * steps / el / 1e6, "unit": "Mcell-RK3-steps/s",
            "cores": po.lib().orc_threads(), "kind": "port",
            "cpu_model": cpu["model"], "physical_cores": cpu["physical_cores"],
            "logical_cpus": cpu["logical_cpus"],
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")},
            "sample": "CPU oracle (C+OpenMP, gcc -O3), viscous TGV %d^3, %d steps after 1 warm-up, "
                      "%.1f s" % (N, steps, el)}


def run_form(case, form, params, thresh, steps, warmup, runs=1, profile_steps=0, reuse=False):
    """`runs` timings of `steps` RK3 steps of one arithmetic form after
    `warmup` steps, with the library's per-kernel event timing off; then
    `profile_steps` more steps with it on for the per-kernel device times.
    (per-run seconds, per-kernel device times, all-finite)."""
    from cbns3d_cuda_amd import cases
    from cbns3d_cuda_amd.block import BUF_Q, F_Q, FORM_HYBRID, Block

    b = Block(case.params)
    try:
        b.calc_metrics(case.x, case.y, case.z, case.period)
        b.set_conservative(cases.q_ref_aos(case))
        b.set_bc_calc_primitive(BUF_Q)
        if form & FORM_HYBRID:
            b.set_indicator(params, thresh, reuse=reuse)
        b.set_form(form)
        t = 0.0
        for _ in range(warmup):
            t, dt = b.rk3_step(1e9, t)
        b.synchronize()
        els = []
        for _ in range(runs):
            b.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                t, dt = b.rk3_step(1e9, t)
            b.synchronize()
            els.append(time.perf_counter() - t0)
        kt = {}
        if profile_steps:
            b.kernel_times(reset=True)
            b.set_profiling(True)
            for _ in range(profile_steps):
                t, dt = b.rk3_step(1e9, t)
            b.synchronize()
            b.set_profiling(False)
            kt = b.kernel_times(reset=True)
        finite = bool(np.isfinite(b.get_field(F_Q)).all())
    finally:
        b.free()
    return els, kt, finite


def rate(cells, steps, els):
    """Median over runs: (Mcell-RK3-steps/s, ms per step)."""
    el = float(np.median(els))
    return cells * steps / el / 1e6, 1e3 * el / steps


def error_fields(err):
    return {k: err[k] for k in ("L2_rho", "Linf_rho", "L2_E", "Linf_E", "troubled_fraction_mean",
                                "diverged_step") if k in err}


def other_forms(case, check_case, args, region=None):
    """FP32 and FP32 + learned hybrid: rates over the same timed steps, and
    each form's rho/E error against the FP64 form after `check_steps` steps
    from the same initial state of `check_case`."""
    from cbns3d_cuda_amd import learned
    from cbns3d_cuda_amd.block import FORM_FP32, FORM_HYBRID

    p = case.params
    cells = p.IM * p.JM * p.KM
    kms = lambda kt: {k: v[0] / max(v[1], 1) for k, v in kt.items() if v[1]}  # noqa: E731
    out = {}
    els, kt, finite = run_form(case, FORM_FP32, None, 0.0, args.steps, args.warmup, args.runs,
                               args.profile_steps)
    v, ms = rate(cells, args.steps, els)
    out["fp32"] = {"value": v, "ms_per_step": ms, "finite": finite, "kernel_ms": kms(kt)}
    if args.check_steps > 0:
        log("bench: fp32 error check, %d steps" % args.check_steps)
        err = learned.compare(check_case, args.check_steps, FORM_FP32, log=log, region=region)
        out["fp32"].update({"check_steps": args.check_steps, "check_case": check_case.name,
                            "check_region": region, **error_fields(err)})
    if not os.path.exists(args.indicator):
        out["fp32_hybrid"] = {"skipped": "no indicator weights at %s" % args.indicator}
        return out
    P, thr, meta = learned.load(args.indicator)
    form = FORM_FP32 | FORM_HYBRID
    for key, reuse in (("fp32_hybrid", False), ("fp32_hybrid_reuse", True)):
        els, kt, finite = run_form(case, form, P, thr, args.steps, args.warmup, args.runs,
                                   args.profile_steps, reuse=reuse)
        v, ms = rate(cells, args.steps, els)
        h = {"value": v, "ms_per_step": ms, "finite": finite, "kernel_ms": kms(kt),
             "indicator": os.path.basename(args.indicator), "threshold": thr}
        if reuse:
            h["indicator_every_steps"] = max(1, int(1.0 / case.params.CFL))
        if args.check_steps > 0:
            log("bench: %s error check, %d steps" % (key, args.check_steps))
            err = learned.compare(check_case, args.check_steps, form, P, thr, log=log, reuse=reuse,
                                  region=region)
            h.update({"check_steps": args.check_steps, **error_fields(err),
                      "meets_L2_rho_1e-6": err["L2_rho"] <= 1e-6})
        out[key] = h
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=("tgv", "shock_vortex", "hit", "slab"), default="tgv",
                    help="tgv: viscous Taylor-Green 256^3 (the metric's workload); "
                         "shock_vortex: Mach-2 shock/vortex 256x128x128 (learned form); "
                         "hit: isotropic turbulence 512^3 (combined form); "
                         "slab: one GPU's share of a 512^3 block cut in 8 slabs, 512x512x64")
    ap.add_argument("--size", type=int, default=0, help="N (default: the workload's own)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-size", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-forms", action="store_true",
                    help="measure only the FP64 plain form")
    ap.add_argument("--indicator", default=os.path.join(ROOT, "cbns3d_cuda_amd",
                                                        "indicator_weights.npz"))
    ap.add_argument("--check-steps", type=int, default=200,
                    help="steps of the FP32 and hybrid error checks against FP64 (0: skip)")
    ap.add_argument("--runs", type=int, default=5,
                    help="timed runs of --steps steps each; the median run is reported")
    ap.add_argument("--profile-steps", type=int, default=0,
                    help="steps of the per-kernel event pass after the timed runs "
                         "(default: --steps)")
    args = ap.parse_args()
    if args.profile_steps <= 0:
        args.profile_steps = args.steps
    if args.gpus != 1:
        print("bench.py: runs on one GPU; the multi-GPU halo exchange is designed, "
              "not a measured path. Use --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from cbns3d_cuda_amd import cases

    wl = args.workload
    region = None
    N = args.size or {"tgv": 256, "shock_vortex": 256, "hit": 512, "slab": 512}[wl]
    if wl == "tgv":
        case = cases.taylor_green(N)
        check = case
        desc = ("viscous Taylor-Green vortex %d^3, Re=1600, M=0.1, curvilinear periodic" % N)
    elif wl == "shock_vortex":
        case = cases.shock_vortex(N, N // 2, N // 2, seed=0)
        check = case
        desc = ("Mach-2 shock / compressible vortex %dx%dx%d, inviscid, clustered curvilinear "
                "mesh" % (N, N // 2, N // 2))
    elif wl == "slab":
        case = cases.smooth_box(N, N, N // 8, viscous=1)
        check = cases.smooth_box(128, 128, 16, viscous=1)
        desc = ("one GPU's share of a %d^3 viscous block in 8 slabs: %dx%dx%d, periodic "
                "(ghost layers stand in for the halo exchange)" % (N, N, N, N // 8))
    else:
        case = cases.isotropic_turbulence(N, seed=0)
        check = case
        region = tuple((0, min(128, n)) for n in (N, N, N))
        desc = ("decaying isotropic turbulence %d^3, Re_lambda=100, Mt=0.6, periodic; FP32 and "
                "hybrid error on the 128^3 subdomain [0,128)^3 of the %d^3 run" % (N, N))
    p = case.params
    cells = p.IM * p.JM * p.KM
    log("bench: building %s" % case.name)
    els, kt, finite = run_form(case, 0, None, 0.0, args.steps, args.warmup, args.runs,
                               args.profile_steps)
    value, ms_step = rate(cells, args.steps, els)
    log("bench: %.3f ms/step (median of %d runs: %s), %.1f Mcell-RK3-steps/s, finite=%s"
        % (ms_step, args.runs, ", ".join("%.3f" % (1e3 * e / args.steps) for e in els), value,
           finite))
    for k, (ms, n) in kt.items():
        if n:
            log("  %-11s %8.3f ms total  %5d launches  %.4f ms avg" % (k, ms, n, ms / n))

    # dominant kernel and its algorithmic work per launch; every kernel's
    # roofline is listed beside it
    dom = max(kt, key=lambda k: kt[k][0])
    roof = roofline_of(dom, kt[dom][0], kt[dom][1], p, wl)
    per_kernel = {k: roofline_of(k, ms, n, p, wl) for k, (ms, n) in kt.items() if n}
    hbm = step_hbm(kt, args.profile_steps, ms_step, p, wl)

    forms = None
    if not args.no_forms:
        forms = other_forms(case, check, args, region)
        for k, v in forms.items():
            log("  form %s: %s" % (k, json.dumps(v)))
    del case, check

    cpu = None
    if not args.no_cpu_baseline and wl == "tgv":
        log("bench: cpu baseline (oracle)")
        cpu = cpu_baseline(args.cpu_size, args.cpu_steps)

    out = {
        "metric": "Mcell-RK3-steps/s at 256^3: FP64 plain vs FP32+learned hybrid @ L2 err<=1e-6",
        "value": value if finite else 0.0,
        "unit": "Mcell-RK3-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "runs": args.runs,
        "ms_per_step": ms_step,
        "ms_per_step_runs": [1e3 * e / args.steps for e in els],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (analytic / seeded initial state and mesh, no dataset)",
        "config": {"workload": desc + "; plain FP64 form", "nodes": cells,
                   "form": "fp64-plain", "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "forms": forms,
        "kernel_ms": {k: v[0] / max(v[1], 1) for k, v in kt.items() if v[1]},
        "kernel_ms_note": "HIP events on the block's stream over %d further steps after the "
                          "timed runs (the timed runs record no events)" % args.profile_steps,
        "kernel_roofline": {k: {"frac": r["frac"], "achieved": r["achieved"], "unit": r["unit"]}
                            for k, r in per_kernel.items()},
        "step_hbm": hbm,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
