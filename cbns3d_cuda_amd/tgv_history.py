"""Kinetic-energy and enstrophy history of the viscous Taylor-Green vortex
(Re = 1600, M = 0.1, Pr = 0.72, Sutherland viscosity, [0, 2 pi]^3 on the 5 %
perturbed curvilinear mesh of cases.taylor_green) on one GPU.

  E_k(t)    volume-averaged rho |u|^2 / 2 (initially 1/8)
  Omega(t)  volume-averaged rho |omega|^2 / 2 (initially 3/8)
  eps(t)    -dE_k/dt by central differences of the recorded history; for the
            incompressible limit eps = 2 Omega / Re, so the difference is the
            scheme's own dissipation
The viscous divergence uses the reference's transverse offset by default
(src/viscous_contribution_kernel.cu:46-65); --centered centres it.

Usage:  python -m cbns3d_cuda_amd.tgv_history --N 128 --t-end 10 --out tgv.json
"""
import argparse
import json
import time

import numpy as np

from . import cases, diagnostics
from .block import BUF_Q, FORM_FP32, Block


def run(N, t_end, dt_out, form=0, max_steps=None, centered=False, log=print):
    case = cases.taylor_green(N, visc_centered=int(centered))
    Re = case.params.Re
    b = Block(case.params)
    try:
        b.calc_metrics(case.x, case.y, case.z, case.period)
        b.set_conservative(cases.q_ref_aos(case))
        b.set_bc_calc_primitive(BUF_Q)
        b.set_form(form)
        hist = []

        def record(t, n):
            prim, met = diagnostics.fields(b)
            ke = diagnostics.kinetic_energy(prim, met)
            om = diagnostics.enstrophy(prim, met)
            hist.append({"t": t, "step": n, "E_k": ke, "Omega": om})
            log("t %8.4f  step %6d  E_k %.6e  Omega %.6e" % (t, n, ke, om))

        t, n = 0.0, 0
        record(t, n)
        t_next = dt_out
        t0 = time.perf_counter()
        while t < t_end - 1e-12 and (max_steps is None or n < max_steps):
            t, _ = b.rk3_step(min(t_next, t_end), t)
            n += 1
            if t >= t_next - 1e-12 or (max_steps is not None and n == max_steps):
                record(t, n)
                t_next += dt_out
        b.synchronize()
        wall = time.perf_counter() - t0
    finally:
        b.free()
    ts = np.array([h["t"] for h in hist])
    ke = np.array([h["E_k"] for h in hist])
    for q, h in enumerate(hist):
        if 0 < q < len(hist) - 1:
            h["eps"] = float(-(ke[q + 1] - ke[q - 1]) / (ts[q + 1] - ts[q - 1]))
            h["eps_2Omega_over_Re"] = 2.0 * h["Omega"] / Re
    return {"N": N, "Re": Re, "visc_centered": int(centered), "steps": n, "t_end": t,
            "wall_s": wall, "history": hist}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--N", type=int, default=128)
    ap.add_argument("--t-end", type=float, default=10.0)
    ap.add_argument("--dt-out", type=float, default=0.25)
    ap.add_argument("--steps", type=int, default=None, help="stop after this many steps")
    ap.add_argument("--fp32", action="store_true")
    ap.add_argument("--centered", action="store_true",
                    help="centred viscous divergence instead of the reference's (-1,-1) offset")
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    res = run(a.N, a.t_end, a.dt_out, FORM_FP32 if a.fp32 else 0, a.steps, a.centered)
    eps = [(h["t"], h["eps"]) for h in res["history"] if "eps" in h]
    if eps:
        tp, ep = max(eps, key=lambda x: x[1])
        print("peak dissipation -dE_k/dt = %.5f at t = %.2f (%d steps, %.1f s)"
              % (ep, tp, res["steps"], res["wall_s"]))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(res, f, indent=1)


if __name__ == "__main__":
    main()
