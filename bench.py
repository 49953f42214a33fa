"""Benchmark: cells/s to chemistry + thermal steady state on one MI355X.

Workload (the single-GPU configuration of the headline metric): a 64^3
log-normal density cube (seeded Gaussian random field, P(k) ~ k^-11/3,
median n_H = 1e4 cm^-3, sigma_ln = 1.0, seed 42), 0.5 pc box, G0 = 1e4
Habing incident from the -x face over 48 rays (nside 2), zeta_CR = 2e-16 s^-1,
the bundled 31-species / 75-reaction network, initial H atomic, C+ 1.4e-4,
O 3.0e-4, e- = C+, T0 = 50 K, all FP64.

One step = one solve of the whole grid to the coupled steady state: the
RT -> shielding -> cosmic rays -> dust -> chemistry -> thermal fixed point
iterated from the initial state until the upstream convergence test passes
(p95 dT/T, p99 dx_H2, p99 dx_CO < 0.05 after > 2 iterations, solver_3d.py:
381-436).  Chemistry in every iteration is the upstream steady-state solve
(chemistry/solver.py:145-225, as refine() calls it): each cell continues
from its current abundances to t_max = min(3e17, max(1e16, 3e17/n_H)) s with
the GPU BDF at rtol 1e-9 (abundances within 1e-6 of the tightly integrated
upstream RHS).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cells/s to chem+thermal steady state at ≤1e-6 rel err; adjoint/forward time"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
N_SP, N_RX = 31, 75


def flops_per_cell(stats):
    """Algorithmic FP64 operations of the BDF solve from its counters.
    LU of the 31x31 iteration matrix 2n^3/3, triangular solves 2n^2, RHS
    ~8 per reaction + 1 per stoichiometric term, Jacobian 2 per term plus
    the matrix build n^2, per-step difference updates ~40 n."""
    n = N_SP
    steps, lu, rhs, jac = [stats[:, k].astype(np.float64) for k in range(4)]
    return (lu * (2.0 * n ** 3 / 3.0 + 2.0 * n * n) + rhs * (2.0 * n * n + 8 * N_RX + 300)
            + jac * (2.0 * 300 + 4 * N_RX) + steps * 40.0 * n)


def make_solver(n=64, variant=0, waves_per_cu=0, autotune=False):
    from prism3d_amd.density import PC, lognormal_median
    from prism3d_amd.grid import directional_G0
    from prism3d_amd.solver3d import PDRSolverGPU
    dens, _ = lognormal_median(n, 0.5 * PC, 1e4, 1.0, seed=42)
    g0r = directional_G0(1e4, [-1.0, 0.0, 0.0], nside=2)
    return PDRSolverGPU(dens, 0.5 * PC, G0_external=g0r, zeta_CR_0=2e-16, nside_rt=2,
                        chemistry='network', x_c_total=1.4e-4, x_o_total=3.0e-4,
                        initial_abundances={'H': 1.0, 'C+': 1.4e-4, 'O': 3.0e-4, 'e-': 1.4e-4},
                        T_init=50.0, chem_atol=1e-20, chem_variant=variant,
                        chem_waves_per_cu=waves_per_cu, chem_autotune=autotune)


def chem_form(variant, n_cells):
    """Which chemistry kernel p3d_chem_solve runs for this configuration."""
    import torch
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    if variant == 4 or (variant == 0 and n_cells >= cus * 8 * 64):
        return "batched: persistent one-cell-per-lane kernel with a device work queue"
    return "plain drop-in: one cell per 64-lane workgroup"


def profiled_traffic(args, lane):
    """HBM bytes per chemistry launch of this workload as measured by rocprofv3
    (FETCH_SIZE / WRITE_SIZE, calibrated; profiles/chem_64cube_traffic.json),
    or None when no profile of this configuration is committed."""
    path = os.path.join(ROOT, "profiles", "chem_64cube_traffic.json")
    if not (lane and args.n == 64 and args.variant in (0, 4) and args.waves_per_cu in (0, 8)
            and os.path.exists(path)):
        return None
    with open(path) as f:
        rec = json.load(f)
    return float(sum(rec["hbm_bytes_per_launch"].values()))


def solve(solver, max_iter):
    solver.reset()
    conv = solver.run(max_iterations=max_iter, convergence_tol=0.05, verbose=True)
    return conv, len(solver.convergence_history)


def cpu_baseline(solver, recorded, budget_s=20.0):
    """The oracle (C restatement, OpenMP over cells) replaying the GPU's
    per-iteration chemistry calls on the sampled cells (same inputs, same
    tolerances), plus the thermal root-find; reported as cells/s to steady
    state.  The sample is cut to about budget_s of CPU time."""
    from oracle import oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n_all = recorded[0]['n_H'].size
    take = n_all
    t_chem = 0.0
    while True:
        sel = np.arange(0, n_all, max(1, n_all // take))[:take]
        t0 = time.perf_counter()
        for r in recorded:
            O.solve_steady_batch(r['y'][:, sel], r['n_H'][sel], r['T'][sel], r['G0'][sel],
                                 r['A_V'][sel], r['zeta'][sel], r['fsh_H2'][sel],
                                 r['fsh_CO'][sel], t_max=r['t_max'], n_checkpoints=2,
                                 rtol=solver.chem_rtol, atol=solver.chem_atol, early_stop=False,
                                 post_mode=-1)
            if time.perf_counter() - t0 > budget_s and sel.size > 16:
                break
        t_chem = time.perf_counter() - t0
        if t_chem <= budget_s or sel.size <= 16:
            break
        take = max(16, int(sel.size * budget_s / t_chem / 2))
    rate = sel.size / t_chem
    return {"value": rate, "unit": "cells/s", "cores": cores, "kind": "port",
            "sample": f"{sel.size} strided cells of the {solver.n_cells}-cell grid: the GPU run's "
                      f"{len(recorded)} per-iteration chemistry solves (31 species, BDF rtol "
                      f"{solver.chem_rtol:g}) replayed by the C oracle in {t_chem:.2f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--n", type=int, default=64, help="grid cells per side")
    ap.add_argument("--max-iter", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--variant", type=int, default=0, help="chemistry kernel variant (0 auto)")
    ap.add_argument("--autotune", action="store_true",
                    help="pick the chemistry kernel configuration online (prism3d_amd.autotune)")
    ap.add_argument("--waves-per-cu", type=int, default=0,
                    help="resident wavefronts per CU of the persistent kernel (0 default)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this size (a known-traffic "
                         "kernel that calibrates profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the per-cell "
              "solve shards by cell range with no exchange; see the design notes)",
              file=sys.stderr)
        sys.exit(2)

    import torch
    if args.calib_copy_mib > 0:
        nb = args.calib_copy_mib * 2 ** 20 // 8
        a = torch.ones(nb, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    solver = make_solver(args.n, args.variant, args.waves_per_cu, args.autotune)
    print(f"[bench] {args.n}^3 = {solver.n_cells} cells; warmup {args.warmup}", file=sys.stderr,
          flush=True)
    for w in range(args.warmup):
        conv, iters = solve(solver, args.max_iter)
        print(f"[bench] warmup {w}: {iters} iterations, converged={conv}", file=sys.stderr,
              flush=True)
    solver.record_chem = True
    solver.chem_records = []
    if not args.no_cpu:
        # keep the chemistry inputs of a strided cell sample (for the CPU
        # replay); the copies are a few MB per iteration
        solver.record_inputs = np.linspace(0, solver.n_cells - 1, 4096).astype(np.int64)
        solver.recorded = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its = []
    for s in range(args.steps):
        conv, iters = solve(solver, args.max_iter)
        its.append(iters)
        print(f"[bench] step {s}: {iters} iterations, converged={conv}", file=sys.stderr,
              flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = wall / args.steps * 1e3
    value = solver.n_cells / (wall / args.steps)

    durs = [a.elapsed_time(b) * 1e-3 for a, b, _ in solver.chem_records]
    fl = [flops_per_cell(st.cpu().numpy()).sum() for _, _, st in solver.chem_records]
    steps_mean = float(np.mean([st[:, 0].float().mean().item() for *_x, st in solver.chem_records]))
    avg_dur = float(np.mean(durs))
    achieved = float(np.mean(fl)) / avg_dur / 1e12
    form = chem_form(args.variant, solver.n_cells)
    lane = form.startswith("batched")
    roof = {"bound": "mfma", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP64_PEAK_TFLOPS, "traffic": profiled_traffic(args, lane),
            "kernel": ("chem_lane_kernel + hand-over chem_solve_kernel (one p3d_chem_solve call)"
                       if lane else "chem_solve_kernel"),
            "avg_launch_ms": avg_dur * 1e3, "launches": len(durs),
            "bdf_steps_per_cell": steps_mean}
    out = {"metric": METRIC, "value": value, "unit": "cells/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic",
           "config": {"workload": f"{args.n}^3 log-normal cube (median n_H 1e4, sigma_ln 1.0, "
                                  "0.5 pc), G0 1e4 from -x, 48 rays, zeta 2e-16, 31-species "
                                  "network, steady-state BDF (t_max 1e16-3e17 s) at rtol 1e-9, "
                                  "coupled thermal",
                      "cells": solver.n_cells, "fixed_point_iterations": its,
                      "chem_form": form},
           "adjoint_over_forward": None,
           "roofline": roof}
    if not args.no_cpu:
        rec = solver.recorded[:its[0]]  # the first timed solve's iterations
        out["cpu_baseline"] = cpu_baseline(solver, rec)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
