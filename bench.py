"""RHMC trajectory seconds (mixed precision vs FP64) and Dslash GB/s on one
MI355X, 32^3x8 lattice.

Workload (one "step" = one RHMC trajectory, all fields resident in HBM):
  32^3x8, Nf=2 staggered, am=0.01, beta=5.6, 8 pseudofermions, 16 poles for
  the heatbath, MD and Metropolis rational functions (Remez on [1e-5,1]),
  Omelyan 2MN with 10 steps of tau=1, links exp(i 0.3 H) seeded,
  momenta/pseudofermions drawn on the device from the seed.  Heatbath and
  Metropolis solves to relative residual 1e-10, MD force solves to 1e-6
  relative.  Mixed precision: MD solves in FP32 (12-real links), heatbath and
  Metropolis solves FP32 + FP64 refinement; Delta H of every mixed trajectory
  is recomputed on the same endpoints with FP64 solves.  The FP64 path runs
  the same trajectories from the same start with every solve in FP64.  The
  value is the mean mixed-precision trajectory time.

Prints ONE JSON line.  --gpus above 1 is refused (one-GPU build).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rhmc-on-gpu_amd"))

from rhmc_amd import Gpu, fields  # noqa: E402
from rhmc_amd.remez import rhmc_approximations  # noqa: E402

HBM_PEAK_GBS = 8000.0


def dslash_bytes(Vh, nrhs, link_b, spin_b):
    """Algorithmic bytes of one half-lattice Dslash: N_out (8 L_b + 2 N_rhs S_b)."""
    return Vh * (8 * link_b + 2 * nrhs * spin_b)


def cpu_baseline(dims, mass, beta, n_pf, shifts, md_a, chi, U, iters_total, n_force, n_q):
    """FP64 oracle (one core) on a bounded sample of the same workload, scaled
    to one trajectory with the FP64 GPU trajectory's own operation counts."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle  # checker, timed here only as the CPU baseline
    o = Oracle(dims)
    k = 4
    t0 = time.perf_counter()
    o.multishift(U, mass, chi, shifts, 0.0, max_cg=k)
    t_it = (time.perf_counter() - t0) / k
    x = np.zeros((len(shifts), o.Vh, 3), np.complex128)
    x[:] = chi
    t0 = time.perf_counter()
    F = o.fermion_force_accum(U, md_a, x)
    t_facc = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.ipdot_total(U, beta, F, True)
    t_ipdot = time.perf_counter() - t0
    P = fields.gauss_momenta(fields.Lattice(*dims), 3)
    t0 = time.perf_counter()
    o.exp_update(U, P, 0.01)
    t_exp = time.perf_counter() - t0
    est = iters_total * n_pf * t_it + n_force * (n_pf * t_facc + t_ipdot) + n_q * t_exp
    sample = (f"oracle FP64 on 32^3x8: {k} multi-shift CG iterations (1 rhs, {len(shifts)} shifts), "
              f"one {len(shifts)}-pole force accumulation, one gauge+TA force, one exp update; scaled by the "
              f"FP64 GPU trajectory's {iters_total} CG iterations x {n_pf} pseudofermions, {n_force} force "
              f"evaluations and {n_q} link updates")
    return est, sample


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-md", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--dims", type=str, default="32,32,32,8")
    ap.add_argument("--npf", type=int, default=8)
    ap.add_argument("--res-md", type=float, default=1e-6)
    ap.add_argument("--res-metro", type=float, default=1e-10)
    ap.add_argument("--dslash-only", action="store_true",
                    help="only the Dslash and copy-probe timings (for counter runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)

    dims = tuple(int(x) for x in args.dims.split(","))
    mass, beta, n_pf, order = 0.01, 5.6, args.npf, 16
    lat = fields.Lattice(*dims)
    U0 = fields.random_links(lat, 1234, 0.3)
    first, md, last = rhmc_approximations(2, n_pf, order, order, order)
    co = dict(first=(first.a0, first.a, first.b), md=(md.a0, md.a, md.b), last=(last.a0, last.a, last.b))
    g = Gpu(dims, mass=mass, beta=beta, n_pf=n_pf, max_order=order)
    g.init0(U0)
    # spectral bounds of MdagM on the start configuration (power iterations)
    if args.dslash_only:
        reps = 50
        res = {}
        for name, prec, nv, lb, sb in (("fp32_12real_8rhs", 1, n_pf, 48, 24), ("fp64_18real_1rhs", 0, 1, 144, 48),
                                       ("fp64_18real_8rhs", 0, n_pf, 144, 48)):
            ms = g.time_dslash(prec, nv, reps)
            res[name] = {"ms": ms, "GBps": dslash_bytes(lat.Vh, nv, lb, sb) / (ms * 1e-3) / 1e9,
                         "bytes": dslash_bytes(lat.Vh, nv, lb, sb)}
        cp = 1 << 30
        ms = g.time_copy(cp, 20)
        res["copy_1GiB"] = {"ms": ms, "GBps": 2 * cp / (ms * 1e-3) / 1e9, "bytes": 2 * cp}
        res["autotune"] = g.tuning_table()
        print(json.dumps(res))
        g.close()
        return
    mx = g.find_max(1e-5, seed=11)
    mn = g.find_min(mx, 1e-5, seed=12)
    mn = max(mn, mass * mass)
    print(f"[bench] spectral bounds {mn:.4e} {mx:.4e}", file=sys.stderr, flush=True)
    traj_kw = dict(n_flavours=2, no_md=args.no_md, integrator="omelyan", res_md=args.res_md,
                   res_metro=args.res_metro, relative=True, bounds=(mn, mx), draw_fields=True,
                   metropolis=False, fp32_threshold=1e-7, **co)

    def run(precision, seed, verify=False):
        g.init0(U0)
        t0 = time.perf_counter()
        r = g.trajectory(precision=precision, seed=seed, verify_fp64=verify, **traj_kw)
        dt = time.perf_counter() - t0
        print(f"[bench] trajectory precision={'mixed' if precision else 'fp64'} seed={seed} {dt:.3f}s "
              f"dH={r.delta_h:.3e} iters={r.cg_iterations} solves={r.n_solves} "
              f"solver={r.seconds_solver:.3f}s force={r.seconds_force:.3f}s", file=sys.stderr, flush=True)
        return dt, r

    for w in range(args.warmup):
        run(1, 100 + w)
    t_mixed, t_fp64, dh_m, dh_d, iters_m, iters_d = [], [], [], [], [], []
    dh_m64 = []
    r_d = None
    for s in range(args.steps):
        dt, r = run(1, 1000 + s)
        t_mixed.append(dt)
        dh_m.append(r.delta_h)
        iters_m.append(r.cg_iterations)
        # same trajectory again (identical), now with the FP64 recomputation of Delta H
        _, rv = run(1, 1000 + s, verify=True)
        dh_m64.append(rv.delta_h_fp64)
        if not args.no_fp64:
            dt, r_d = run(0, 1000 + s)
            t_fp64.append(dt)
            dh_d.append(r_d.delta_h)
            iters_d.append(r_d.cg_iterations)
    val = float(np.mean(t_mixed))
    fp64 = float(np.mean(t_fp64)) if t_fp64 else None

    # Dslash roofline: FP32 12-real links, 8 right-hand sides (the mixed
    # solver's operator) and FP64 18-real, timed with HIP events on the
    # library stream.
    reps = 200
    ms32 = g.time_dslash(1, n_pf, reps)
    ms64 = g.time_dslash(0, 1, reps)
    b32 = dslash_bytes(lat.Vh, n_pf, 48, 24)
    b64 = dslash_bytes(lat.Vh, 1, 144, 48)
    gbs32 = b32 / (ms32 * 1e-3) / 1e9
    gbs64 = b64 / (ms64 * 1e-3) / 1e9

    cpu = None
    if not args.no_cpu and r_d is not None:
        n_force = 2 * args.no_md
        n_q = 2 * args.no_md + 1
        chi = fields.gauss_spinor(lat, 5)[0]
        a0, a, b, _ = md.rescaled(mn, mx, -2.0 / 4.0 / n_pf)
        est, sample = cpu_baseline(dims, mass, beta, n_pf, np.array(b), np.array(a), chi, U0,
                                   int(np.mean(iters_d)), n_force, n_q)
        cpu = {"value": est, "unit": "s/traj", "cores": 1, "kind": "port", "sample": sample}

    out = {
        "metric": "RHMC traj sec + Dslash GB/s (% HBM), mixed vs FP64 on GPU, 32^3x8",
        "value": val,
        "unit": "s/traj",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": val * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic (seeded exp(i 0.3 H) links, device-drawn momenta and pseudofermions)",
        "config": {"workload": f"RHMC trajectory {'x'.join(map(str, dims))}, Nf=2, am=0.01, beta=5.6, "
                               f"{n_pf} pseudofermions x {order} poles, Omelyan {args.no_md} steps, "
                               f"residual 1e-10 (heatbath/Metropolis), {args.res_md:g} (MD), relative",
                   "lattice": list(dims), "n_pf": n_pf, "poles": order, "no_md": args.no_md},
        "fp64_s_per_traj": fp64,
        "mixed_vs_fp64_speedup": (fp64 / val) if fp64 else None,
        "delta_h_mixed": dh_m,
        "delta_h_fp64": dh_d,
        "delta_h_mixed_fp64_recomputed": dh_m64,
        "delta_h_error_vs_fp64_recomputation": [abs(a - b) for a, b in zip(dh_m, dh_m64)],
        "delta_h_abs_diff_mixed_vs_fp64_trajectory": [abs(a - b) for a, b in zip(dh_m, dh_d)],
        "cg_iterations_mixed": iters_m,
        "cg_iterations_fp64": iters_d,
        "spectral_bounds": [mn, mx],
        "dslash": {"fp32_12real_8rhs_ms": ms32, "fp32_12real_8rhs_GBps": gbs32,
                   "fp64_18real_1rhs_ms": ms64, "fp64_18real_1rhs_GBps": gbs64,
                   "frac_hbm_fp32": gbs32 / HBM_PEAK_GBS, "frac_hbm_fp64": gbs64 / HBM_PEAK_GBS},
        "roofline": {"bound": "hbm", "kernel": "k_dslash<float,S12,8 rhs> (half-lattice hop)",
                     "achieved": gbs32, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs32 / HBM_PEAK_GBS,
                     "bytes_per_launch": b32, "traffic": None},
        "cpu_baseline": cpu,
        "dslash_autotune": g.tuning_table(),
    }
    print(json.dumps(out))
    g.close()


if __name__ == "__main__":
    main()
