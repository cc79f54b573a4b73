"""Benchmark: solves/s to relative residual 1e-6 on the 128^3 grid, mixed vs FP64.

Workload: 128^3 cells, cell 2 nm, layered slab (z-bands eps 2.1 / 3.6+0.35i /
-15+1i) with a 4x4 array of embedded spheres (R = 6 cells, eps -10+0.8i),
normal-incidence x-polarised plane wave at 550 nm, zero-padded 256^3
transforms, BiCGSTAB to relative residual 1e-6.  A "step" is one complete
solve from x = 0 with the right-hand side, medium and Green's function
already resident in HBM.  `value` is the mixed-precision solve rate
(FP32-complex BiCGSTAB recurrences and operator, FP64 solution, residual
replaced by the true FP64 residual every 10x reduction, FP64 stop rule) of the
reference's unpreconditioned BiCGSTAB; the plain FP64 solve is reported beside
it, and the key `jacobi` repeats both with the Jacobi preconditioner (same
true-residual stop rule, fewer iterations).  The roofline entry covers the
dominant kernel of the mixed solve (the FP32 x-axis pass: forward transform,
spectral Green's-function product, inverse transform, over the x window of
the internal axis order on its pencil length), timed per launch with HIP
events that its dispatch records on the library's stream, in one more mixed
solve after the timed ones.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
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

METRIC = "solves/s to rel. residual 1e-6 @128³ grid, mixed vs FP64; % of HBM roofline"
HBM_PEAK = 8.0e12  # MI355X HBM3E, B/s
# HBM bytes per launch of the FP32 x pass of the 128^3 workload (x window of
# 48 planes on 64-length pencils), measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE (scaled by a plain copy of known size):
# profiles/k_xreg_f32_128cube_hbm_traffic.md
XPASS_F32_128_TRAFFIC = 168.5e6
XPASS_F32_128_PLAN = {"axis_order": (2, 0, 1), "x_window": (40, 88), "x_pencil": 64, "grid": (128, 128, 128),
                      "y_window": (0, 128), "y_pencil": 128, "z_window": (0, 128), "z_pencil": 128}
# VALU instructions per launch of the same kernel (rocprofv3 SQ_INSTS_VALU, median launch):
# profiles/k_xreg_f32_128cube_valu.md.  The pass is VALU-bound; MI355X issues at most
# 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles = 614.4e9 wave64 VALU instructions per second.
XPASS_F32_128_VALU = 24502272
VALU_ISSUE_PEAK = 256 * 4 * 2.4e9 / 4

# register-resident x-pass kernels by pencil length (FP32), without the last
# template argument: the pencil entries that can hold data (the launcher picks
# n/2 or 3n/4 when the window fits, else n)
XREG_F32 = {32: "k_xreg<float,32,8,XPlan<8,8>,0,2,false", 64: "k_xreg<float,64,16,XPlan<8,4,4>,0,0,false",
            96: "k_xreg<float,96,16,XPlan<4,4,4,3>,0,2,true", 128: "k_xreg<float,128,16,XPlan<16,16>,0,2,true",
            256: "k_xreg<float,256,32,XPlan<16,16,2>,0,0,true"}


def xreg_support(plan):
    """Pencil entries the FP32 x pass's instantiation covers (launch_x)."""
    n = plan["x_pencil"]
    x0, x1 = plan["x_window"]
    need = x1 - x0 if n < plan["grid"][0] else x1
    if n % 4 == 0 and need <= n // 2:
        return n // 2
    if n % 4 == 0 and need <= 3 * n // 4:
        return 3 * n // 4
    return n


def product_bytes(plan, s):
    """Algorithmic bytes of one Krylov product: 81 entries of s bytes per cell
    of the x window (input, output and diagonal; the padded W1/W2 arrays each
    pass reads and writes), the G~ octant of the x pass's pencil length, and
    4 bytes of flags and row ranges per window cell (vie_model_bytes_matvec)."""
    _, ny, nz = plan["grid"]
    vw = (plan["x_window"][1] - plan["x_window"][0]) * ny * nz
    oct_ = (plan["x_pencil"] + 1) * (ny + 1) * (nz + 1)
    return 81.0 * vw * s + oct_ * s + 4.0 * vw


def solve_models(plan, s, iters, outer=0):
    """Solve-level algorithmic bytes of a BiCGSTAB solve (the path's byte model):
    per iteration 2 products + 16 fused 3-component vector streams; a mixed
    solve adds, per residual replacement, one FP64 product and 12 V FP64
    vector entries.  `dense` charges every cell of the grid (81 V s per
    product, 48 V s of vectors); `window` charges the cells of the x window
    that the products and updates actually cover."""
    nx, ny, nz = plan["grid"]
    V = nx * ny * nz
    vw = (plan["x_window"][1] - plan["x_window"][0]) * ny * nz
    oct_full = (nx + 1) * (ny + 1) * (nz + 1)
    dense = iters * (2 * (81.0 * V * s + oct_full * s + 4.0 * V) + 48.0 * V * s)
    window = iters * (2 * product_bytes(plan, s) + 48.0 * vw * s)
    if outer:
        dense += outer * (81.0 * V * 16 + oct_full * 16 + 4.0 * V + 12.0 * V * 16)
        window += outer * (product_bytes(plan, 16.0) + 12.0 * V * 16)
    return dense, window


def xpass_bytes(plan, s):
    """x pass per launch: reads and writes the 3 components of the x window's
    planes of the padded (2Ny x 2Nz) spectrum in place, and reads the G~ octant
    of its pencil length once."""
    _, ny, nz = plan["grid"]
    w = plan["x_window"][1] - plan["x_window"][0]
    oct_ = (plan["x_pencil"] + 1) * (ny + 1) * (nz + 1)
    return 2.0 * 3 * (2 * ny) * (2 * nz) * w * s + oct_ * s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the Jacobi-preconditioned, 96^3 dimer sweep and 256^3 legs")
    ap.add_argument("--xpass-timing", choices=["live", "separate"], default="separate",
                    help="time the FP32 x pass in one more mixed solve after the timed ones (separate), or "
                         "during the timed solves (live: dispatch events on one launch in eight, no graph replay)")
    ap.add_argument("--n", type=int, default=128, help="grid edge (128 = the metric workload)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only time a plain device copy of the x pass's working set (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark measures one GPU; multi-GPU runs are not part of it "
                                   "(independent wavelengths/polarisations shard with no communication)"}))
        sys.exit(2)

    import torch

    from vie_fft_amd import geometry as G
    from vie_fft_amd import viefft as VF

    assert torch.cuda.is_available(), "bench needs a GPU"
    if args.copy_probe:
        copy_probe(torch, args.n)
        return
    sc = G.metric_case() if args.n == 128 else G.layered_case(n=(args.n,) * 3, band_start=args.n // 2 - 24,
                                                               sphere_z=args.n // 2 + 4)
    n = sc.n
    flags = sc.flags()
    s = VF.Solver(n, sc.d, sc.k0, VF.FP64)
    s.set_medium(sc.material(), sc.eps_table, flags)
    s.green_compute()
    s.set_rhs(s.incident(*sc.rhs_angles()))

    def sync():
        torch.cuda.synchronize()

    # ---- warm-up
    for _ in range(args.warmup):
        s.solve_resident(VF.BICGSTAB, tol=args.tol)
        s.solve_mixed(tol=args.tol)

    # ---- timed: FP64 plain solves (no instrumentation inside the timed region)
    sync()
    t0 = time.perf_counter()
    infos64 = [s.solve_resident(VF.BICGSTAB, tol=args.tol) for _ in range(args.steps)]
    sync()
    t64 = (time.perf_counter() - t0) / args.steps
    x64 = s.solution()
    q64 = cross_sections(s, sc)

    # ---- timed: mixed solves (the metric).  The FP32 x pass (the roofline
    # kernel) is timed per launch with HIP events recorded by the dispatch
    # itself (one launch in eight, on the library's stream), by default in one
    # more solve after the timed ones so that the timed solves run
    # uninstrumented (with timing on, iterations are enqueued one by one
    # instead of replayed as graphs)
    live = args.xpass_timing == "live"
    if live:
        s.time_kernel(13)
    sync()
    t0 = time.perf_counter()
    infosm = [s.solve_mixed(tol=args.tol) for _ in range(args.steps)]
    sync()
    tm = (time.perf_counter() - t0) / args.steps
    if not live:
        s.time_kernel(13)
        s.solve_mixed(tol=args.tol)
    n13, ms13 = s.kernel_time()
    s.time_kernel(0)
    xm = s.solution()
    qm = cross_sections(s, sc)

    # ---- the FP64 x pass: one more FP64 solve with the same events (outside
    # the timed regions)
    s.time_kernel(3)
    s.solve_resident(VF.BICGSTAB, tol=args.tol)
    n3, ms3_64 = s.kernel_time()
    s.time_kernel(0)
    tuning = s.tuning()

    it64 = infos64[-1]["iterations"]
    itm = infosm[-1]["iterations"]
    outer = infosm[-1]["outer"]
    plan = s.plan()
    xb = xpass_bytes(plan, 8.0)
    achieved = xb / (ms13 * 1e-3) if ms13 > 0 else 0.0
    nxp = plan["x_pencil"]
    kernel = XREG_F32[nxp] + f",{xreg_support(plan)}>" if nxp in XREG_F32 else f"k_xmix<float,{nxp},B>"
    w = plan["x_window"][1] - plan["x_window"][0]

    out = {
        "metric": METRIC,
        "value": 1.0 / tm,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tm * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 recurrences / c128 solution and residual",
        "data": "synthetic",
        "config": {
            "workload": f"{n[0]}x{n[1]}x{n[2]} layered slab + 4x4 spheres, cell 2 nm, 550 nm x-pol, "
                        f"padded {2*n[0]}x{2*n[1]}x{2*n[2]} transforms, BiCGSTAB to rel. residual {args.tol:g}",
            "grid": list(n),
        },
        "mixed": {"solves_per_s": 1.0 / tm, "ms_per_solve": tm * 1e3, "iterations": itm,
                  "us_per_iteration": tm * 1e6 / itm, "residual_replacements": outer,
                  "true_relres": infosm[-1]["true_relres"]},
        "fp64": {"solves_per_s": 1.0 / t64, "ms_per_solve": t64 * 1e3, "iterations": it64,
                 "us_per_iteration": t64 * 1e6 / it64, "true_relres": infos64[-1]["true_relres"],
                 "xpass_ms": ms3_64, "xpass_frac": xpass_bytes(plan, 16.0) / (ms3_64 * 1e-3) / HBM_PEAK
                 if ms3_64 > 0 else None},
        "speedup_mixed_vs_fp64": t64 / tm,
        "mixed_vs_fp64": {"solution_rel_l2": float(np.linalg.norm(xm - x64) / np.linalg.norm(x64)),
                          "q_ext_rel": abs(qm[0] - q64[0]) / abs(q64[0]),
                          "q_abs_rel": abs(qm[1] - q64[1]) / abs(q64[1]),
                          "q_ext_fp64": q64[0], "q_abs_fp64": q64[1], "q_ext_mixed": qm[0], "q_abs_mixed": qm[1],
                          "note": "both solves to the same FP64 true relative residual 1e-6; the difference is "
                                  "bounded by the two solves' residuals, not by FP32 rounding"},
        "solve_roofline": solve_roofline(plan, tm, itm, outer, t64, it64),
        "plan": {"axis_order": list(plan["axis_order"]), "x_window": list(plan["x_window"]),
                 "y_window": list(plan["y_window"]), "y_pencil": plan["y_pencil"],
                 "z_window": list(plan["z_window"]), "z_pencil": plan["z_pencil"],
                 "x_pencil": nxp, "grid": list(plan["grid"]),
                 "product_bytes_f32": product_bytes(plan, 8.0), "product_bytes_f64": product_bytes(plan, 16.0),
                 "tuning": tuning},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK,
                     "traffic": XPASS_F32_128_TRAFFIC if plan == XPASS_F32_128_PLAN else None,
                     "kernel": f"{kernel} (x pass: forward transform, spectral product, inverse, "
                               f"{w}-plane window on {nxp}-length pencils)",
                     "bytes_per_launch": xb, "mean_ms": ms13, "launches": n13},
    }
    if plan == XPASS_F32_128_PLAN and ms13 > 0:
        out["roofline"]["valu"] = {"insts_per_launch": XPASS_F32_128_VALU, "issue_peak_per_s": VALU_ISSUE_PEAK,
                                   "frac": XPASS_F32_128_VALU / VALU_ISSUE_PEAK / (ms13 * 1e-3)}
    if not args.no_extra:
        _leg(out, "jacobi", lambda: jacobi_leg(s, sc, args, sync, plan, x64, q64, t64))
    # measured practical HBM rate (device copy) beside the published peak
    _, _, copy_gbps = copy_rate(torch, n[0])
    out["roofline"]["copy_GBps"] = copy_gbps
    out["roofline"]["frac_of_copy"] = achieved / 1e9 / copy_gbps
    s.close()
    if not args.no_extra:
        _leg(out, "dimer_sweep_96cube", lambda: dimer_sweep_leg(args.tol))
        _leg(out, "spheres_256cube", spheres_256_leg)
        _leg(out, "small_configs", lambda: small_configs_leg(args.tol), merge=True)
    if not args.no_cpu:
        _leg(out, "cpu_baseline", lambda: cpu_baseline(sc, flags, it64))
    print(json.dumps(out))


def _leg(out, key, fn, merge=False):
    """One more configuration beside the metric: its figures under `key` (or
    merged, for a leg that returns several keys), or the error that stopped it,
    so that the metric's line is printed whatever happens to the others."""
    try:
        r = fn()
    except Exception as e:  # noqa: BLE001 -- reported in the line, not hidden
        out[key] = {"error": f"{type(e).__name__}: {e}"}
        return
    if merge:
        out.update(r)
    else:
        out[key] = r


def jacobi_leg(s, sc, args, sync, plan, x64, q64, t64_plain):
    """The same workload with the Jacobi preconditioner (vie_set_preconditioner:
    right-preconditioned BiCGSTAB on A (m (.) y) = b, m = 1/d_f, x = m (.) y;
    the stop rule reads the same true residual): K timed mixed and K timed
    FP64 solves after one untimed pair, their accuracy against the plain FP64
    solution, and the whole-solve roofline fractions (the preconditioned
    product moves the same bytes: its z forward pass reads m, its z inverse
    pass no longer reads d)."""
    from vie_fft_amd import viefft as VF

    s.set_preconditioner(VF.PRECOND_JACOBI)
    s.solve_mixed(tol=args.tol)
    s.solve_resident(VF.BICGSTAB, tol=args.tol)
    sync()
    t0 = time.perf_counter()
    im = [s.solve_mixed(tol=args.tol) for _ in range(args.steps)]
    sync()
    tm = (time.perf_counter() - t0) / args.steps
    xm = s.solution()
    qm = cross_sections(s, sc)
    sync()
    t0 = time.perf_counter()
    i64 = [s.solve_resident(VF.BICGSTAB, tol=args.tol) for _ in range(args.steps)]
    sync()
    t64 = (time.perf_counter() - t0) / args.steps
    xj = s.solution()
    qj = cross_sections(s, sc)
    s.set_preconditioner(VF.PRECOND_NONE)
    itm, outer, it64 = im[-1]["iterations"], im[-1]["outer"], i64[-1]["iterations"]

    def acc(x, q):
        return {"solution_rel_l2": float(np.linalg.norm(x - x64) / np.linalg.norm(x64)),
                "q_ext_rel": abs(q[0] - q64[0]) / abs(q64[0]), "q_abs_rel": abs(q[1] - q64[1]) / abs(q64[1])}

    return {"mixed": {"solves_per_s": 1.0 / tm, "ms_per_solve": tm * 1e3, "iterations": itm,
                      "us_per_iteration": tm * 1e6 / itm, "residual_replacements": outer,
                      "true_relres": im[-1]["true_relres"], "vs_plain_fp64": acc(xm, qm)},
            "fp64": {"solves_per_s": 1.0 / t64, "ms_per_solve": t64 * 1e3, "iterations": it64,
                     "us_per_iteration": t64 * 1e6 / it64, "true_relres": i64[-1]["true_relres"],
                     "vs_plain_fp64": acc(xj, qj)},
            "speedup_mixed_vs_fp64": t64 / tm,
            "speedup_mixed_vs_plain_fp64": t64_plain / tm,
            "solve_roofline": solve_roofline(plan, tm, itm, outer, t64, it64)}


def _timed(fn, sync, reps):
    """Mean seconds of reps calls of fn after one untimed call, and the last result."""
    fn()
    sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        r = fn()
    sync()
    return (time.perf_counter() - t0) / reps, r


def small_configs_leg(tol):
    """The remaining configurations, each timed on the device with its data
    resident (mean of 3 solves after one untimed):
    * sphere_32cube: the CPU-runnable plain case, 32^3, sphere R = 10 cells,
      eps -4+0.7i, 400 nm, BiCGSTAB to 1e-6 (FP64 and mixed), Q_ext / Q_abs;
    * layered_128x128x64: the plain FP64 case, layered slab + 4x4 spheres,
      550 nm, BiCGSTAB to 1e-6 (FP64 and mixed) and CORS with a divergence
      stop (factor 1e3, at most 2000 iterations: CORS diverges on this medium,
      as the oracle's CORS does);
    * design_64cube: the adjoint-gradient case, 64^3 per-voxel eps, 600 nm:
      forward FP64 solve to 1e-10, then the absorbed power and its gradient
      with respect to all 262144 voxel permittivities (one adjoint solve to
      1e-10 and the fused gradient kernel); and the same with both solves
      mixed precision (vie_solve_mixed_resident, vie_absorption_grad_mixed),
      its gradient against the FP64 one."""
    import torch

    from vie_fft_amd import geometry as G
    from vie_fft_amd import viefft as VF

    sync = torch.cuda.synchronize
    out = {}
    sc = G.sphere_case(32, radius=10.0, eps_phys=-4.0 + 0.7j, delta_nm=2.0, wavelength_nm=400.0)
    with VF.Solver(sc.n, sc.d, sc.k0, VF.FP64) as s:
        s.set_medium(sc.material(), sc.eps_table, sc.flags())
        s.green_compute()
        s.set_rhs_incident(*sc.rhs_angles())
        t64, i64 = _timed(lambda: s.solve_resident(VF.BICGSTAB, tol=tol), sync, 3)
        q = cross_sections(s, sc)
        tm, im = _timed(lambda: s.solve_mixed(tol=tol), sync, 3)
        qm = cross_sections(s, sc)
    out["sphere_32cube"] = {"fp64_ms": t64 * 1e3, "fp64_iterations": i64["iterations"],
                            "fp64_true_relres": i64["true_relres"], "mixed_ms": tm * 1e3,
                            "mixed_iterations": im["iterations"], "mixed_true_relres": im["true_relres"],
                            "q_ext": q[0], "q_abs": q[1], "q_ext_mixed_rel": abs(qm[0] - q[0]) / abs(q[0]),
                            "q_abs_mixed_rel": abs(qm[1] - q[1]) / abs(q[1])}
    sc = G.layered_case()
    with VF.Solver(sc.n, sc.d, sc.k0, VF.FP64) as s:
        s.set_medium(sc.material(), sc.eps_table, sc.flags())
        s.green_compute()
        s.set_rhs_incident(*sc.rhs_angles())
        t64, i64 = _timed(lambda: s.solve_resident(VF.BICGSTAB, tol=tol), sync, 3)
        tm, im = _timed(lambda: s.solve_mixed(tol=tol), sync, 3)
        s.set_divergence_stop(1e3)

        def cors():
            st = s.lib.vie_solve_resident(s.h, VF.CORS, tol, 2000, None)
            return st, s.history()

        tc, (cst, hist) = _timed(cors, sync, 1)
        s.set_divergence_stop(0.0)
        tuning = s.tuning()
    out["layered_128x128x64"] = {"tuning": tuning, "fp64_solves_per_s": 1.0 / t64, "fp64_iterations": i64["iterations"],
                                 "fp64_us_per_iteration": t64 * 1e6 / i64["iterations"],
                                 "fp64_true_relres": i64["true_relres"], "mixed_solves_per_s": 1.0 / tm,
                                 "mixed_iterations": im["iterations"], "mixed_true_relres": im["true_relres"],
                                 "cors_status": VF.STATUS.get(cst, cst), "cors_iterations": len(hist),
                                 "cors_min_relres": float(hist.min()) if len(hist) else None,
                                 "cors_last_relres": float(hist[-1]) if len(hist) else None, "cors_ms": tc * 1e3}
    sc = G.design_case()
    eps, flags = sc.eps(), sc.flags()
    with VF.Solver(sc.n, sc.d, sc.k0, VF.FP64) as s:
        s.set_medium_eps(eps, flags)
        s.green_compute()
        s.set_rhs(s.incident(*sc.rhs_angles()) * flags)

        def grad():
            fwd = s.solve_resident(VF.BICGSTAB, tol=1e-10)
            P, gre, gim, adj = s.absorption_grad(tol=1e-10)
            return fwd, adj, P, gre

        def grad_mixed():
            fwd = s.solve_mixed(tol=1e-10)
            P, gre, gim, adj = s.absorption_grad(tol=1e-10, mixed=True)
            return fwd, adj, P, gre

        tg, (fwd, adj, P, g64) = _timed(grad, sync, 3)
        tm, (fwdm, adjm, Pm, gm) = _timed(grad_mixed, sync, 3)
    out["design_64cube"] = {"forward_and_gradient_ms": tg * 1e3, "forward_iterations": fwd["iterations"],
                            "adjoint_iterations": adj["iterations"], "absorbed_power": P,
                            "voxels": int(np.prod(sc.n)), "tol": 1e-10,
                            "mixed_forward_and_gradient_ms": tm * 1e3,
                            "mixed_forward_iterations": fwdm["iterations"],
                            "mixed_adjoint_iterations": adjm["iterations"],
                            "mixed_adjoint_true_relres": adjm["true_relres"],
                            "mixed_gradient_rel_l2": float(np.linalg.norm(gm - g64) / np.linalg.norm(g64)),
                            "mixed_absorbed_power_rel": abs(Pm - P) / P}
    return out


def cross_sections(s, sc):
    """(Q_ext, Q_abs) of the resident solution: FarField1's forward scattering
    and NearField's Ene_Metal of the spheres, over the geometric area."""
    th, ph, ps = sc.rhs_angles()
    return (s.forward_scatter(th, ph, ps) / sc.area, s.absorbed(sc.absorb_mat) * 120.0 * np.pi / sc.area)


def dimer_sweep_leg(tol):
    """The 96^3 Au-dimer sweep (Drude eps, 64 wavelengths 400-800 nm x 2
    polarisations = 128 right-hand sides, mixed precision to 1e-6) as one
    batched solve: right-hand sides per second of a whole sweep (context,
    per-instance operators and G~, right-hand sides, solve, Q_ext / Q_abs of
    every instance), the median of three sweeps after one untimed sweep (the
    wall time of one sweep varies by about 20% from call to call; its solve
    and set-up phases do not)."""
    from vie_fft_amd import geometry as G
    from vie_fft_amd import sweep as SW

    points = G.dimer_sweep(64)
    scene = lambda lam, psi: G.dimer_case(lam, psi=psi)  # noqa: E731
    SW.run_sweep_batched(scene, points, tol=tol, mixed=True)
    runs = []
    for _ in range(3):
        tm = {}
        t0 = time.perf_counter()
        res = SW.run_sweep_batched(scene, points, tol=tol, mixed=True, timings=tm)
        runs.append((time.perf_counter() - t0, tm, res))
    runs.sort(key=lambda r: r[0])
    wall, tm, res = runs[1]
    its = np.array([r["iterations"] for r in res])
    return {"rhs": len(res), "rhs_per_s": len(res) / wall, "wall_s": wall,
            "wall_s_of_3": [r[0] for r in runs], "solve_s": tm["solve_s"],
            "solve_rhs_per_s": len(res) / tm["solve_s"], "setup_s": tm["setup_s"], "post_s": tm["post_s"],
            "instances_per_batch": tm["chunk"], "iterations_min": int(its.min()),
            "iterations_median": float(np.median(its)), "iterations_max": int(its.max()),
            "all_converged": all(r["converged"] for r in res),
            "max_true_relres": max(r["true_relres"] for r in res)}


def spheres_256_leg(tol=1e-8):
    """256^3 grid, 60 seeded random spheres (cell 1.5 nm, 550 nm), mixed
    precision to the configuration's true residual 1e-8: every axis window is
    the whole grid, so the passes run at their full lengths.  One untimed
    mixed solve, then 2 timed mixed solves, 1 timed FP64 solve and 1 timed
    Jacobi-preconditioned mixed solve; per-iteration times, the dense-model
    fraction of 8 TB/s and the mixed solution's error against the FP64 one."""
    from vie_fft_amd import geometry as G
    from vie_fft_amd import viefft as VF

    sc = G.random_spheres_case()
    with VF.Solver(sc.n, sc.d, sc.k0, VF.FP64) as s:
        s.set_medium(sc.material(), sc.eps_table, sc.flags())
        s.green_compute()
        s.set_rhs_incident(*sc.rhs_angles())
        s.solve_mixed(tol=tol)
        t0 = time.perf_counter()
        im = [s.solve_mixed(tol=tol) for _ in range(2)]
        tm = (time.perf_counter() - t0) / 2
        xm = s.solution()
        t0 = time.perf_counter()
        i64 = s.solve_resident(VF.BICGSTAB, tol=tol)
        t64 = time.perf_counter() - t0
        x64 = s.solution()
        s.set_preconditioner(VF.PRECOND_JACOBI)
        t0 = time.perf_counter()
        ij = s.solve_mixed(tol=tol)
        tj = time.perf_counter() - t0
        xj = s.solution()
        s.set_preconditioner(VF.PRECOND_NONE)
        plan = s.plan()
        tuning = s.tuning()
    dm, _ = solve_models(plan, 8.0, im[-1]["iterations"], im[-1]["outer"])
    d6, _ = solve_models(plan, 16.0, i64["iterations"])
    nrm = float(np.linalg.norm(x64))
    return {"tol": tol, "mixed_s": tm, "mixed_iterations": im[-1]["iterations"],
            "mixed_replacements": im[-1]["outer"],
            "mixed_vs_fp64_rel_l2": float(np.linalg.norm(xm - x64)) / nrm,
            "jacobi_mixed_s": tj, "jacobi_mixed_iterations": ij["iterations"],
            "jacobi_mixed_true_relres": ij["true_relres"],
            "jacobi_mixed_vs_fp64_rel_l2": float(np.linalg.norm(xj - x64)) / nrm,
            "mixed_us_per_iteration": tm * 1e6 / im[-1]["iterations"], "mixed_true_relres": im[-1]["true_relres"],
            "fp64_s": t64, "fp64_iterations": i64["iterations"], "fp64_us_per_iteration": t64 * 1e6 / i64["iterations"],
            "fp64_true_relres": i64["true_relres"], "mixed_dense_frac": dm / tm / HBM_PEAK,
            "fp64_dense_frac": d6 / t64 / HBM_PEAK, "x_window": list(plan["x_window"]), "x_pencil": plan["x_pencil"],
            "tuning": tuning}


def solve_roofline(plan, tm, itm, outer, t64, it64):
    """Whole-solve % of the HBM roofline: model bytes / solve time / 8 TB/s,
    for the dense model (every cell; above 1 means faster than a dense solve
    could be at peak bandwidth) and for the window the solve covers."""
    dm, wm = solve_models(plan, 8.0, itm, outer)
    d6, w6 = solve_models(plan, 16.0, it64)
    return {"mixed": {"dense_bytes": dm, "dense_frac": dm / tm / HBM_PEAK,
                      "window_bytes": wm, "window_frac": wm / tm / HBM_PEAK},
            "fp64": {"dense_bytes": d6, "dense_frac": d6 / t64 / HBM_PEAK,
                     "window_bytes": w6, "window_frac": w6 / t64 / HBM_PEAK}}


def copy_rate(torch, n):
    """Device-to-device copy of 3 x 4 n^3 complex64 (the FP32 x pass's array),
    20 times: (bytes per copy, ms per copy, read + written GB/s)."""
    nbytes = 3 * 4 * n ** 3 * 8
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").uniform_()
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(20):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / 20
    del a, b
    return nbytes, ms, 2 * nbytes / (ms * 1e-3) / 1e9


def copy_probe(torch, n):
    """The copy of copy_rate alone: the reference kernel for scaling the
    FETCH_SIZE / WRITE_SIZE counters and the practical HBM copy rate."""
    nbytes, ms, gbps = copy_rate(torch, n)
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "ms": ms, "GB_per_s": gbps}))


def _component_seconds(sc, flags):
    """Seconds for one X_FFT-equivalent component product of the CPU oracle in
    the reference schedule (one core)."""
    from oracle import oracle as O

    op = O.Operator(sc.n, sc.d, sc.k0, sc.eps(), flags, O.SCHED_REFERENCE)
    rng = np.random.default_rng(0)
    x = (rng.standard_normal((3,) + sc.n) + 1j * rng.standard_normal((3,) + sc.n)) * flags
    t = op.time_component(x, 0)
    del op
    return t


def _cpu_child():
    """Child process of the all-cores leg: one independent component product."""
    from vie_fft_amd import geometry as G

    sc = G.metric_case()
    print(json.dumps({"t_comp": _component_seconds(sc, sc.flags())}), flush=True)


def _cpu_model():
    """model name, family/model/stepping and MHz of the first CPU in /proc/cpuinfo."""
    keys = {"model name": None, "cpu family": None, "model": None, "stepping": None, "cpu MHz": None}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" not in line:
                if all(v is not None for v in keys.values()):
                    break
                continue
            k, v = (t.strip() for t in line.split(":", 1))
            if k in keys and keys[k] is None:
                keys[k] = v
    except OSError:
        return "unknown"
    return (f"{keys['model name']} (family {keys['cpu family']}, model {keys['model']}, "
            f"stepping {keys['stepping']}, {keys['cpu MHz']} MHz)")


def cpu_baseline(sc, flags, iterations):
    """Time the CPU oracle in the reference's schedule on a bounded sample:
    one output component of one product (3 forward + 3 inverse padded FFTs and
    3 per-element symbol loops, the work of one X_FFT call,
    vie_fft/CG_Iteration.h:131-143); a BiCGSTAB iteration is 6 such calls (two
    products), a solve is `iterations` iterations.  Two legs: one core (the
    reference's MKL-sequential solve, README.md:37), and K cores each running
    an independent sample at once (its frequency-farm mode, vie_fft/main.cpp:27-36),
    K = min(16, visible cores) child processes (16: one GPU's share of the host)."""
    import subprocess

    t_comp = _component_seconds(sc, flags)
    t_solve = 6.0 * t_comp * iterations
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    # the GPU host gives one GPU's job a share of 16 cores: the all-cores leg
    # runs one process per core of that share (all of them on smaller hosts)
    k = max(1, min(16, nproc))
    here = os.path.dirname(os.path.abspath(__file__))
    procs = [subprocess.Popen([sys.executable, "-c", "from bench import _cpu_child; _cpu_child()"], cwd=here,
                              stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True) for _ in range(k)]
    ts = []
    for p in procs:
        out, _ = p.communicate()
        if p.returncode == 0 and out.strip():
            ts.append(json.loads(out.strip().splitlines()[-1])["t_comp"])
    par = None
    if len(ts) == k:
        t_par = 6.0 * float(np.mean(ts)) * iterations
        par = {"value": k / t_par, "unit": "solves/s", "cores": k, "host_cores": nproc,
               "cpu_count": os.cpu_count(),
               "sample": f"{k} concurrent processes, one component product each: mean {np.mean(ts):.2f} s "
                         f"(max {max(ts):.2f} s); throughput = {k} solves per extrapolated solve time; "
                         f"{k} = the 16-core CPU share of one GPU's job (of {nproc} cores visible)"}
    return {"value": 1.0 / t_solve, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"one X_FFT-equivalent component product at {sc.n[0]}^3 (reference 9+9 FFT schedule, "
                      f"per-element trig) took {t_comp:.2f} s on 1 core; solve extrapolated as 6 x that x "
                      f"{iterations} BiCGSTAB iterations ({t_solve:.0f} s); set-up not counted",
            "cpu_model": _cpu_model(), "all_cores": par}


if __name__ == "__main__":
    main()
