"""Wavelength and polarisation sweeps over one grid, and the reference's result files.

The reference runs one process per wavelength under an MPI farm
(``vie_fft/main.cpp:27-36``), each process writing ``Ene_Metal{m}.txt``,
``Ene_Active{m}.txt`` and ``Grid{m}.txt`` (``vie_fft/NearFar.h:273-280``), and
rank 1 merges them (``vie_fft/main.cpp:55-100``).  Here one GPU context stays
resident across the sweep: per wavelength the wavenumber, the dispersive
permittivities and G~ are updated in place (``vie_set_k0``,
``vie_set_medium``, ``vie_green_compute``), per polarisation only the
right-hand side changes.  Permittivity tables are read in the format of the
reference's ``data/*_epr.txt`` (``vie_fft/ReadFile.h:23-45``).

    python -m vie_fft_amd.sweep [--count 64] [--n 96] [--fp64] [--out DIR] [--sequential]

runs the 96^3 Au-dimer sweep (Drude permittivity, 64 wavelengths from 400 to
800 nm, two polarisations) as one batched solve (``run_sweep_batched``; with
--sequential one solve after the other, ``run_sweep``) and prints one JSON line
per right-hand side and a summary line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from vie_fft_amd import geometry as G  # noqa: E402
    from vie_fft_amd import viefft as VF  # noqa: E402
else:
    from . import geometry as G
    from . import viefft as VF


def read_eps_table(path, count=41):
    """Permittivity rows ``re im`` (engineering sign, as the reference's files
    store them) for wavelengths ``400 + 10 m`` nm, ``m < count``
    (``vie_fft/ReadFile.h:33-40``)."""
    rows = []
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 2:
                rows.append(complex(float(parts[0]), float(parts[1])))
            if len(rows) == count:
                break
    if len(rows) < count:
        raise ValueError(f"{path}: {len(rows)} rows, {count} expected")
    return np.array(rows, dtype=np.complex128)


def reference_wavelengths(count=41, begin=400, step=10):
    """Incident wavelengths in nm (``vie_fft/ReadFile.h:6-7,39``)."""
    return [begin + m * step for m in range(count)]


def active_face_volume(material, mat, d):
    """Absorb_Volume of NearField: cell volume times the number of faces with
    i, j, k < N - 1 whose two cells both carry `mat` (``vie_fft/NearFar.h:186-250``)."""
    m = np.asarray(material) == mat
    core = m[:-1, :-1, :-1]
    cnt = (core & m[1:, :-1, :-1]).sum() + (core & m[:-1, 1:, :-1]).sum() + (core & m[:-1, :-1, 1:]).sum()
    return float(cnt) * d[0] * d[1] * d[2]


def write_point_files(out_dir, m, wavelength_nm, q_ext, q_abs, active_density, n, d):
    """Ene_Metal{m}.txt, Ene_Active{m}.txt and Grid{m}.txt of one wavelength,
    in the reference's formats (``vie_fft/NearFar.h:273-280``)."""
    os.makedirs(out_dir, exist_ok=True)
    wl = int(round(wavelength_nm))
    with open(os.path.join(out_dir, f"Ene_Metal{m}.txt"), "w") as f:
        f.write("%d %.15f %.15f\n" % (wl, q_ext, q_abs))
    with open(os.path.join(out_dir, f"Ene_Active{m}.txt"), "w") as f:
        f.write("%d %.15f\n" % (wl, active_density))
    with open(os.path.join(out_dir, f"Grid{m}.txt"), "w") as f:
        f.write("%d %d %d %d %.15f %.15f %.15f\n" % (m, n[0], n[1], n[2], d[0], d[1], d[2]))


def merge_point_files(out_dir, count):
    """Concatenate the per-wavelength files into Ene_Metal.txt, Ene_Active.txt
    and Grid.txt (``vie_fft/main.cpp:55-100``) and remove the pieces."""
    for stem in ("Ene_Metal", "Ene_Active", "Grid"):
        with open(os.path.join(out_dir, f"{stem}.txt"), "w") as out:
            for m in range(count):
                p = os.path.join(out_dir, f"{stem}{m}.txt")
                with open(p) as f:
                    out.write(f.read())
                os.remove(p)


def scene_voxels(s, sc):
    """(material, flags) of a scene: its analytic shapes voxelised on the
    device (vie_voxelize: byte-equal to geometry.py's rules, a few
    milliseconds at 96^3), or the host rules for per-voxel media."""
    from . import geometry as G

    if sc.eps_voxel is None and all(isinstance(sh, (G.Sphere, G.Band)) for sh in sc.shapes):
        return s.voxelize(sc)
    return sc.material(), sc.flags()


def run_sweep(scene_at, points, tol=1e-6, mixed=True, method=VF.BICGSTAB, metal=1, active=None, maxit=0,
              out_dir=None, log=None, warm=False):
    """Solve every (wavelength_nm, psi) of `points` on one resident context.

    scene_at(wavelength_nm, psi) -> geometry.Scene (same grid and cells for
    every point).  Returns one dict per point: Q_ext, Q_abs (metal id
    `metal`), the active-material absorption density (id `active`, if given),
    iterations, residual replacements, true residual and seconds; the first
    point of each wavelength also holds the context's plan (Solver.plan).  With
    out_dir the per-wavelength files are written (first polarisation of each
    wavelength, as the reference writes one file per process) and merged.
    With warm=True every solve starts from the previous wavelength's solution
    of the same polarisation instead of x = 0: the points are then visited
    polarisation by polarisation (wavelengths in their given order), so that
    guess is the solution still resident on the device (``vie_set_guess`` with
    no host copy); results come back in the order of `points`.
    """
    results = [None] * len(points)
    s = None
    cur = None
    material = flags = None
    wl_index = {}
    order = list(range(len(points)))
    if warm:
        first = {}
        for i, (_, psi) in enumerate(points):
            first.setdefault(psi, i)
        order.sort(key=lambda i: (first[points[i][1]], i))
    last_psi = None  # polarisation of the resident solution (warm starts)
    wl_plan = {}
    try:
        for idx in order:
            lam, psi = points[idx]
            sc = scene_at(lam, psi)
            if s is None:
                s = VF.Solver(sc.n, sc.d, sc.k0, VF.FP64)
                material, flags = scene_voxels(s, sc)
                vol_active = active_face_volume(material, active, sc.d) if active is not None else 0.0
            if cur != lam:
                t0 = time.perf_counter()
                if cur is not None:
                    s.set_k0(sc.k0)
                s.set_medium(material, sc.eps_table, flags)
                s.green_compute()
                cur = lam
                setup = time.perf_counter() - t0
                plan = s.plan()
            else:
                setup = 0.0
                plan = None
            t0 = time.perf_counter()
            s.set_rhs_incident(sc.theta, sc.phi, psi)
            if warm and last_psi == psi:
                s.set_guess(None)
            info = s.solve_mixed(tol=tol, maxit=maxit) if mixed else s.solve_resident(method, tol=tol, maxit=maxit)
            solve_s = time.perf_counter() - t0
            last_psi = psi
            q_ext = s.forward_scatter(sc.theta, sc.phi, psi) / sc.area
            q_abs = s.absorbed(metal) * 120.0 * np.pi / sc.area
            act = 3.0 * s.absorbed(active) / vol_active if active is not None and vol_active > 0 else 0.0
            r = {"wavelength_nm": float(lam), "psi": float(psi), "q_ext": q_ext, "q_abs": q_abs,
                 "active_density": act, "iterations": info["iterations"], "outer": info["outer"],
                 "true_relres": info["true_relres"], "converged": bool(info["converged"]),
                 "solve_s": solve_s, "setup_s": setup}
            if plan is not None and "plan" not in wl_plan.get(lam, {}):
                r["plan"] = plan
                wl_plan[lam] = r
            results[idx] = r
            if log:
                log(r)
            if out_dir is not None and lam not in wl_index:
                wl_index[lam] = len(wl_index)
                write_point_files(out_dir, wl_index[lam], lam, q_ext, q_abs, act, sc.n, sc.d)
    finally:
        if s is not None:
            s.close()
    if out_dir is not None:
        merge_point_files(out_dir, len(wl_index))
    return results


def run_sweep_batched(scene_at, points, tol=1e-6, mixed=True, method=VF.BICGSTAB, metal=1, active=None, maxit=0,
                      out_dir=None, log=None, chunk=None, timings=None):
    """run_sweep with every point an instance of one batched solve.

    One context holds the geometry; each (wavelength_nm, psi) point becomes an
    instance with its own wavenumber and permittivity table
    (``vie_batch_set_operators``: G~ for all wavelengths in one set of
    launches) and its own masked incident field, and the instances are solved
    together (``vie_batch_solve``) in batches of at most `chunk` (default: as
    many as the device holds).  Same result records as run_sweep; solve_s and
    setup_s are the batch's times shared evenly.  `timings` (a dict) receives
    the set-up, solve and post-processing seconds.
    """
    scenes = [scene_at(lam, psi) for lam, psi in points]
    sc0 = scenes[0]
    results = []
    wl_index = {}
    t_setup = t_solve = t_post = 0.0
    with VF.Solver(sc0.n, sc0.d, sc0.k0, VF.FP64) as s:
        material, flags = scene_voxels(s, sc0)
        vol_active = active_face_volume(material, active, sc0.d) if active is not None else 0.0
        t0 = time.perf_counter()
        s.set_medium(material, sc0.eps_table, flags)
        s.green_compute()
        t_setup += time.perf_counter() - t0
        plan = s.plan()
        cap = chunk or max(1, min(len(points), s.batch_capacity()))
        for b0 in range(0, len(points), cap):
            part = scenes[b0:b0 + cap]
            nb = len(part)
            t0 = time.perf_counter()
            s.batch_reserve(nb)
            s.batch_set_operators(0, [sc.k0 for sc in part], np.stack([sc.eps_table for sc in part]))
            for i, (sc, (lam, psi)) in enumerate(zip(part, points[b0:b0 + nb])):
                s.batch_set_rhs_incident(i, sc.theta, sc.phi, psi)
            t1 = time.perf_counter()
            _, infos = s.batch_solve(nb, method=method, tol=tol, maxit=maxit, mixed=mixed)
            t2 = time.perf_counter()
            # every instance's outputs with one read-back
            angles = ([sc.theta for sc in part], [sc.phi for sc in part], [psi for _, psi in points[b0:b0 + nb]])
            fw, ab = s.batch_outputs(*angles, mat=metal)
            if active is not None and vol_active > 0:
                _, aa = s.batch_outputs(*angles, mat=active)
            for i, (sc, (lam, psi)) in enumerate(zip(part, points[b0:b0 + nb])):
                q_ext = fw[i] / sc.area
                q_abs = ab[i] * 120.0 * np.pi / sc.area
                act = 3.0 * aa[i] / vol_active if active is not None and vol_active > 0 else 0.0
                info = infos[i]
                r = {"wavelength_nm": float(lam), "psi": float(psi), "q_ext": q_ext, "q_abs": q_abs,
                     "active_density": act, "iterations": info["iterations"], "outer": info["outer"],
                     "true_relres": info["true_relres"], "converged": bool(info["converged"]),
                     "solve_s": (t2 - t1) / nb, "setup_s": (t1 - t0) / nb}
                if b0 + i == 0:
                    r["plan"] = plan
                results.append(r)
                if log:
                    log(r)
                if out_dir is not None and lam not in wl_index:
                    wl_index[lam] = len(wl_index)
                    write_point_files(out_dir, wl_index[lam], lam, q_ext, q_abs, act, sc.n, sc.d)
            t_setup += t1 - t0
            t_solve += t2 - t1
            t_post += time.perf_counter() - t2
    if out_dir is not None:
        merge_point_files(out_dir, len(wl_index))
    if timings is not None:
        timings.update({"setup_s": t_setup, "solve_s": t_solve, "post_s": t_post, "chunk": cap})
    return results


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=64, help="wavelengths from 400 to 800 nm")
    ap.add_argument("--n", type=int, default=96)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--fp64", action="store_true", help="plain FP64 solves instead of mixed precision")
    ap.add_argument("--out", default=None, help="directory for the reference-format result files")
    ap.add_argument("--sequential", action="store_true", help="one solve after the other instead of one batch")
    ap.add_argument("--warm", action="store_true",
                    help="with --sequential: start each solve from the previous wavelength's solution")
    args = ap.parse_args()
    points = G.dimer_sweep(args.count)
    t0 = time.perf_counter()
    kw = {"warm": args.warm} if args.sequential else {"timings": {}}
    run = run_sweep if args.sequential else run_sweep_batched
    res = run(lambda lam, psi: G.dimer_case(lam, psi=psi, n=args.n), points, tol=args.tol,
              mixed=not args.fp64, out_dir=args.out, log=lambda r: print(json.dumps(r), flush=True), **kw)
    wall = time.perf_counter() - t0
    its = np.array([r["iterations"] for r in res])
    print(json.dumps({"workload": f"{args.n}^3 Au dimer, {args.count} wavelengths x 2 polarisations",
                      "rhs": len(res), "wall_s": wall, "rhs_per_s": len(res) / wall,
                      "solve_s": sum(r["solve_s"] for r in res), "setup_s": sum(r["setup_s"] for r in res),
                      "iterations_min": int(its.min()), "iterations_median": float(np.median(its)),
                      "iterations_max": int(its.max()),
                      "all_converged": all(r["converged"] for r in res), "plan": res[0].get("plan"),
                      "batched": not args.sequential, "batch_timings": kw.get("timings")}), flush=True)


if __name__ == "__main__":
    main()
