"""Benchmark: DOF/s solved to a 1e-8 FP64 relative residual on one MI355X.

Workload (default): the 2000 x 500 Q1 plane-stress cantilever beam, nodes
jittered by +-20% of the cell size (seed 0), 2,005,002 DOFs, E = 211 GPa,
nu = 0.3, left edge clamped, traction -1e6 N/m on the top edge, 2x2 Gauss.
One step = one Jacobi-PCG solve from u = 0 on the device-resident assembled
system until ||b - K u||_2 / ||b||_2 <= 1e-8 (true residual recomputed in
FP64). value = DOFs / mean step time.

Also reported: the SpMV kernel's HBM roofline (algorithmic bytes / average
launch time from HIP events on the solver's stream), the end-to-end time
(upload + renumbering + assembly + conditions + solve), the mixed-precision
solve, and a CPU baseline: the reference libbfm compiled from its own C
sources (oracle/_ref), single-threaded as libbfm is, on a bounded sample (its
dense path cannot hold the 2M-DOF matrix: 32 TB).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload q1_2m]
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CPU_SAMPLE = (180, 45)   # Q1 beam cells for the reference CPU path (16,652 DOFs)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline_reference():
    """libbfm (reference C) on a bounded beam sample, through its own C API."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from bfm_amd import meshgen
    from bfm_amd import pybfm as pb
    ref = os.path.join(ROOT, "oracle", "_ref", "libbfm_ref.so")
    nx, ny = CPU_SAMPLE
    mesh = meshgen.beam(nx, ny, kind=4, jitter=0.2, seed=0)
    mat = meshgen.BEAM_MATERIAL
    if os.path.exists(ref):
        L = pb.Library(ref)
        cm = pb.Mesh_arrays.from_generated(mesh, lib=L)
        m = pb.Material("steel", mat["rho"], mat["E"], mat["nu"], lib=L)
        rule = pb.Rule_gauss_legendre(2, 4, lib=L)
        obj = pb.Obj(cm, m, rule, lib=L)
        inst = pb.Instance(obj, lib=L)
        conds = []
        kinds, values, flags = meshgen.beam_conditions(mesh)
        for k, v, f in zip(kinds, values, flags):
            c = pb.Condition(cm, int(k), float(v), lib=L)
            c.set_nodes(f)
            inst.add_condition(c)
            conds.append(c)
        sim = pb.Sim(pb.Sim.PLANAR_STRESS, lib=L)
        sim.add_instance(inst)
        t0 = time.perf_counter()
        sim.run()
        dt = time.perf_counter() - t0
        kind = "reference"
        what = "libbfm compiled from the reference C sources"
    else:
        import oracle as O
        kinds, values, flags = meshgen.beam_conditions(mesh)
        t0 = time.perf_counter()
        O.solve_dense(mesh, True, mat["E"], mat["nu"], mat["rho"], kinds, values, flags)
        dt = time.perf_counter() - t0
        kind = "port"
        what = "the oracle's C restatement of libbfm's path"
    return {
        "value": mesh.n_dof / dt,
        "unit": "DOF/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"{what}, single-threaded: dense assembly, ordered conditions, RCM, band LU on the "
                   f"{nx}x{ny} jittered Q1 beam ({mesh.n_dof} DOFs) in {dt:.2f} s; the 2,005,002-DOF "
                   f"workload is infeasible for libbfm's dense path (32 TB matrix, int index overflow)"),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="q1_2m")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); multi-GPU sharding is "
              "not built (see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from bfm_amd import device as D
    from bfm_amd import meshgen

    mesh, stress = meshgen.workload(args.workload)
    log(f"workload {args.workload}: {mesh.n_nodes} nodes, {mesh.n_dof} DOFs, {mesh.n_elems} elements")

    t0 = time.perf_counter()
    s = D.beam_system(mesh, stress=stress)
    info = s.info()
    t_setup = time.perf_counter() - t0
    log(f"setup {t_setup:.2f} s: {info}")

    def step(precision):
        t = time.perf_counter()
        _, st = s.solve(precision=precision, rtol=args.rtol, want_u=False)
        return time.perf_counter() - t, st

    for _ in range(args.warmup):
        dt, st = step(D.FP64)
        log(f"warmup {dt * 1e3:.1f} ms, {st['iterations']} iterations")
    times, stats = [], []
    for k in range(args.steps):
        dt, st = step(D.FP64)
        assert st["status"] == 0 and st["rel_residual"] <= args.rtol, st
        times.append(dt)
        stats.append(st)
        log(f"step {k}: {dt * 1e3:.1f} ms, {st['iterations']} iterations, rel residual {st['rel_residual']:.3e}")
    ms = 1e3 * float(np.mean(times))
    value = mesh.n_dof / (ms / 1e3)

    spmv_ms, spmv_bytes = s.time_spmv(D.FP64, reps=100)
    achieved = spmv_bytes / (spmv_ms / 1e3) / 1e9
    log(f"spmv fp64: {spmv_ms * 1e3:.1f} us, {spmv_bytes / 1e6:.1f} MB, {achieved:.0f} GB/s")

    extra = {}
    if not args.no_mixed:
        try:
            dt, st = step(D.MIXED)
            m_ms, m_bytes = s.time_spmv(D.MIXED, reps=100)
            extra["mixed"] = {
                "ms_per_step": 1e3 * dt, "value": mesh.n_dof / dt, "iterations": st["iterations"],
                "rel_residual": st["rel_residual"], "speedup_vs_fp64_pcg": (ms / 1e3) / dt,
                "spmv_us": m_ms * 1e3, "spmv_gbs_algorithmic": m_bytes / (m_ms / 1e3) / 1e9,
            }
            log(f"mixed: {1e3 * dt:.1f} ms, {st['iterations']} iterations, rel residual {st['rel_residual']:.3e}")
        except D.BfmxError as e:
            extra["mixed"] = {"status": str(e)}
            log(f"mixed: {e}")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline_reference()
        log(f"cpu baseline {cpu['value']:.0f} DOF/s")

    out = {
        "metric": "DOF/s solved to 1e-8 FP64 rel. residual; learned-FCG vs plain PCG speedup",
        "value": value,
        "unit": "DOF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: 2000x500 Q1 plane-stress cantilever beam, 20% jitter (seed 0), "
                        f"{mesh.n_dof} DOFs, Jacobi-PCG FP64 to rtol {args.rtol:g}",
            "n_dof": mesh.n_dof,
            "n_elems": mesh.n_elems,
            "preconditioner": "jacobi",
            "iterations": int(stats[-1]["iterations"]),
            "rel_residual": stats[-1]["rel_residual"],
            "setup_s": t_setup,
            "end_to_end_s": t_setup + ms / 1e3,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_cg_spmv64",
            "kernel_us": spmv_ms * 1e3,
            "algorithmic_bytes": spmv_bytes,
        },
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
