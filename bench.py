"""Benchmark: DOF/s solved to a 1e-8 FP64 relative residual on one MI355X.

Workload (default): the 2000 x 500 Q1 plane-stress cantilever beam, nodes
jittered by +-20% of the cell size (seed 0), 2,005,002 DOFs, E = 211 GPa,
nu = 0.3, left edge clamped, traction -1e6 N/m on the top edge, 2x2 Gauss.

One step = the first solve, from u = 0, of a freshly assembled device-resident
system until ||b - K u||_2 / ||b||_2 <= 1e-8 (true residual recomputed in
FP64): every per-system cost of the solver is inside it (FP32 matrix copy,
multigrid topology and numeric set-up, solver buffers, captured chunks), as
in each bfm_sim_run. value = DOFs / mean step time of the headline solver,
flexible CG in FP64 preconditioned by one FP32 smoothed-aggregation V-cycle.
A re-solve on the same system (hierarchy cached) is reported beside it, and
plain FP64 Jacobi-PCG is timed on the system (the speedups are reported).

Also reported: the batched workload (4096 seeded holed panels of 1k-60k DOFs,
assembled and solved by block-Jacobi PCG in one graph launch, one workgroup per
panel: DOF/s and the launch's algorithmic GB/s, and the same batch in the
multigrid mode); the roofline of the headline solver's largest kernel (the
flexible-CG FP64 SpMV), of its finest-level post-smoothing step (FP32 SpMV +
block-Jacobi Chebyshev step) and of the FP64 CG SpMV (algorithmic bytes /
mean launch time from HIP events on the solver's stream), the end-to-end time (upload + renumbering + assembly + conditions +
solve), and a CPU baseline: the reference libbfm compiled from its own C
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
# HBM bytes per launch on the q1_2m beam measured by two rocprofv3 counter runs
# (FETCH_SIZE x 1.9998 + WRITE_SIZE, the factors read off a 256 MiB
# device-to-device copy in the same runs): profiles/hbm_traffic_q1_2m.md
TRAFFIC_Q1_2M = {"k_cg_spmv": 389.3e6, "k_fcg_spmv": 397.4e6, "k_cheb": 254.9e6}
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


def copy_calibration(nbytes):
    """One hipMemcpy of nbytes between two device buffers (the runtime's copy
    kernel reads and writes nbytes): the reference that scales the
    FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) == 0 and hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3) == 0          # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    log(f"copy calibration: {nbytes} bytes device to device")


def batched_leg(D, meshgen, count, rtol):
    """The batched workload: `count` seeded holed panels (1k-60k DOFs, plane
    strain, traction U(1e5, 1e7) N/m on top, bottom corners pinned), built
    into one batch from their meshes; one timed graph launch assembles every
    panel (one workgroup each) and solves it by block-Jacobi PCG (one
    workgroup each, the autotuner's configurations) to rtol."""
    t = time.perf_counter()
    inst, sizes, dofs = [], [], 0
    for k in range(count):
        m, tr = meshgen.panel(k)
        inst.append(D.panel_instance(m, tr))
        sizes.append(m.n_nodes)
        dofs += m.n_dof
    t_mesh = time.perf_counter() - t
    t = time.perf_counter()
    b = D.Batch.from_instances(inst)
    del inst
    t_create = time.perf_counter() - t
    log(f"batched: {count} panels, {dofs} DOFs; meshes {t_mesh:.1f} s, batch set-up {t_create:.1f} s")
    pre = D.PRECOND_BLOCK_JACOBI
    t = time.perf_counter()
    picks, times = b.autotune(precond=pre)
    t_tune = time.perf_counter() - t
    log(f"batched: autotune {t_tune:.1f} s, picks {picks.tolist()}")
    b.solve(precond=pre, rtol=rtol, picks=picks)         # graph capture + first launch
    t = time.perf_counter()
    st, ms = b.solve(precond=pre, rtol=rtol, picks=picks)
    wall = time.perf_counter() - t
    assert all(x["status"] == 0 and x["rel_residual"] <= rtol for x in st)
    its = np.array([x["iterations"] for x in st])
    nnzb = np.array([b.system(k).info()["nnzb"] for k in range(count)])
    byt = sum(D.batch_bytes(int(nnzb[k]), sizes[k], int(st[k]["iterations"]), int(st[k]["restarts"]), pre)
              for k in range(count))
    gbs = byt / (ms / 1e3) / 1e9
    log(f"batched: {ms:.1f} ms device ({wall * 1e3:.1f} ms wall), {dofs / (ms / 1e3):.3e} DOF/s, "
        f"iterations mean {its.mean():.0f} max {its.max()}, {gbs:.0f} GB/s algorithmic")
    # multigrid mode: the same batch solved as one block-diagonal system by
    # flexible CG + the multigrid V-cycle (each load scaled to unit norm, the
    # tolerance checks testing every panel's own residual); the first launch
    # builds the union and its hierarchy on the host
    t = time.perf_counter()
    b.solve(precond=D.PRECOND_AMG, rtol=rtol)
    t_mg_setup = time.perf_counter() - t
    t = time.perf_counter()
    st_mg, ms_mg = b.solve(precond=D.PRECOND_AMG, rtol=rtol)
    wall_mg = time.perf_counter() - t
    assert all(x["status"] == 0 and x["rel_residual"] <= rtol for x in st_mg)
    worst = max(x["rel_residual"] for x in st_mg)
    log(f"batched multigrid: {ms_mg:.1f} ms device ({wall_mg * 1e3:.1f} ms wall), {dofs / (ms_mg / 1e3):.3e} DOF/s, "
        f"{st_mg[0]['iterations']} iterations, worst instance residual {worst:.2e}, {ms / ms_mg:.2f}x the "
        f"persistent launch; first launch (host hierarchy build) {t_mg_setup:.1f} s")
    return {
        "workload": f"{count} holed panels (seeds 0-{count - 1}), 1k-60k DOFs, plane strain, block-Jacobi PCG "
                    f"to rtol {rtol:g}, assembled and solved in one graph launch",
        "n_dof": int(dofs), "ms_per_step": ms, "wall_ms": wall * 1e3, "value": dofs / (ms / 1e3),
        "unit": "DOF/s", "iterations_mean": float(its.mean()), "iterations_max": int(its.max()),
        "picks": picks.tolist(), "autotune_ms": times.tolist(), "setup_s": {"meshes": t_mesh,
                                                                              "batch_create": t_create,
                                                                              "autotune": t_tune},
        "roofline": {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": gbs / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_batch_pcg (block Jacobi, autotuned) + k_batch_assemble (whole graph launch)",
                     "algorithmic_bytes": byt},
        "multigrid": {
            "solver": "the batch as one block-diagonal system: flexible CG (FP64) + FP32 smoothed-aggregation "
                      "V-cycle, one dense coarsest inverse per panel, loads scaled to unit norm, every panel's "
                      f"own true residual checked against rtol {rtol:g}",
            "ms_per_step": ms_mg, "wall_ms": wall_mg * 1e3, "value": dofs / (ms_mg / 1e3), "unit": "DOF/s",
            "iterations": int(st_mg[0]["iterations"]), "worst_rel_residual": worst,
            "speedup_vs_persistent_launch": ms / ms_mg, "first_launch_s": t_mg_setup},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="q1_2m")
    ap.add_argument("--rtol", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-pcg", action="store_true", help="skip the plain Jacobi-PCG leg")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-preconditioner leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched holed-panel leg")
    ap.add_argument("--batch-count", type=int, default=4096, help="panels in the batched leg")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the batched leg and print its record (for profiling; not the bench line)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run one device-to-device copy of BYTES (the HBM counters' reference)")
    ap.add_argument("--mixed-cg", action="store_true",
                    help="also run CG on the FP32 matrix with FP64 reliable updates")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): the solver has no "
              "multi-GPU sharding; value is one GPU's throughput", file=sys.stderr)
        sys.exit(2)

    from bfm_amd import device as D
    from bfm_amd import meshgen

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
    if args.batch_only:
        print(json.dumps(batched_leg(D, meshgen, args.batch_count, args.rtol)), flush=True)
        return
    mesh, stress = meshgen.workload(args.workload)
    log(f"workload {args.workload}: {mesh.n_nodes} nodes, {mesh.n_dof} DOFs, {mesh.n_elems} elements")

    def fresh():
        t = time.perf_counter()
        sys_ = D.beam_system(mesh, stress=stress)
        return sys_, time.perf_counter() - t

    # headline: the first solve of a freshly assembled system, so every
    # per-system cost of the solver is inside the step (the FP32 matrix copy,
    # the multigrid hierarchy's topology and numeric set-up, the solver's
    # buffers and its captured chunks), as each bfm_sim_run pays it; the
    # system itself (upload, renumbering, pattern, assembly, conditions) is
    # built before the step's clock starts
    first_ms, first_st, setups = [], None, []
    s = None
    for k in range(args.warmup + args.steps):
        if s is not None:
            s.close()
        s, t_setup = fresh()
        setups.append(t_setup)
        t = time.perf_counter()
        _, st0 = s.solve(precond=D.PRECOND_AMG, rtol=args.rtol, want_u=False)
        dt = time.perf_counter() - t
        assert st0["status"] == 0 and st0["rel_residual"] <= args.rtol, st0
        tag = "warmup" if k < args.warmup else f"step {k - args.warmup}"
        log(f"fcg_amg first solve of a fresh system, {tag}: {dt * 1e3:.1f} ms, {st0['iterations']} iterations, "
            f"rel residual {st0['rel_residual']:.3e} (system set-up {t_setup:.2f} s)")
        if k >= args.warmup:
            first_ms.append(dt * 1e3)
            first_st = st0
    info = s.info()
    t_setup = float(np.mean(setups))
    log(f"setup {t_setup:.2f} s: {info}")

    def leg(name, precond, precision=D.FP64, steps=args.steps):
        for _ in range(args.warmup):
            t = time.perf_counter()
            _, st = s.solve(precond=precond, precision=precision, rtol=args.rtol, want_u=False)
            log(f"{name} warmup {(time.perf_counter() - t) * 1e3:.1f} ms, {st['iterations']} iterations")
        times, stats = [], []
        for k in range(steps):
            t = time.perf_counter()
            _, st = s.solve(precond=precond, precision=precision, rtol=args.rtol, want_u=False)
            dt = time.perf_counter() - t
            assert st["status"] == 0 and st["rel_residual"] <= args.rtol, st
            times.append(dt)
            stats.append(st)
            log(f"{name} step {k}: {dt * 1e3:.1f} ms, {st['iterations']} iterations, "
                f"rel residual {st['rel_residual']:.3e}")
        ms = 1e3 * float(np.mean(times))
        return ms, stats[-1]

    ms_first = float(np.mean(first_ms))
    value = mesh.n_dof / (ms_first / 1e3)
    # the same solver re-solving on a system whose hierarchy exists
    ms, st = leg("fcg_amg (cached hierarchy)", D.PRECOND_AMG)

    def roof(kernel, label):
        k_ms, k_bytes = s.time_kernel(kernel, reps=100)
        gbs = k_bytes / (k_ms / 1e3) / 1e9
        log(f"{label}: {k_ms * 1e3:.1f} us, {k_bytes / 1e6:.1f} MB, {gbs:.0f} GB/s")
        traffic = TRAFFIC_Q1_2M.get(label.split("<")[0].split(" ")[0]) if args.workload == "q1_2m" else None
        return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": gbs / HBM_PEAK_GBS, "traffic": traffic, "kernel": label,
                "kernel_us": k_ms * 1e3, "algorithmic_bytes": k_bytes}

    roof_fcg = roof(D.KERNEL_FCG_SPMV, "k_fcg_spmv (flexible-CG FP64 SpMV, the solve's largest kernel)")
    roof_smooth = roof(D.KERNEL_AMG_SMOOTH, "k_cheb<2,double,false,true> (finest-level post-smoothing step)")
    roof_spmv = roof(D.KERNEL_CG_SPMV64, "k_cg_spmv<double> (FP64 CG SpMV)")

    extra = {}
    if not args.no_pcg:
        p_ms, p_st = leg("pcg_jacobi", D.PRECOND_JACOBI, steps=max(1, min(args.steps, 2)))
        extra["pcg_jacobi_fp64"] = {
            "ms_per_step": p_ms, "value": mesh.n_dof / (p_ms / 1e3), "iterations": int(p_st["iterations"]),
            "rel_residual": p_st["rel_residual"],
        }
        extra["speedup_vs_pcg"] = p_ms / ms_first
        extra["speedup_vs_pcg_cached_hierarchy"] = p_ms / ms
        log(f"speedup of multigrid FCG over Jacobi-PCG: {p_ms / ms_first:.2f}x first solve, "
            f"{p_ms / ms:.2f}x re-solve")
    weights = os.path.join(ROOT, "bfm_amd", "weights", "mpnn_elasticity.f32")
    if not args.no_learned and os.path.exists(weights):
        # the learned preconditioner beside the method without it (block-Jacobi PCG)
        s.set_model(D.Model.load(weights))
        w_save, args.warmup = args.warmup, 0
        b_ms, b_st = leg("pcg_block_jacobi", D.PRECOND_BLOCK_JACOBI, steps=1)
        l_ms, l_st = leg("learned_fcg", D.PRECOND_LEARNED, steps=1)
        args.warmup = w_save
        extra["learned_fcg"] = {
            "ms_per_step": l_ms, "value": mesh.n_dof / (l_ms / 1e3), "iterations": int(l_st["iterations"]),
            "rel_residual": l_st["rel_residual"], "fallback": int(l_st["fallback"]),
            "block_jacobi_pcg": {"ms_per_step": b_ms, "iterations": int(b_st["iterations"])},
            "iteration_ratio_vs_block_jacobi": b_st["iterations"] / max(1, l_st["iterations"]),
        }
        if "pcg_jacobi_fp64" in extra:
            extra["learned_fcg"]["speedup_vs_pcg"] = extra["pcg_jacobi_fp64"]["ms_per_step"] / l_ms
        log(f"learned FCG: {l_ms:.1f} ms, {l_st['iterations']} iterations (block-Jacobi PCG "
            f"{b_st['iterations']}, {b_ms:.1f} ms)")
    if args.mixed_cg:
        try:
            m_ms, m_st = leg("cg_fp32_matrix", D.PRECOND_JACOBI, D.MIXED, steps=1)
            extra["cg_fp32_matrix"] = {"ms_per_step": m_ms, "iterations": int(m_st["iterations"])}
        except (D.BfmxError, AssertionError) as e:
            extra["cg_fp32_matrix"] = {"status": str(e)}
            log(f"cg_fp32_matrix: {e}")

    if not args.no_batch:
        extra["batched_panels"] = batched_leg(D, meshgen, args.batch_count, args.rtol)

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
        "ms_per_step": ms_first,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32",
        "data": "synthetic",
        "config": {
            "workload": f"{args.workload}: 2000x500 Q1 plane-stress cantilever beam, 20% jitter (seed 0), "
                        f"{mesh.n_dof} DOFs, solved to rtol {args.rtol:g} from u = 0",
            "solver": "flexible CG (FP64) + FP32 smoothed-aggregation multigrid V-cycle",
            "n_dof": mesh.n_dof,
            "n_elems": mesh.n_elems,
            "iterations": int(first_st["iterations"]),
            "rel_residual": first_st["rel_residual"],
            "step": "first solve of a freshly assembled system: FP32 matrix copy, multigrid topology and "
                    "numeric set-up, solver buffers and captured chunks, and the iterations to the tolerance",
            "setup_s": t_setup,
            "end_to_end_s": t_setup + ms_first / 1e3,
            "resolve_cached_hierarchy": {"ms_per_step": ms, "value": mesh.n_dof / (ms / 1e3),
                                         "iterations": int(st["iterations"]), "rel_residual": st["rel_residual"]},
        },
        "roofline": roof_fcg,
        "roofline_smoother": roof_smooth,
        "roofline_spmv_fp64": roof_spmv,
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
