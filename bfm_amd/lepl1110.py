"""Headless counterpart of the reference's command-line driver.

    python -m bfm_amd.lepl1110 MESH PROBLEM [--out DIR] [--solver NAME] [--rtol R]

Reads MESH (a .lepl1110 mesh, or a Wavefront .obj when the name ends in
.obj) and the PROBLEM file, runs the simulation through this package's
libbfm (the planar path solved on the GPU), and writes the nodal
displacements as DIR/U.txt (x components) and DIR/V.txt (y components) in
libbfm's format: a "Number of nodes N" header, then %14.7e fields, three per
line. DIR defaults to "data", as in the reference.

Reference: lepl1110.py:1-16 (Mesh_lepl1110, Ez_lepl1110, sim.run, two
ez.write calls) and libbfm/src/ez.c:211-229 (the file format). The display
step of the reference (bfm.show, pyglet) is not part of this driver.
"""

import argparse
import ctypes as C
import os
import sys

from . import _ffi as F
from . import pybfm as pb

SOLVERS = {"jacobi": 0, "block-jacobi": 1, "multigrid": 2, "learned": 3}


def run(mesh_path, problem_path, out_dir="data", solver=None, rtol=None, lib=None):
    """Runs one problem and writes out_dir/U.txt and out_dir/V.txt; returns
    (the Ez_lepl1110 object, the bfm_solve_stats_t of its instance)."""
    L = lib or pb.default_library()
    if mesh_path.endswith(".obj"):
        mesh = pb.Mesh_wavefront(mesh_path, lib=L)
    else:
        mesh = pb.Mesh_lepl1110(mesh_path, lib=L)
    ez = pb.Ez_lepl1110(mesh, problem_path, lib=L)
    if solver is not None or rtol is not None:
        cfg = F.bfm_solver_t()
        L.check(L.bfm_sim_get_solver(C.byref(ez.c_ez.sim), C.byref(cfg)))
        if solver is not None:
            cfg.precond = SOLVERS[solver]
        if rtol is not None:
            cfg.rtol = rtol
        if cfg.precond == SOLVERS["learned"]:
            from .learned import default_params
            p = default_params()
            L.check(L.bfm_sim_set_model(C.byref(ez.c_ez.sim), p.ctypes.data_as(C.POINTER(C.c_float)), p.size))
        L.check(L.bfm_sim_set_solver(C.byref(ez.c_ez.sim), C.byref(cfg)))
    ez.sim.run()
    stats = F.bfm_solve_stats_t()
    L.check(L.bfm_sim_get_stats(C.byref(ez.c_ez.sim), 0, C.byref(stats)))
    os.makedirs(out_dir, exist_ok=True)
    ez.write(os.path.join(out_dir, "U.txt"), 0)
    ez.write(os.path.join(out_dir, "V.txt"), 1)
    ez.keep_mesh = mesh
    return ez, stats


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m bfm_amd.lepl1110", description=__doc__.split("\n\n")[0])
    ap.add_argument("mesh")
    ap.add_argument("problem")
    ap.add_argument("--out", default="data", help="directory of U.txt and V.txt (default: data)")
    ap.add_argument("--solver", choices=sorted(SOLVERS), default=None,
                    help="preconditioner of the GPU solve (default: the library's, point Jacobi)")
    ap.add_argument("--rtol", type=float, default=None, help="relative residual target (default 1e-10)")
    a = ap.parse_args(argv)
    ez, st = run(a.mesh, a.problem, a.out, a.solver, a.rtol)
    n = ez.c_ez.instance.n_effects // 2
    print(f"{n} nodes: {st.iterations} iterations, relative residual {st.rel_residual:.3e}, "
          f"solve {st.solve_ms:.1f} ms; wrote {os.path.join(a.out, 'U.txt')} and {os.path.join(a.out, 'V.txt')}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
