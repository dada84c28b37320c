"""The drop-in libbfm.so on the GPU, driven as the reference's own clients
drive libbfm: pybfm-style wrapper classes, a C program written against the
public headers, and the system-level calls of run_elasticity
(libbfm/src/sim.c:103-135). Displacements are compared with the reference
(fixtures, the data/U.txt and data/V.txt files, and the reference library
compiled from its sources when oracle/_ref holds it) within 1e-8 relative.
"""

import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import REF_DATA, ROOT, dense_from_golden, rel_err

from bfm_amd import _ffi as F
from bfm_amd import meshgen
from bfm_amd import pybfm as pb

pytestmark = pytest.mark.gpu

LEPL = os.path.join(REF_DATA, "8.lepl1110")
PROBLEM = os.path.join(REF_DATA, "problem.txt")
OBJ = os.path.join(REF_DATA, "bridge-nothing.obj")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libbfm_ref.so")
U_TOL = 1e-8


@pytest.fixture(scope="module")
def L():
    return pb.Library()


def test_ez_lepl1110_run_and_write(L, golden, tmp_path):
    mesh = pb.Mesh_lepl1110(LEPL, lib=L)
    ez = pb.Ez_lepl1110(mesh, PROBLEM, lib=L)
    ez.sim.run()
    u = ez.effects
    err = rel_err(u, golden["lepl8/u"])
    print(f"8.lepl1110: error vs libbfm {err:.2e}")
    assert err <= U_TOL
    ez.write(str(tmp_path / "U.txt"), 0)
    ez.write(str(tmp_path / "V.txt"), 1)
    for name, comp in (("U.txt", 0), ("V.txt", 1)):
        mine = pb.read_effects_file(str(tmp_path / name))
        ref = pb.read_effects_file(os.path.join(REF_DATA, name))
        assert np.abs(mine - ref).max() <= 5e-8 * np.abs(ref).max()


def test_c_client_program(golden, tmp_path):
    exe = tmp_path / "lepl1110_run"
    lib_dir = os.path.join(ROOT, "bfm_amd", "lib")
    subprocess.run(["gcc", "-std=gnu99", "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "c",
                    "lepl1110_run.c"), "-o", str(exe), "-L", lib_dir, "-lbfm", f"-Wl,-rpath,{lib_dir}"], check=True)
    r = subprocess.run([str(exe), LEPL, PROBLEM, str(tmp_path / "U.txt"), str(tmp_path / "V.txt")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    U = pb.read_effects_file(str(tmp_path / "U.txt"))
    assert np.abs(U - golden["lepl8/u"][0::2]).max() <= 5e-8 * np.abs(U).max()


def c1_sim(L, grav=False, funky=None, solver=None):
    mesh = meshgen.beam(40, 10, kind=3)
    mat = meshgen.BEAM_MATERIAL
    objs = {}
    objs["mesh"] = cm = pb.Mesh_arrays.from_generated(mesh, lib=L)
    objs["mat"] = m = pb.Material("steel", mat["rho"], mat["E"], mat["nu"], lib=L)
    objs["rule"] = rule = pb.Rule_gauss_legendre(2, 3, lib=L)
    objs["obj"] = obj = pb.Obj(cm, m, rule, lib=L)
    objs["inst"] = inst = pb.Instance(obj, lib=L)
    conds = []
    for kind, value, where in [(3, meshgen.BEAM_TRACTION, "top"), (0, 0.0, "left"), (1, 0.0, "left")]:
        c = pb.Condition(cm, kind, value, lib=L)
        c.set_nodes(mesh.node_sets[where])
        inst.add_condition(c)
        conds.append(c)
    objs["conds"] = conds
    sim = pb.Sim(pb.Sim.PLANAR_STRESS, lib=L)
    sim.add_instance(inst)
    if grav:
        f = pb.Force_linear(pb.Force_linear.EARTH_GRAVITY_2D, lib=L)
        sim.add_force(f)
        objs["f"] = f
    if funky is not None:
        f = pb.Force(2, lib=L)
        assert L.bfm_force_set_funky(C.byref(f.c_force), C.cast(funky, C.c_void_p), None) == 0
        sim.add_force(f)
        objs["funky"] = f
    if solver is not None:
        L.lib.bfm_sim_set_solver.argtypes = [C.c_void_p, C.c_void_p]
        assert L.lib.bfm_sim_set_solver(C.byref(sim.c_sim), C.byref(solver)) == 0
    return sim, inst, objs


class Solver(C.Structure):
    _fields_ = [("precond", C.c_int), ("precision", C.c_int), ("rtol", C.c_double), ("max_iter", C.c_size_t)]


class SolveStats(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", C.c_size_t), ("rel_residual", C.c_double),
                ("solve_ms", C.c_double), ("total_ms", C.c_double)]


@pytest.mark.parametrize("grav", [False, True])
def test_c1_through_wrapper_classes(L, golden, grav):
    sim, inst, _ = c1_sim(L, grav)
    sim.run()
    name = "c1_gravity" if grav else "c1"
    err = rel_err(inst.effects, golden[f"{name}/u"])
    assert err <= U_TOL
    st = SolveStats()
    L.lib.bfm_sim_get_stats.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]
    assert L.lib.bfm_sim_get_stats(C.byref(sim.c_sim), 0, C.byref(st)) == 0
    assert st.iterations > 0 and st.rel_residual <= 1e-10


def test_solver_settings_and_funky_force(L, golden):
    calls = []

    @F.FUNKY
    def gravity(force, pos, out, data):
        calls.append(1)
        out.contents.data[0] = 0.0
        out.contents.data[1] = -9.81
        return 0

    sim, inst, _ = c1_sim(L, funky=gravity, solver=Solver(1, 0, 1e-10, 0))
    sim.run()
    assert calls, "the callback was not evaluated"
    assert rel_err(inst.effects, golden["c1_gravity/u"]) <= U_TOL


def test_system_level_calls(L, golden):
    """create_planar_stress, renumber, matrix_solve, perm_vec(inv): the
    sequence of run_elasticity, on the device-resident matrix."""
    sim, inst, objs = c1_sim(L)
    sysm = F.bfm_system_t()
    no_forces = (C.POINTER(F.bfm_force_t) * 1)()
    assert L.bfm_system_create_planar_stress(C.byref(sysm), C.byref(inst.c_instance), 0, no_forces) == 0
    assert sysm.A.kind == F.MATRIX_KIND_BSR and sysm.A.m == 902
    K = dense_from_golden(golden, "c1", 902)
    for (i, j) in [(0, 0), (5, 7), (100, 101), (901, 901), (400, 2)]:
        assert abs(L.bfm_matrix_get(C.byref(sysm.A), i, j) - K[i, j]) <= 1e-13 * np.abs(K).max()
    b = np.ctypeslib.as_array(sysm.b.data, shape=(902,)).copy()
    assert np.abs(b - golden["c1/b"]).max() <= 1e-13 * np.abs(b).max()
    assert L.bfm_system_renumber(C.byref(sysm)) == 0
    i, j = 10, 11
    pi, pj = sysm.perm.perm[i], sysm.perm.perm[j]
    assert abs(L.bfm_matrix_get(C.byref(sysm.A), pi, pj) - K[i, j]) <= 1e-13 * np.abs(K).max()
    assert L.bfm_matrix_solve(C.byref(sysm.A), C.byref(sysm.b)) == 0
    assert L.bfm_perm_perm_vec(C.byref(sysm.perm), C.byref(sysm.b), True) == 0
    u = np.ctypeslib.as_array(sysm.b.data, shape=(902,)).copy()
    assert rel_err(u, golden["c1/u"]) <= U_TOL
    assert L.bfm_system_destroy(C.byref(sysm)) == 0


def test_several_instances_in_one_sim(L, golden):
    sim, inst, objs = c1_sim(L)
    inst2 = pb.Instance(objs["obj"], lib=L)
    for c in objs["conds"]:
        inst2.add_condition(c)
    sim.add_instance(inst2)
    sim.run()
    assert rel_err(inst.effects, golden["c1/u"]) <= U_TOL
    assert np.array_equal(inst.effects, inst2.effects)


def obj_conditions(mesh, lib):
    xy = mesh.coords.reshape(-1, 2)
    x0, x1 = xy[:, 0].min(), xy[:, 0].max()
    y0, y1 = xy[:, 1].min(), xy[:, 1].max()
    ends = (xy[:, 0] <= x0 + 0.02 * (x1 - x0)) | (xy[:, 0] >= x1 - 0.02 * (x1 - x0))   # both supports
    top = xy[:, 1] >= y1 - 0.05 * (y1 - y0)
    out = []
    for kind, value, flags in [(0, 0.0, ends), (1, 0.0, ends), (3, -1e5, top), (5, 2e4, top)]:
        c = pb.Condition(mesh, kind, value, lib=lib)
        c.set_nodes(flags)
        out.append(c)
    return out


def run_obj(lib):
    mesh = pb.Mesh_wavefront(OBJ, lib=lib)
    mat = pb.Material("AA7075", 2.81e3, 71.7e9, 0.33, lib=lib)
    rule = pb.Rule_gauss_legendre(2, 3, lib=lib)
    obj = pb.Obj(mesh, mat, rule, lib=lib)
    inst = pb.Instance(obj, lib=lib)
    conds = obj_conditions(mesh, lib)
    for c in conds:
        inst.add_condition(c)
    sim = pb.Sim(pb.Sim.PLANAR_STRAIN, lib=lib)
    sim.add_instance(inst)
    g = pb.Force_linear(pb.Force_linear.EARTH_GRAVITY_2D, lib=lib)
    sim.add_force(g)
    sim.run()
    return inst.effects


def test_wavefront_bridge_against_reference_library(L):
    if not os.path.exists(REF_LIB):
        pytest.skip("reference library not built")
    u = run_obj(L)
    u_ref = run_obj(pb.Library(REF_LIB))
    err = rel_err(u, u_ref)
    print(f"bridge-nothing.obj: error vs libbfm {err:.2e}")
    assert err <= U_TOL


def test_lepl1110_command_line(tmp_path):
    """python -m bfm_amd.lepl1110 MESH PROBLEM writes U.txt and V.txt as the
    reference's lepl1110.py:9-14 does (libbfm/src/ez.c:211-229 format); they
    match the reference's own data/U.txt and data/V.txt to their 8 digits."""
    import sys
    out = tmp_path / "data"
    r = subprocess.run([sys.executable, "-m", "bfm_amd.lepl1110", LEPL, PROBLEM, "--out", str(out)],
                       cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    print(r.stdout.strip())
    for name in ("U.txt", "V.txt"):
        mine = pb.read_effects_file(str(out / name))
        ref = pb.read_effects_file(os.path.join(REF_DATA, name))
        assert np.abs(mine - ref).max() <= 5e-8 * np.abs(ref).max(), name
        head_mine = open(out / name).readline()
        assert head_mine == open(os.path.join(REF_DATA, name)).readline()


def beam_sim(L, meshes, solver=None, params=None):
    """One sim holding one instance per generated beam mesh (plane stress,
    traction on top, left edge clamped)."""
    mat = meshgen.BEAM_MATERIAL
    keep = []
    sim = pb.Sim(pb.Sim.PLANAR_STRESS, lib=L)
    insts = []
    for mesh in meshes:
        cm = pb.Mesh_arrays.from_generated(mesh, lib=L)
        m = pb.Material("steel", mat["rho"], mat["E"], mat["nu"], lib=L)
        rule = pb.Rule_gauss_legendre(2, mesh.kind, lib=L)
        obj = pb.Obj(cm, m, rule, lib=L)
        inst = pb.Instance(obj, lib=L)
        kinds, values, flags = meshgen.beam_conditions(mesh)
        for k, v, f in zip(kinds, values, flags):
            c = pb.Condition(cm, int(k), float(v), lib=L)
            c.set_nodes(f)
            inst.add_condition(c)
            keep.append(c)
        keep += [cm, m, rule, obj]
        sim.add_instance(inst)
        insts.append(inst)
    if params is not None:
        assert L.bfm_sim_set_model(C.byref(sim.c_sim), params.ctypes.data_as(C.POINTER(C.c_float)),
                                   params.size) == 0
    if solver is not None:
        assert L.bfm_sim_set_solver(C.byref(sim.c_sim), C.byref(solver)) == 0
    sim.keep = keep
    return sim, insts


@pytest.mark.parametrize("precond", [0, 1, 2])
def test_sim_solves_several_instances_in_one_batch(L, precond):
    """Several instances of uneven sizes (P1 and Q1) in one sim run as one
    batch (the instance loop of libbfm/src/sim.c:103-135 in one launch); each
    matches the same instance run alone (a single-instance sim, one device
    system) within 1e-8."""
    meshes = [meshgen.beam(40, 10, kind=3), meshgen.beam(30, 8, kind=4, jitter=0.2, seed=1),
              meshgen.beam(60, 15, kind=3, jitter=0.2, seed=2), meshgen.beam(12, 3, kind=4)]
    solver = F.bfm_solver_t(precond, 0, 1e-10, 0)
    sim, insts = beam_sim(L, meshes, solver)
    sim.run()
    for k, mesh in enumerate(meshes):
        solo, (one,) = beam_sim(L, [mesh], F.bfm_solver_t(1, 0, 1e-10, 0))
        solo.run()
        st = F.bfm_solve_stats_t()
        assert L.bfm_sim_get_stats(C.byref(sim.c_sim), k, C.byref(st)) == 0
        err = rel_err(insts[k].effects, one.effects)
        print(f"precond {precond}, instance {k} ({mesh.n_dof} DOFs): {st.iterations} iterations, "
              f"residual {st.rel_residual:.2e}, error vs solo {err:.2e}")
        assert st.status == 0 and st.rel_residual <= 1e-10 and err <= U_TOL


def test_sim_batch_below_the_rounding_floor(L):
    """A tolerance below what FP64 can reach on some instances (the 902-DOF
    beam's floor is ~2e-11): the batch leaves them above rtol, each is solved
    again on its own, and the run is accepted at the floor (status 1), as a
    single-instance run is."""
    meshes = [meshgen.beam(40, 10, kind=3), meshgen.beam(12, 3, kind=4)]
    sim, insts = beam_sim(L, meshes, F.bfm_solver_t(1, 0, 1e-14, 20000))
    sim.run()
    for k in range(len(meshes)):
        st = F.bfm_solve_stats_t()
        assert L.bfm_sim_get_stats(C.byref(sim.c_sim), k, C.byref(st)) == 0
        print(f"instance {k}: status {st.status}, residual {st.rel_residual:.2e}")
        assert st.status in (0, 1) and st.rel_residual <= 1e-9


def test_sim_learned_solver(L, golden):
    """precond 3 (learned flexible CG on the Jacobi base) through the libbfm
    API with the packaged parameters trained for that base; refused without
    parameters."""
    from bfm_amd.learned import WEIGHTS, default_params
    mesh = meshgen.beam(40, 10, kind=3)
    sim, (inst,) = beam_sim(L, [mesh], F.bfm_solver_t(3, 0, 1e-10, 0))
    with pytest.raises(AssertionError, match="bfm_sim_set_model"):
        sim.run()
    sim2, (inst2,) = beam_sim(L, [mesh], F.bfm_solver_t(3, 0, 1e-10, 0), params=default_params(WEIGHTS))
    sim2.run()
    st = F.bfm_solve_stats_t()
    assert L.bfm_sim_get_stats(C.byref(sim2.c_sim), 0, C.byref(st)) == 0
    err = rel_err(inst2.effects, golden["c1/u"])
    print(f"learned through bfm_sim_run: {st.iterations} iterations, error vs libbfm {err:.2e}")
    assert err <= U_TOL
