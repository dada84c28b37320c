"""GPU parity: the HIP path (through the bfmx C ABI) against the reference
fixtures (tests/golden, written by libbfm compiled from its own sources) and
the CPU oracle.

Tolerances (floating point; the GPU sums Gauss points in registers, fuses
multiply-adds and mirrors the element matrix, so it rounds differently):
  K, b entries     |gpu - ref| <= 1e-13 * max|ref|
  displacements    ||u_gpu - u_ref|| / ||u_ref|| <= 1e-8
  residuals        ||b - K u|| / ||b|| <= the solver tolerance, K and b from
                   the oracle's own assembly
"""

import numpy as np
import pytest

from conftest import REF_DATA, dense_from_golden, rel_err
from make_golden import TINY_CONDITIONS, node_sets

from bfm_amd import device as D
from bfm_amd import meshgen
from bfm_amd.pybfm import read_effects_file

pytestmark = pytest.mark.gpu

K_TOL = 1e-13
U_TOL = 1e-8


def close_entries(a, ref, tol=K_TOL):
    scale = np.abs(ref).max()
    err = np.abs(a - ref).max()
    assert err <= tol * scale, f"max entry error {err:.3e} > {tol:g} * {scale:.3e}"


def tiny_mesh(kind):
    return meshgen.beam(3, 2, kind=kind, length=3.0, height=2.0, jitter=0.2, seed=11)


def tiny_conditions(mesh):
    sets = node_sets(mesh)
    kinds = [c[0] for c in TINY_CONDITIONS]
    values = [c[1] for c in TINY_CONDITIONS]
    flags = np.stack([sets[c[2]] for c in TINY_CONDITIONS]).astype(np.uint8)
    return kinds, values, flags


def test_device_visible():
    assert D.device_count() >= 1


@pytest.mark.parametrize("kind,tag", [(4, "q1"), (3, "p1")])
@pytest.mark.parametrize("stress", [0, 1])
@pytest.mark.parametrize("renumber", [True, False])
def test_tiny_assembly_and_conditions(golden, kind, tag, stress, renumber):
    name = f"tiny_{tag}_{'stress' if stress else 'strain'}"
    mesh = tiny_mesh(kind)
    s = D.System.from_mesh(mesh, renumber=renumber)
    s.assemble(stress, 7.0e10, 0.33, 2700.0, g=(0.0, -9.81))
    close_entries(s.get_dense(), golden[f"{name}/K_pre"])
    close_entries(s.get_rhs(), golden[f"{name}/b_pre"])
    # every condition kind, in the golden order
    s.assemble(stress, 7.0e10, 0.33, 2700.0, g=(0.0, -9.81))
    s.apply_conditions(*tiny_conditions(mesh))
    close_entries(s.get_dense(), golden[f"{name}/K"])
    close_entries(s.get_rhs(), golden[f"{name}/b"])
    u, st = s.solve(rtol=1e-14)
    assert st["status"] == 0
    assert rel_err(u, golden[f"{name}/u"]) <= U_TOL


@pytest.mark.parametrize("tag", ["q1", "p1"])
@pytest.mark.parametrize("t", [0, 1, 2])
def test_single_element_matrix(golden, tag, t):
    pre = f"elem_{tag}_{t}"
    coords = golden[f"{pre}/coords"]
    elems = golden[f"{pre}/elems"]
    kind = elems.shape[1]
    s = D.System(kind, coords, elems)
    s.assemble(int(golden[f"{pre}/stress"]), 2.0e11, float(golden[f"{pre}/nu"]), 7800.0, g=(0.0, -9.81))
    close_entries(s.get_dense(), golden[f"{pre}/K"])
    close_entries(s.get_rhs(), golden[f"{pre}/b"])


C1_CONDS = [(3, meshgen.BEAM_TRACTION, "top"), (0, 0.0, "left"), (1, 0.0, "left")]


@pytest.mark.parametrize("name,conds,grav", [
    ("c1", C1_CONDS, False),
    ("c1_gravity", C1_CONDS, True),
    ("c1_neumann_last", C1_CONDS[1:] + C1_CONDS[:1], False),
])
@pytest.mark.parametrize("precond", [D.PRECOND_JACOBI, D.PRECOND_BLOCK_JACOBI, D.PRECOND_AMG])
def test_c1_beam(golden, name, conds, grav, precond):
    mesh = meshgen.beam(40, 10, kind=3)
    mat = meshgen.BEAM_MATERIAL
    s = D.System.from_mesh(mesh)
    s.assemble(1, mat["E"], mat["nu"], mat["rho"], g=(0.0, -9.81) if grav else (0.0, 0.0))
    for kind, value, where in conds:
        s.apply_condition(kind, value, mesh.node_sets[where])
    K = dense_from_golden(golden, name, mesh.n_dof)
    close_entries(s.get_dense(), K)
    close_entries(s.get_rhs(), golden[f"{name}/b"])
    u, st = s.solve(precond=precond, rtol=1e-10)
    assert st["status"] == 0 and st["rel_residual"] <= 1e-10
    err = rel_err(u, golden[f"{name}/u"])
    print(f"{name}: {st['iterations']} iterations, residual {st['rel_residual']:.2e}, error vs libbfm {err:.2e}")
    assert err <= U_TOL


@pytest.mark.parametrize("precond", [D.PRECOND_JACOBI, D.PRECOND_AMG])
def test_lepl1110_mesh_problem(golden, precond):
    """8.lepl1110 + problem.txt as parsed by the reference, solved on the GPU,
    against the reference's solve and the 8-digit data/U.txt, data/V.txt."""
    coords = golden["lepl8/coords"]
    elems = golden["lepl8/elems"]
    E, nu, rho = golden["lepl8/E_nu_rho"]
    s = D.System(elems.shape[1], coords, elems, edges=golden["lepl8/edge_nodes"],
                 edge_boundary=(golden["lepl8/edge_elems"][:, 1] == -1).astype(np.uint8))
    stress = int(golden["lepl8/sim_kind"]) == 2
    s.assemble(stress, E, nu, rho, g=(0.0, -9.81))
    flags = golden["lepl8/cond_flags"]
    for i in range(int(golden["lepl8/n_conditions"])):
        kind, value = golden[f"lepl8/cond{i}_kind_value"]
        s.apply_condition(int(kind), value, flags[i])
    u, st = s.solve(precond=precond, rtol=1e-10)
    assert st["status"] == 0
    err = rel_err(u, golden["lepl8/u"])
    print(f"lepl8: {st['iterations']} iterations, residual {st['rel_residual']:.2e}, error vs libbfm {err:.2e}")
    assert err <= U_TOL
    U = read_effects_file(f"{REF_DATA}/U.txt")
    V = read_effects_file(f"{REF_DATA}/V.txt")
    # the files hold 8 significant digits
    assert np.abs(u[0::2] - U).max() <= 5e-8 * np.abs(U).max() + 1e-20
    assert np.abs(u[1::2] - V).max() <= 5e-8 * np.abs(V).max() + 1e-20


def oracle_csr(mesh, stress, forces=None):
    import oracle as O
    mat = meshgen.BEAM_MATERIAL
    kinds, values, flags = meshgen.beam_conditions(mesh)
    return O.assemble_csr(mesh, stress, mat["E"], mat["nu"], mat["rho"], forces=forces, kinds=kinds,
                          values=values, flags=flags)


@pytest.mark.parametrize("kind", [3, 4])
def test_mid_mesh_matrix_and_solution(kind):
    """A jittered 120 x 30 beam: every matrix entry against the oracle's CSR
    assembly, and u against a direct sparse solve of the oracle's system."""
    import scipy.sparse.linalg as spla
    mesh = meshgen.beam(120, 30, kind=kind, jitter=0.2, seed=3)
    s = D.beam_system(mesh, gravity=(0.0, -9.81))
    A, b = oracle_csr(mesh, True, forces=[(0.0, -9.81)])
    Kg = s.get_csr_scipy()
    Ko = A.to_scipy()
    diff = abs(Kg - Ko)
    assert diff.max() <= K_TOL * abs(Ko).max()
    close_entries(s.get_rhs(), b)
    u_ref = spla.spsolve(Ko.tocsc(), b)
    u, st = s.solve(rtol=1e-9)
    assert st["status"] == 0
    err = rel_err(u, u_ref)
    # block Jacobi reaches the same solution
    u2, st2 = s.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-9)
    err2 = rel_err(u2, u_ref)
    print(f"mid kind {kind}: jacobi {st['iterations']} it, error {err:.2e}; block {st2['iterations']} it, "
          f"error {err2:.2e}")
    assert err <= U_TOL and err2 <= U_TOL


def test_determinism():
    mesh = meshgen.beam(64, 16, kind=4, jitter=0.2, seed=9)
    s1 = D.beam_system(mesh, gravity=(0.0, -9.81))
    s2 = D.beam_system(mesh, gravity=(0.0, -9.81))
    K1, K2 = s1.get_bsr()[2], s2.get_bsr()[2]
    assert np.array_equal(K1, K2)
    u1, _ = s1.solve(rtol=1e-9)
    u2, _ = s2.solve(rtol=1e-9)
    u3, _ = s1.solve(rtol=1e-9)
    assert np.array_equal(u1, u2) and np.array_equal(u1, u3)


def test_renumbering_invariance():
    mesh = meshgen.beam(50, 12, kind=4, jitter=0.2, seed=4)
    a = D.beam_system(mesh, renumber=True)
    b = D.beam_system(mesh, renumber=False)
    p = a.node_perm()
    assert np.array_equal(np.sort(p), np.arange(mesh.n_nodes))
    ua, _ = a.solve(rtol=1e-9)
    ub, _ = b.solve(rtol=1e-9)
    err = rel_err(ua, ub)
    print(f"renumbered vs natural order: {err:.2e}")
    assert err <= U_TOL


@pytest.mark.parametrize("kind", [4, 3])
def test_heterogeneous_E_against_oracle(kind):
    """Per-element E drawn log-uniform over [1e9, 1e11] (the mixed-precision
    and learned workloads' material) on a jittered 120 x 30 beam with a body
    load: every K entry and b against the oracle's CSR assembly with the same
    elem_E (reference: the single material E of libbfm/src/system.c:454-458,
    taken per element), u against a direct solve of the oracle's system; then
    the same through a batch built from the mesh (its own assembly kernel)."""
    import scipy.sparse.linalg as spla
    mesh = meshgen.beam(120, 30, kind=kind, jitter=0.2, seed=5)
    mesh.elem_E = meshgen.loguniform_E(mesh.n_elems, seed=5)
    assert mesh.elem_E.min() >= 1e9 and mesh.elem_E.max() <= 1e11
    g = (0.0, -9.81)
    s = D.beam_system(mesh, gravity=g)
    A, b = oracle_csr(mesh, True, forces=[g])
    Ko = A.to_scipy()
    Kg = s.get_csr_scipy()
    assert abs(Kg - Ko).max() <= K_TOL * abs(Ko).max()
    close_entries(s.get_rhs(), b)
    u_ref = spla.spsolve(Ko.tocsc(), b)
    u, st = s.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-9)
    assert st["status"] == 0
    err = rel_err(u, u_ref)
    mat = meshgen.BEAM_MATERIAL
    kinds, values, flags = meshgen.beam_conditions(mesh)
    pb = D.Batch.from_instances([D.planar_instance(mesh, True, mat, kinds, values, flags, g=g,
                                                   elem_E=mesh.elem_E)])
    stb, _ = pb.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-9)
    own = pb.system(0)
    assert abs(own.get_csr_scipy() - Ko).max() <= K_TOL * abs(Ko).max()
    close_entries(own.get_rhs(), b)
    err_b = rel_err(pb.get_u(0), u_ref)
    print(f"heterogeneous E, kind {kind}: E in [{mesh.elem_E.min():.3g}, {mesh.elem_E.max():.3g}], "
          f"system error {err:.2e} ({st['iterations']} it), batch error {err_b:.2e}")
    assert stb[0]["status"] == 0
    assert err <= U_TOL and err_b <= U_TOL


def test_mixed_precision_reaches_fp64_residual():
    mesh = meshgen.beam(200, 50, kind=4, jitter=0.2, seed=2)
    mesh.elem_E = meshgen.loguniform_E(mesh.n_elems, seed=2)
    s = D.beam_system(mesh)
    u64, st64 = s.solve(rtol=1e-9)
    u32, st32 = s.solve(precision=D.MIXED, rtol=1e-9)
    assert st32["status"] == 0 and st32["rel_residual"] <= 1e-9
    # error of the mixed solution against the FP64 one, stated
    err = rel_err(u32, u64)
    print(f"mixed vs fp64: rel error {err:.3e}, iterations {st32['iterations']} vs {st64['iterations']}")
    assert err <= 1e-7


def test_zero_load_gives_zero_solution():
    mesh = meshgen.beam(8, 4, kind=4)
    s = D.System.from_mesh(mesh)
    s.assemble(1, 1e9, 0.3, 0.0)
    s.apply_condition(0, 0.0, mesh.node_sets["left"])
    s.apply_condition(1, 0.0, mesh.node_sets["left"])
    u, st = s.solve()
    assert st["status"] == 0 and st["iterations"] == 0
    assert np.all(u == 0)


def test_singular_system_reports_failure():
    mesh = meshgen.beam(8, 4, kind=4)
    s = D.System.from_mesh(mesh)
    s.assemble(1, 1e9, 0.3, 1000.0, g=(0.3, -9.81))
    with pytest.raises(D.BfmxError) as e:
        s.solve(rtol=1e-12, max_iter=4000)
    assert e.value.code in (-4, -5, -8)


def test_tolerance_below_rounding_floor_reports_stagnation():
    """A tolerance below what the FP64 iterate can reach (the residual is
    evaluated compensated, so the floor is the iterate's own: a few 1e-12 on
    this beam) ends with BFMX_E_STAGNATION, not the iteration cap."""
    mesh = meshgen.beam(40, 10, kind=3)
    s = D.beam_system(mesh)
    with pytest.raises(D.BfmxError) as e:
        s.solve(rtol=1e-15)
    assert e.value.code == -8


def test_invalid_inputs_are_rejected():
    with pytest.raises(D.BfmxError) as e:
        D.System(4, np.zeros((4, 2)), np.array([[0, 1, 2, 7]]))
    assert e.value.code == -1
    with pytest.raises(D.BfmxError):
        D.System(5, np.zeros((4, 2)), np.array([[0, 1, 2, 3, 0]]))
    mesh = meshgen.beam(4, 2, kind=4)
    s = D.System.from_mesh(mesh)
    with pytest.raises(D.BfmxError):
        s.solve()          # not assembled


def ld_residual_norm(rowptr, col, val, b, u):
    """||b - K u|| for a CSR system, with the products and sums in long
    double (64-bit significand): the residual of the FP64 u itself, not the
    rounding of its evaluation in FP64."""
    ld = np.longdouble
    prod = val.astype(ld) * u[col].astype(ld)
    r = b.astype(ld) - np.add.reduceat(prod, rowptr[:-1])
    return float(np.sqrt((r * r).sum()))


def check_against_oracle(s, A, b, u, rtol):
    """The solve's claim, then parity: u meets rtol on the GPU's own system
    (K and b read back), and on the oracle's independently assembled system
    within rtol plus the two systems' own difference ||dK u - db|| / ||b||
    (their entries agree to ~1e-15 relative, summed in another order).
    Returns (residual on the GPU's system, on the oracle's, the difference)."""
    Kg = s.get_csr_scipy()
    Kg.sort_indices()
    bg = s.get_rhs()
    nb = np.linalg.norm(b)
    res_g = ld_residual_norm(Kg.indptr, Kg.indices, Kg.data, bg, u) / np.linalg.norm(bg)
    res_o = ld_residual_norm(A.rowptr, A.col, A.val, b, u) / nb
    dK = (Kg - A.to_scipy()).tocsr()
    dK.sort_indices()
    diff = ld_residual_norm(dK.indptr, dK.indices, dK.data, bg - b, u) / nb   # ||db - dK u||
    assert res_g <= rtol, res_g
    assert res_o <= rtol + diff, (res_o, diff)
    return res_g, res_o, diff


def test_full_size_beam_residual_against_oracle():
    """The 2000 x 500 Q1 beam (2,005,002 DOFs): the GPU solutions (Jacobi PCG
    and multigrid FCG, rtol 1e-8) satisfy the oracle's own assembled system
    to 1e-8, the residual evaluated in long double."""
    mesh, stress = meshgen.workload("q1_2m")
    s = D.beam_system(mesh, stress=stress)
    u, st = s.solve(rtol=1e-8)
    assert st["status"] == 0 and st["rel_residual"] <= 1e-8
    A, b = oracle_csr(mesh, stress)
    res_g, res_o, diff = check_against_oracle(s, A, b, u, 1e-8)
    print(f"Jacobi PCG ({st['iterations']} iterations): relative residual {res_g:.3e} on the GPU's system, "
          f"{res_o:.3e} on the oracle's (systems differ by {diff:.1e})")
    assert rel_err(s.get_rhs(), b) <= 1e-14
    u2, st2 = s.solve(precond=D.PRECOND_AMG, rtol=1e-8)
    assert st2["status"] == 0 and st2["rel_residual"] <= 1e-8
    res_g, res_o, diff = check_against_oracle(s, A, b, u2, 1e-8)
    print(f"multigrid FCG ({st2['iterations']} iterations, {st2['solve_ms']:.1f} ms vs {st['solve_ms']:.1f} ms): "
          f"relative residual {res_g:.3e} on the GPU's system, {res_o:.3e} on the oracle's; difference to PCG "
          f"{rel_err(u2, u):.2e}")
    assert rel_err(u2, u) <= 1e-7


def test_full_size_heterogeneous_beam_to_1e8():
    """The mixed-precision workload at full size: the 2000 x 500 Q1 beam
    (2,005,002 DOFs) with E per element log-uniform over [1e9, 1e11] (contrast
    100), solved by flexible CG in FP64 preconditioned by the FP32 multigrid
    V-cycle (every SpMV of the V-cycle in FP32) to a 1e-8 FP64 relative
    residual. The residual holds against the oracle's independently assembled
    system (long double evaluation), and the error against a tightly
    converged FP64 solution (multigrid FCG down to its rounding floor,
    ~7e-10) is stated."""
    mesh, stress = meshgen.workload("q1_2m_het")
    assert mesh.elem_E.min() >= 1e9 and mesh.elem_E.max() <= 1e11
    s = D.beam_system(mesh, stress=stress)
    u, st = s.solve(precond=D.PRECOND_AMG, rtol=1e-8)
    assert st["status"] == 0 and st["rel_residual"] <= 1e-8
    try:
        u_t, st_t = s.solve(precond=D.PRECOND_AMG, rtol=1e-11)
        floor = st_t["rel_residual"]
    except D.BfmxError as e:   # stops at the rounding floor, well below 1e-8
        assert e.code == -8, e
        u_t = s.get_u()
        floor = float(str(e).split("residual ")[1].split()[0])
    assert floor <= 2e-9
    err = rel_err(u, u_t)
    A, b = oracle_csr(mesh, stress)
    res_g, res_o, diff = check_against_oracle(s, A, b, u, 1e-8)
    print(f"heterogeneous 2M beam: {st['iterations']} iterations, residual {res_g:.3e} on the GPU's system, "
          f"{res_o:.3e} on the oracle's (systems differ by {diff:.1e}); floor {floor:.2e}; error vs the tight "
          f"solve {err:.2e}")
    assert err <= 1e-8


# ---------------------------------------------------------------- multigrid FCG


def test_amg_heterogeneous_beam_matches_fp64_pcg():
    """Per-element E over two decades (log-uniform [1e9, 1e11]): the FP32
    multigrid inside FP64 flexible CG reaches the same FP64 residual as plain
    PCG, in far fewer iterations; the error between the two is stated."""
    mesh = meshgen.beam(200, 50, kind=4, jitter=0.2, seed=3)
    mesh.elem_E = meshgen.loguniform_E(mesh.n_elems, seed=3)
    s = D.beam_system(mesh)
    u_pcg, st_pcg = s.solve(rtol=1e-9)    # the residual floor of this system is ~6e-10
    u_amg, st_amg = s.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    assert st_amg["status"] == 0 and st_amg["rel_residual"] <= 1e-9
    err = rel_err(u_amg, u_pcg)
    print(f"amg vs pcg: {st_amg['iterations']} vs {st_pcg['iterations']} iterations, rel difference {err:.2e}")
    assert err <= 1e-6
    assert st_amg["iterations"] * 5 < st_pcg["iterations"]


def test_amg_p1_mesh_against_oracle():
    import oracle as O
    mesh = meshgen.beam(60, 15, kind=3, jitter=0.15, seed=4)
    mat = meshgen.BEAM_MATERIAL
    kinds, values, flags = meshgen.beam_conditions(mesh)
    s = D.beam_system(mesh, stress=False)
    u, st = s.solve(precond=D.PRECOND_AMG, rtol=1e-10)
    assert st["status"] == 0
    u_ref, *_ = O.solve_dense(mesh, False, mat["E"], mat["nu"], mat["rho"], kinds, values, flags)
    err = rel_err(u, u_ref)
    print(f"P1 multigrid: {st['iterations']} iterations, error vs oracle {err:.2e}")
    assert err <= 1e-8


def test_amg_is_deterministic_and_reusable():
    """Two solves give bitwise identical solutions; re-assembly with other
    conditions re-runs the numeric set-up (the hierarchy topology and the
    captured graphs stay)."""
    mesh = meshgen.beam(120, 30, kind=4, jitter=0.2, seed=5)
    s = D.beam_system(mesh)
    u1, st1 = s.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    u2, st2 = s.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    assert st1["iterations"] == st2["iterations"]
    assert np.array_equal(u1, u2)
    # clamp the right edge too: new Dirichlet set, new matrix
    mat = meshgen.BEAM_MATERIAL
    s.assemble(1, mat["E"], mat["nu"], mat["rho"])
    kinds, values, flags = meshgen.beam_conditions(mesh)
    s.apply_conditions(kinds, values, flags)
    s.apply_condition(D.DIRICHLET_X, 0.0, mesh.node_sets["right"])
    s.apply_condition(D.DIRICHLET_Y, 0.0, mesh.node_sets["right"])
    u3, st3 = s.solve(precond=D.PRECOND_AMG, rtol=1e-10)
    u4, st4 = s.solve(precond=D.PRECOND_JACOBI, rtol=1e-10)
    assert st3["status"] == 0
    err = rel_err(u3, u4)
    print(f"re-assembled: {st3['iterations']} vs {st4['iterations']} iterations, rel difference {err:.2e}")
    assert err <= 1e-8
    assert np.abs(u3).max() < np.abs(u1).max()   # both ends clamped: stiffer


def test_amg_smoothed_against_unsmoothed_aggregation(monkeypatch):
    """Smoothed aggregation (the default: P = (I - w D^-1 A) P_t) and plain
    aggregation (BFMX_AMG_SMOOTH=0, read when a system builds its hierarchy)
    reach the same solution; smoothing the prolongator cuts the iterations."""
    mesh = meshgen.beam(240, 60, kind=4, jitter=0.2, seed=8)
    s_sa = D.beam_system(mesh)
    u_sa, st_sa = s_sa.solve(precond=D.PRECOND_AMG, rtol=1e-10)
    monkeypatch.setenv("BFMX_AMG_SMOOTH", "0")
    s_ua = D.beam_system(mesh)
    u_ua, st_ua = s_ua.solve(precond=D.PRECOND_AMG, rtol=1e-10)
    assert st_sa["status"] == 0 and st_ua["status"] == 0
    err = rel_err(u_sa, u_ua)
    print(f"smoothed {st_sa['iterations']} vs unsmoothed {st_ua['iterations']} iterations, difference {err:.2e}")
    assert err <= 1e-8
    assert st_sa["iterations"] * 2 <= st_ua["iterations"]


@pytest.mark.parametrize("kind", [4, 3])
@pytest.mark.parametrize("host_agg", ["0", "levels", None])
def test_amg_device_and_host_hierarchies(monkeypatch, kind, host_agg):
    """The hierarchy topology built on the GPU (the default: aggregates by
    host sweeps up to 131072 rows, the transfer operators and the two-step
    Galerkin product on the GPU; with BFMX_AMG_HOST_AGG=0 MIS-2 aggregates on
    the GPU at every level; with BFMX_AMG_HOST_LEVELS=16384 the levels from
    16384 rows down by the host build) and on the host (BFMX_AMG_BUILD=host:
    sequential aggregation, contribution lists) all lead flexible CG to the
    oracle's solution; two fresh systems built on the GPU give bitwise the
    same solution (the device build does not depend on scheduling)."""
    if host_agg == "0":
        monkeypatch.setenv("BFMX_AMG_HOST_AGG", "0")
    elif host_agg == "levels":
        monkeypatch.setenv("BFMX_AMG_HOST_LEVELS", "16384")
    import scipy.sparse.linalg as spla
    mesh = meshgen.beam(300, 75, kind=kind, jitter=0.2, seed=13)
    mesh.elem_E = meshgen.loguniform_E(mesh.n_elems, seed=13)
    A, b = oracle_csr(mesh, True)
    u_ref = spla.spsolve(A.to_scipy().tocsc(), b)
    dev1 = D.beam_system(mesh)
    u1, st1 = dev1.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    dev1.close()
    dev2 = D.beam_system(mesh)
    u2, st2 = dev2.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    monkeypatch.setenv("BFMX_AMG_BUILD", "host")
    host = D.beam_system(mesh)
    u3, st3 = host.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    print(f"kind {kind}, host_agg {host_agg}: device build {st1['iterations']} iterations "
          f"(error {rel_err(u1, u_ref):.2e}), "
          f"host build {st3['iterations']} (error {rel_err(u3, u_ref):.2e})")
    assert st1["status"] == 0 and st3["status"] == 0
    assert st1["iterations"] == st2["iterations"] and np.array_equal(u1, u2)
    assert rel_err(u1, u_ref) <= U_TOL and rel_err(u3, u_ref) <= U_TOL


@pytest.mark.parametrize("wgs", ["512", "0", "3"])
def test_amg_numeric_overlap_matches_one_stream(monkeypatch, wgs):
    """The first build queues each level's numeric set-up on a second stream
    while the host builds the levels below, the finest level's largest
    kernels held to BFMX_AMG_NUMERIC_WGS workgroups (grid-stride loops: the
    default 512, the natural grids, and 3, which makes each workgroup take
    many rows); the same set-up queued after the build on the system's
    stream (BFMX_AMG_OVERLAP=0, natural grids) gives bitwise the same solve,
    on a heterogeneous beam and after a reassembly with new E."""
    mesh = meshgen.beam(300, 75, kind=4, jitter=0.2, seed=17)
    mesh.elem_E = meshgen.loguniform_E(mesh.n_elems, seed=17)
    monkeypatch.setenv("BFMX_AMG_NUMERIC_WGS", wgs)
    out = []
    for overlap in ("1", "0"):
        monkeypatch.setenv("BFMX_AMG_OVERLAP", overlap)
        s = D.beam_system(mesh)
        u, st = s.solve(precond=D.PRECOND_AMG, rtol=1e-9)
        assert st["status"] == 0
        mat = meshgen.BEAM_MATERIAL
        s.assemble(1, mat["E"], mat["nu"], mat["rho"], elem_E=meshgen.loguniform_E(mesh.n_elems, seed=18))
        s.apply_conditions(*meshgen.beam_conditions(mesh))
        u2, st2 = s.solve(precond=D.PRECOND_AMG, rtol=1e-9)
        assert st2["status"] == 0
        s.close()
        out.append((u, st["iterations"], u2, st2["iterations"]))
    print(f"overlap ({wgs} workgroups): {out[0][1]} / {out[0][3]} iterations, one stream: {out[1][1]} / {out[1][3]}")
    assert out[0][1] == out[1][1] and np.array_equal(out[0][0], out[1][0])
    assert out[0][3] == out[1][3] and np.array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("kind", [3, 4])
def test_small_levels_host_transfer_matches_the_device_build(monkeypatch, kind):
    """Levels of at most 4096 rows swept on the host get their transfer
    topology (prolongator, restriction, A P and coarse patterns, coarse
    sliced ELL, mirrors) from the host in one upload; the device build of the
    same arrays (BFMX_AMG_HOST_TRANSFER=0) gives bitwise the same solve."""
    mesh = meshgen.beam(300, 75, kind=kind, jitter=0.2, seed=19)
    mesh.elem_E = meshgen.loguniform_E(mesh.n_elems, seed=19)
    out = []
    for rows in ("4096", "0"):
        monkeypatch.setenv("BFMX_AMG_HOST_TRANSFER", rows)
        s = D.beam_system(mesh)
        u, st = s.solve(precond=D.PRECOND_AMG, rtol=1e-9)
        assert st["status"] == 0
        s.close()
        out.append((u, st["iterations"]))
    print(f"kind {kind}: host transfer {out[0][1]} iterations, device {out[1][1]}")
    assert out[0][1] == out[1][1] and np.array_equal(out[0][0], out[1][0])


def test_finest_sweeps_do_not_depend_on_threads(monkeypatch):
    """A beam of 90,000 nodes: its finest level is aggregated by one
    sequential sweep on the host; one host thread and sixteen give bitwise
    the same solve, which meets the tolerance on the host's residual."""
    mesh = meshgen.beam(599, 149, kind=4, jitter=0.2, seed=21)
    assert mesh.n_nodes > 65536
    out = []
    for threads in ("1", "16"):
        monkeypatch.setenv("BFMX_AMG_THREADS", threads)
        s = D.beam_system(mesh)
        u, st = s.solve(precond=D.PRECOND_AMG, rtol=1e-8)
        assert st["status"] == 0
        out.append((st["iterations"], u))
        K, b = s.get_csr_scipy(), s.get_rhs()
        assert np.linalg.norm(b - K @ u) <= 1e-8 * np.linalg.norm(b)
        s.close()
    print(f"{2 * mesh.n_nodes} DOFs: {out[0][0]} iterations")
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])


def test_amg_reuses_its_set_up_only_for_unchanged_values():
    """A second multigrid solve of an unchanged system reuses the hierarchy's
    numeric set-up (bitwise the same solution); after the system is assembled
    again with other values (per-element E spread over two decades) the set-up
    follows the new matrix: the solve meets rtol on the host's residual of the
    new system and matches a direct solve of it."""
    mesh = meshgen.beam(120, 30, kind=4, jitter=0.2, seed=31)
    s = D.beam_system(mesh)
    u1, st1 = s.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    u1b, st1b = s.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    assert st1["status"] == 0 and np.array_equal(u1, u1b)
    mat = meshgen.BEAM_MATERIAL
    eE = meshgen.loguniform_E(mesh.n_elems, seed=31)
    s.assemble(1, mat["E"], mat["nu"], mat["rho"], elem_E=eE)
    kinds, values, flags = meshgen.beam_conditions(mesh)
    s.apply_conditions(kinds, values, flags)
    u2, st2 = s.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    K, b = s.get_csr_scipy(), s.get_rhs()
    assert st2["status"] == 0
    assert np.linalg.norm(b - K @ u2) <= 1e-9 * np.linalg.norm(b)
    import scipy.sparse.linalg as spla
    u_ref = spla.spsolve(K.tocsc(), b)
    print(f"uniform E {st1['iterations']} iterations; reassembled {st2['iterations']}; "
          f"error vs a direct solve {rel_err(u2, u_ref):.2e}")
    assert rel_err(u2, u_ref) <= 1e-5


def test_amg_single_level_is_a_direct_solve():
    """A mesh small enough for one level: the preconditioner is the dense
    FP32 inverse, so FCG stops after very few iterations."""
    mesh = meshgen.beam(6, 3, kind=4)
    s = D.beam_system(mesh)
    u, st = s.solve(precond=D.PRECOND_AMG, rtol=1e-10)
    u_ref, _ = s.solve(precond=D.PRECOND_JACOBI, rtol=1e-12)
    assert st["status"] == 0 and st["iterations"] <= 4
    assert rel_err(u, u_ref) <= 1e-9


def test_kernel_timing_entry_point():
    mesh = meshgen.beam(100, 25, kind=4)
    s = D.beam_system(mesh)
    for k in (D.KERNEL_CG_SPMV64, D.KERNEL_CG_SPMV32, D.KERNEL_FCG_SPMV, D.KERNEL_AMG_SMOOTH):
        ms, by = s.time_kernel(k, reps=5)
        assert ms > 0 and by > 0
    u, st = s.solve(precond=D.PRECOND_AMG, rtol=1e-9)   # timing leaves the system solvable
    assert st["status"] == 0
