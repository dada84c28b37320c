"""The CPU oracle pinned against the reference and against the mathematics.

1. Bit-exact against fixtures written by libbfm compiled from its own sources
   (tests/golden/make_golden.py): dense K and b before and after every kind
   of boundary condition, element matrices, and the displacements of the
   full run_elasticity path (RCM + band LU).
2. Against the reference's own data files: data/U.txt and data/V.txt for
   meshes/8.lepl1110 + problems/problem.txt (8 significant digits).
3. Against closed forms that a plausible slip (dropped term, wrong sign,
   transposed operand, wrong quadrature point) would break: rigid-body null
   space, the patch test, element areas, Neumann totals, the D-constant
   identity a - b = 2c, band LU against numpy's dense solve, and the
   Euler-Bernoulli tip deflection as an order-of-magnitude check. One test
   plants mistakes into oracle output and shows these checks reject them.
4. Self-consistency of the sparse path: CSR entries equal the dense ones,
   and the CSR Jacobi-PCG matches the direct solve.
"""

import os

import numpy as np
import pytest

import oracle as O
from conftest import REF_DATA, dense_from_golden, rel_err
from make_golden import TINY_CONDITIONS, node_sets

from bfm_amd import meshgen
from bfm_amd.pybfm import read_effects_file

MAT = meshgen.BEAM_MATERIAL
C1_CONDS = [(3, meshgen.BEAM_TRACTION, "top"), (0, 0.0, "left"), (1, 0.0, "left")]


def conds_arrays(mesh, conds, sets=None):
    sets = sets or mesh.node_sets
    kinds = [c[0] for c in conds]
    values = [c[1] for c in conds]
    flags = np.stack([sets[c[2]] for c in conds]).astype(np.uint8)
    return kinds, values, flags


# ------------------------------------------------------------ 1. reference


@pytest.mark.parametrize("name,conds,grav", [
    ("c1", C1_CONDS, False),
    ("c1_gravity", C1_CONDS, True),
    ("c1_neumann_last", C1_CONDS[1:] + C1_CONDS[:1], False),
])
def test_c1_bit_exact(golden, name, conds, grav):
    mesh = meshgen.beam(40, 10, kind=3)
    kinds, values, flags = conds_arrays(mesh, conds)
    u, K, b, perm, k, rc = O.solve_dense(mesh, True, MAT["E"], MAT["nu"], MAT["rho"], kinds, values, flags,
                                         forces=[(0.0, -9.81)] if grav else None)
    assert rc == 0
    assert np.array_equal(K, dense_from_golden(golden, name, mesh.n_dof))
    assert np.array_equal(b, golden[f"{name}/b"])
    assert np.array_equal(u, golden[f"{name}/u"])


@pytest.mark.parametrize("kind,tag", [(4, "q1"), (3, "p1")])
@pytest.mark.parametrize("stress", [0, 1])
def test_tiny_bit_exact(golden, kind, tag, stress):
    name = f"tiny_{tag}_{'stress' if stress else 'strain'}"
    mesh = meshgen.beam(3, 2, kind=kind, length=3.0, height=2.0, jitter=0.2, seed=11)
    kinds, values, flags = conds_arrays(mesh, TINY_CONDITIONS, node_sets(mesh))
    K0, b0 = O.assemble_dense(mesh, stress, 7.0e10, 0.33, 2700.0, forces=[(0.0, -9.81)])
    assert np.array_equal(K0, golden[f"{name}/K_pre"])
    assert np.array_equal(b0, golden[f"{name}/b_pre"])
    u, K, b, perm, k, rc = O.solve_dense(mesh, stress, 7.0e10, 0.33, 2700.0, kinds, values, flags,
                                         forces=[(0.0, -9.81)])
    assert rc == 0
    assert np.array_equal(K, golden[f"{name}/K"])
    assert np.array_equal(b, golden[f"{name}/b"])
    assert np.array_equal(u, golden[f"{name}/u"])


@pytest.mark.parametrize("tag", ["q1", "p1"])
@pytest.mark.parametrize("t", [0, 1, 2])
def test_element_matrices_bit_exact(golden, tag, t):
    pre = f"elem_{tag}_{t}"
    elems = golden[f"{pre}/elems"]
    mesh = meshgen.Mesh(elems.shape[1], golden[f"{pre}/coords"], elems.astype(np.uint64),
                        np.zeros((0, 2), np.uint64), np.zeros(0, np.int64))
    K, b = O.assemble_dense(mesh, int(golden[f"{pre}/stress"]), 2.0e11, float(golden[f"{pre}/nu"]), 7800.0,
                            forces=[(0.0, -9.81)])
    assert np.array_equal(K, golden[f"{pre}/K"])
    assert np.array_equal(b, golden[f"{pre}/b"])


def lepl8_mesh(golden):
    return meshgen.Mesh(4, golden["lepl8/coords"].reshape(-1, 2), golden["lepl8/elems"].astype(np.uint64),
                        golden["lepl8/edge_nodes"].astype(np.uint64), golden["lepl8/edge_elems"][:, 0])


def test_lepl1110_against_reference_and_data_files(golden):
    mesh = lepl8_mesh(golden)
    E, nu, rho = golden["lepl8/E_nu_rho"]
    n = int(golden["lepl8/n_conditions"])
    kinds = [int(golden[f"lepl8/cond{i}_kind_value"][0]) for i in range(n)]
    values = [golden[f"lepl8/cond{i}_kind_value"][1] for i in range(n)]
    u, *_, rc = O.solve_dense(mesh, int(golden["lepl8/sim_kind"]) == 2, E, nu, rho, kinds, values,
                              golden["lepl8/cond_flags"], forces=[(0.0, -9.81)])
    assert rc == 0
    assert np.array_equal(u, golden["lepl8/u"])
    U = read_effects_file(os.path.join(REF_DATA, "U.txt"))
    V = read_effects_file(os.path.join(REF_DATA, "V.txt"))
    assert np.abs(u[0::2] - U).max() <= 5e-8 * np.abs(U).max()
    assert np.abs(u[1::2] - V).max() <= 5e-8 * np.abs(V).max()


# ------------------------------------------------------------ 3. closed forms


def rigid_modes(coords):
    x, y = coords[:, 0], coords[:, 1]
    n = coords.shape[0]
    tx = np.zeros(2 * n)
    tx[0::2] = 1
    ty = np.zeros(2 * n)
    ty[1::2] = 1
    rot = np.zeros(2 * n)
    rot[0::2] = -y
    rot[1::2] = x
    return [tx, ty, rot]


def element_invariant_violations(K, coords):
    """Symmetry, positive semi-definiteness, 3 rigid-body modes in the kernel."""
    scale = np.abs(K).max()
    out = []
    if np.abs(K - K.T).max() > 1e-12 * scale:
        out.append("symmetry")
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    if ev.min() < -1e-10 * scale:
        out.append("psd")
    for v in rigid_modes(coords):
        if np.abs(K @ v).max() > 1e-9 * scale * np.abs(v).max():
            out.append("rigid")
    if (np.abs(ev) <= 1e-10 * scale).sum() != 3:
        out.append("nullity")
    return out


@pytest.mark.parametrize("kind", [3, 4])
@pytest.mark.parametrize("stress", [0, 1])
def test_element_invariants(kind, stress):
    rng = np.random.default_rng(kind * 10 + stress)
    for _ in range(5):
        m = meshgen.beam(1, 1, kind=kind)
        if kind == 3:
            m.elems = m.elems[:1]
            m.coords = m.coords[:3] if False else m.coords
        m.coords = m.coords + rng.uniform(-0.2, 0.2, size=m.coords.shape)
        K, _ = O.assemble_dense(m, stress, 1e9, 0.3, 1.0)
        used = np.unique(m.elems.astype(np.int64))
        dofs = np.stack([2 * used, 2 * used + 1], 1).reshape(-1)
        Ke = K[np.ix_(dofs, dofs)]
        assert element_invariant_violations(Ke, m.coords[used]) == []


def test_invariants_reject_planted_mistakes():
    m = meshgen.beam(1, 1, kind=4)
    m.coords = m.coords + np.random.default_rng(0).uniform(-0.2, 0.2, size=m.coords.shape)
    K, _ = O.assemble_dense(m, 1, 1e9, 0.3, 1.0)
    coords = m.coords
    assert element_invariant_violations(K, coords) == []
    bad = K.copy()
    bad[0:2, 2:4] = bad[0:2, 2:4].T          # one transposed coupling block
    assert element_invariant_violations(bad, coords)
    bad = K.copy()
    bad[1::2, 0::2] *= -1                    # wrong sign of the shear/Poisson coupling rows
    bad[0::2, 1::2] *= -1
    assert element_invariant_violations(bad, coords)
    a, b, c = O.dconst(1, 1e9, 0.3)
    # dropping the c (shear) term: K_drop = K - c * (shear part); emulate by a
    # material with a wrong constant: a - b != 2c breaks isotropy
    assert abs((a - b) - 2 * c) <= 1e-6 * a
    assert abs((a - 0.5 * b) - 2 * c) > 1e-3 * a


@pytest.mark.parametrize("kind", [3, 4])
def test_patch_test(kind):
    """A linear displacement field produces no interior nodal forces."""
    mesh = meshgen.beam(6, 4, kind=kind, jitter=0.2, seed=1)
    K, _ = O.assemble_dense(mesh, 1, 2e11, 0.3, 0.0)
    x, y = mesh.coords[:, 0], mesh.coords[:, 1]
    u = np.zeros(mesh.n_dof)
    u[0::2] = 1e-3 * x + 2e-3 * y + 0.1
    u[1::2] = -3e-3 * x + 0.5e-3 * y - 0.2
    f = K @ u
    interior = ~(mesh.node_sets["left"] | mesh.node_sets["right"] | mesh.node_sets["top"] |
                 mesh.node_sets["bottom"])
    fi = np.stack([f[0::2], f[1::2]], 1)[interior]
    assert np.abs(fi).max() <= 1e-9 * np.abs(K).max() * np.abs(u).max()


@pytest.mark.parametrize("kind", [3, 4])
def test_body_load_integrates_area(kind):
    mesh = meshgen.beam(9, 5, kind=kind, length=3.0, height=2.0, jitter=0.2, seed=2)
    _, b = O.assemble_dense(mesh, 0, 1e9, 0.3, 2.0, forces=[(0.0, 1.0)])
    assert b[0::2].sum() == 0
    assert abs(b[1::2].sum() - 2.0 * 3.0 * 2.0) <= 1e-12 * 12
    # two forces add up like their sum
    _, b2 = O.assemble_dense(mesh, 0, 1e9, 0.3, 2.0, forces=[(0.0, 0.25), (0.0, 0.75)])
    assert np.allclose(b2, b, rtol=1e-14, atol=0)


def test_neumann_totals():
    mesh = meshgen.beam(13, 4, kind=4, length=4.0, height=1.0, jitter=0.2, seed=3)
    m = mesh.n_dof
    Z = np.eye(m)
    b0 = np.zeros(m)
    top = mesh.node_sets["top"].astype(np.uint8)[None]
    _, b = O.apply_conditions_dense(mesh, Z, b0, [3], [-2.5e5], top)
    assert abs(b[1::2].sum() - (-2.5e5 * 4.0)) <= 1e-9 * 1e6
    assert b[0::2].sum() == 0
    # normal traction on the top edge (edges run right to left): along +y... the
    # reference takes (-dy, dx) of the edge vector n1 - n2, i.e. (0, +h) here
    _, bn = O.apply_conditions_dense(mesh, Z, b0, [4], [1.0], top)
    assert abs(bn[1::2].sum() - 4.0) <= 1e-12 and abs(bn[0::2].sum()) <= 1e-12
    _, bt = O.apply_conditions_dense(mesh, Z, b0, [5], [1.0], top)
    assert abs(bt[0::2].sum() - 4.0) <= 1e-12 and abs(bt[1::2].sum()) <= 1e-12


def test_dirichlet_values_and_symmetric_elimination():
    mesh = meshgen.beam(5, 3, kind=3, jitter=0.1, seed=4)
    K, b = O.assemble_dense(mesh, 1, 1e9, 0.3, 1.0, forces=[(0.0, -1.0)])
    left = mesh.node_sets["left"].astype(np.uint8)
    Kc, bc = O.apply_conditions_dense(mesh, K, b, [0, 1], [2e-3, -1e-3], np.stack([left, left]))
    fixed = np.zeros(mesh.n_dof, bool)
    fixed[0::2] = left.astype(bool)
    fixed[1::2] = left.astype(bool)
    assert np.abs(Kc - Kc.T).max() <= 1e-12 * np.abs(K).max()
    assert np.array_equal(Kc[np.ix_(fixed, fixed)], np.eye(fixed.sum()))
    u = np.linalg.solve(Kc, bc)
    assert np.allclose(u[0::2][left.astype(bool)], 2e-3, rtol=1e-12)
    assert np.allclose(u[1::2][left.astype(bool)], -1e-3, rtol=1e-12)
    # the free part solves the original equations with the prescribed values moved right
    free = ~fixed
    r = (K @ u - b)[free]
    assert np.abs(r).max() <= 1e-12 * np.abs(K).max() * np.abs(u).max()


def test_dconst_identity():
    for stress in (0, 1):
        for nu in (0.0, 0.25, 0.3, 0.45):
            a, b, c = O.dconst(stress, 7e10, nu)
            assert abs((a - b) - 2 * c) <= 1e-12 * a
            assert abs(b - (nu / (1 - nu) if not stress else nu) * a) <= 1e-12 * a


def test_band_lu_matches_numpy():
    rng = np.random.default_rng(0)
    m, k = 40, 5
    A = np.zeros((m, m))
    for i in range(m):
        for j in range(max(0, i - k), min(m, i + k + 1)):
            A[i, j] = rng.normal()
    A = A @ A.T + m * np.eye(m)           # SPD, half-bandwidth 2k
    y = rng.normal(size=m)
    x, rc, rc2 = O.band_solve(A, y)
    assert rc == 0 and rc2 == 0
    assert rel_err(x, np.linalg.solve(A, y)) <= 1e-12


def test_band_lu_rejects_zero_pivot():
    A = np.array([[0.0, 1.0], [1.0, 0.0]])
    _, rc, _ = O.band_solve(A, np.ones(2))
    assert rc != 0


def test_rcm_is_a_bijection_and_reduces_bandwidth():
    mesh = meshgen.beam(20, 5, kind=4, jitter=0.2, seed=5)
    K, _ = O.assemble_dense(mesh, 1, 1e9, 0.3, 0.0)
    perm, inv = O.rcm(K)
    assert np.array_equal(np.sort(perm), np.arange(mesh.n_dof))
    assert np.array_equal(perm[inv], np.arange(mesh.n_dof))
    Kp = np.zeros_like(K)
    Kp[np.ix_(perm, perm)] = K
    bw = lambda A: max(abs(i - j) for i, j in zip(*np.nonzero(A)))
    assert bw(Kp) < bw(K)


def test_beam_tip_deflection_order_of_magnitude():
    """Euler-Bernoulli qL^4/(8EI) = 1.82e-3 m; coarse P1 is stiffer (locking)."""
    mesh = meshgen.beam(40, 10, kind=3)
    kinds, values, flags = meshgen.beam_conditions(mesh)
    u, *_, rc = O.solve_dense(mesh, True, MAT["E"], MAT["nu"], MAT["rho"], kinds, values, flags)
    tip = mesh.node_sets["right"]
    v_tip = u[1::2][tip].mean()
    eb = 1e6 * 4.0 ** 4 / (8 * MAT["E"] * (1.0 / 12))
    assert -1.1 * eb < v_tip < -0.3 * eb


# ------------------------------------------------------------ 4. sparse path


@pytest.mark.parametrize("kind", [3, 4])
def test_csr_equals_dense(kind):
    mesh = meshgen.beam(12, 5, kind=kind, jitter=0.2, seed=6)
    kinds, values, flags = meshgen.beam_conditions(mesh)
    A, b = O.assemble_csr(mesh, 1, MAT["E"], MAT["nu"], MAT["rho"], forces=[(0.1, -9.81)], kinds=kinds,
                          values=values, flags=flags)
    K, bd = O.assemble_dense(mesh, 1, MAT["E"], MAT["nu"], MAT["rho"], forces=[(0.1, -9.81)])
    Kd, bdd = O.apply_conditions_dense(mesh, K, bd, kinds, values, flags)
    assert np.array_equal(A.to_scipy().toarray(), Kd)
    assert np.array_equal(b, bdd)


def test_csr_pcg_matches_direct():
    mesh = meshgen.beam(30, 8, kind=4, jitter=0.2, seed=7)
    kinds, values, flags = meshgen.beam_conditions(mesh)
    A, b = O.assemble_csr(mesh, 1, MAT["E"], MAT["nu"], MAT["rho"], kinds=kinds, values=values, flags=flags)
    u_d, *_ = O.solve_dense(mesh, True, MAT["E"], MAT["nu"], MAT["rho"], kinds, values, flags)
    x, it, rr = O.pcg_csr(A, b, rtol=1e-11, threads=2)
    assert it > 0 and rr <= 1e-11
    assert rel_err(x, u_d) <= 1e-8
    r = b - A.matvec(x)
    assert np.linalg.norm(r) / np.linalg.norm(b) <= 1e-10


@pytest.mark.parametrize("kind,nx,ny", [(3, 40, 10), (4, 24, 6)])
def test_sparse_band_path_equals_the_dense_path(kind, nx, ny):
    """libbfm's direct path with the matrix kept sparse (RCM over the CSR
    pattern, band storage built from it, band LU and substitution) is the
    dense path bit for bit: the same permutation, half-bandwidth and u as
    run_elasticity restated (itself bit-identical to the compiled reference
    on the 40 x 10 P1 beam). A window of band rows away from the top holds
    exactly those rows' band entries whose columns fall in the window."""
    mesh = meshgen.beam(nx, ny, kind=kind, jitter=0.2 if kind == 4 else 0.0, seed=8)
    kinds, values, flags = meshgen.beam_conditions(mesh)
    u_d, _, _, perm_d, k_d, rc = O.solve_dense(mesh, True, MAT["E"], MAT["nu"], MAT["rho"], kinds, values, flags)
    assert rc == 0
    A, b = O.assemble_csr(mesh, 1, MAT["E"], MAT["nu"], MAT["rho"], kinds=kinds, values=values, flags=flags)
    perm, inv = O.rcm_csr(A)
    assert np.array_equal(perm, perm_d)
    k = O.bandwidth_csr(A, perm)
    assert k == k_d
    m = A.m
    band = O.band_window(A, perm, inv, k, 0, m)
    assert O.band_lu_steps(band, m, k, m) == 0
    y = np.zeros(m)
    y[perm.astype(np.int64)] = b
    assert O.band_lu_solve(band, m, k, y) == 0
    u = y[perm.astype(np.int64)]
    assert np.array_equal(u, u_d)
    r0, nrows = m // 3, min(m - m // 3, 2 * k + 5)
    win = O.band_window(A, perm, inv, k, r0, nrows)
    Pd = A.to_scipy().toarray()[np.ix_(inv.astype(np.int64), inv.astype(np.int64))]
    for li in range(nrows):
        for lj in range(max(0, li - k), min(nrows, li + k + 1)):
            assert win[lj + li * 2 * k] == Pd[r0 + li, r0 + lj]


@pytest.mark.parametrize("kind", [3, 4])
def test_per_element_E_is_linear_sum_of_unit_elements(kind):
    """The oracle's per-element-E branch (elem_E[e] in place of the material
    E of libbfm/src/system.c:454-458) pinned by linearity: at fixed nu the D
    constants are proportional to E, so K(elem_E) = sum_e E_e K_e(E = 1),
    each K_e assembled from a one-element mesh by the constant-E path (the
    path checked bit for bit against the reference above). The body load
    does not depend on E. An element-index slip in the elem_E branch
    (elem_E[e + 1], or E taken from the wrong element) breaks the sum."""
    mesh = meshgen.beam(6, 3, kind=kind, jitter=0.2, seed=8)
    rng = np.random.default_rng(8)
    mesh.elem_E = 10.0 ** rng.uniform(9, 11, mesh.n_elems)
    nu, rho, g = 0.3, 7850.0, [(0.0, -9.81)]
    for stress in (0, 1):
        K, b = O.assemble_dense(mesh, stress, 0.0, nu, rho, forces=g)
        K_sum = np.zeros_like(K)
        for e in range(mesh.n_elems):
            nodes = mesh.elems[e].astype(np.int64)
            one = meshgen.Mesh(kind, mesh.coords[nodes].copy(), np.arange(kind, dtype=np.uint64)[None, :],
                               np.zeros((0, 2), np.uint64), np.zeros(0, np.int64))
            Ke, _ = O.assemble_dense(one, stress, 1.0, nu, rho)
            dof = np.stack([2 * nodes, 2 * nodes + 1], axis=1).reshape(-1)
            K_sum[np.ix_(dof, dof)] += mesh.elem_E[e] * Ke
        scale = np.abs(K).max()
        assert np.abs(K - K_sum).max() <= 1e-13 * scale
        # the per-element values are the ones used: a shifted elem_E differs
        shifted = meshgen.Mesh(kind, mesh.coords, mesh.elems, mesh.edges, mesh.edge_elem,
                               elem_E=np.roll(mesh.elem_E, 1))
        K_shift, _ = O.assemble_dense(shifted, stress, 0.0, nu, rho, forces=g)
        assert np.abs(K_shift - K_sum).max() > 1e-3 * scale
        mesh_c = meshgen.Mesh(kind, mesh.coords, mesh.elems, mesh.edges, mesh.edge_elem)
        _, b_c = O.assemble_dense(mesh_c, stress, 2.0e11, nu, rho, forces=g)
        assert np.array_equal(b, b_c)
