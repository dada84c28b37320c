"""The libbfm host layer (bfm_amd/lib/libbfm.so) without a GPU.

Readers and the problem-file front end are compared with what the reference
parsed from the same files (tests/golden, written by libbfm compiled from its
sources); the U/V writer byte for byte with the reference's output; the host
dense matrix API with numpy and the oracle; failure paths are forced from
the host (allocator failure, missing GPU).
"""

import ctypes as C
import os

import numpy as np
import pytest

import oracle as O
from conftest import REF_DATA

from bfm_amd import _ffi as F
from bfm_amd import device as D
from bfm_amd import meshgen
from bfm_amd import pybfm as pb

LEPL = os.path.join(REF_DATA, "8.lepl1110")
PROBLEM = os.path.join(REF_DATA, "problem.txt")
OBJ = os.path.join(REF_DATA, "bridge-nothing.obj")


@pytest.fixture(scope="module")
def L():
    return pb.Library()


def test_lepl1110_reader_matches_reference(L, golden):
    m = pb.Mesh_lepl1110(LEPL, lib=L)
    assert m.c_mesh.dim == 2 and m.c_mesh.kind == 4
    assert np.array_equal(m.coords, golden["lepl8/coords"])
    assert np.array_equal(m.elems.astype(np.int64), golden["lepl8/elems"])
    en, ee = m.edges
    assert np.array_equal(en.astype(np.int64), golden["lepl8/edge_nodes"])
    assert np.array_equal(ee, golden["lepl8/edge_elems"])
    assert m.c_mesh.n_domains == 8
    names = [m.c_mesh.domains[i].name.decode() for i in range(8)]
    assert names[0] == "Symmetry" and names[7] == "Bottom"


def test_wavefront_reader_and_edges_match_reference(L, golden):
    m = pb.Mesh_wavefront(OBJ, lib=L)
    assert np.array_equal(m.coords, golden["obj/coords"])
    assert np.array_equal(m.elems.astype(np.int64), golden["obj/elems"])
    en, ee = m.edges
    assert np.array_equal(en.astype(np.int64), golden["obj/edge_nodes"])
    assert np.array_equal(ee, golden["obj/edge_elems"])


def test_reader_rejects_missing_sections(L, tmp_path):
    bad = tmp_path / "bad.lepl1110"
    bad.write_text("Number of nodes 1\n 0 : 0.0 0.0\nNumber of triangles 0\n")
    with pytest.raises(AssertionError):
        pb.Mesh_lepl1110(str(bad), lib=L)
    assert L.state.err.has


def test_ez_problem_file_matches_reference(L, golden):
    m = pb.Mesh_lepl1110(LEPL, lib=L)
    ez = pb.Ez_lepl1110(m, PROBLEM, lib=L)
    assert ez.c_ez.sim.kind == int(golden["lepl8/sim_kind"])
    mat = ez.c_ez.material
    assert [mat.E, mat.nu, mat.rho] == list(golden["lepl8/E_nu_rho"])
    n = int(golden["lepl8/n_conditions"])
    assert ez.c_ez.n_conditions == n and ez.c_ez.instance.n_conditions == n
    for i in range(n):
        c = ez.c_ez.conditions[i]
        kind, value = golden[f"lepl8/cond{i}_kind_value"]
        assert (c.kind, c.value) == (int(kind), value)
        flags = np.ctypeslib.as_array(C.cast(c.nodes, C.POINTER(C.c_uint8)), shape=(m.n_nodes,))
        assert np.array_equal(flags, golden["lepl8/cond_flags"][i])
    assert ez.c_ez.sim.n_forces == 1
    assert ez.c_ez.gravity.linear.force.data[1] == -9.81


def test_ez_rejects_unknown_condition(L, tmp_path):
    m = pb.Mesh_lepl1110(LEPL, lib=L)
    bad = tmp_path / "p.txt"
    bad.write_text("Type of problem    :  Planar strains\nBoundary condition :  Dirichlet-Q = 0.0 : Bottom\n")
    with pytest.raises(AssertionError):
        pb.Ez_lepl1110(m, str(bad), lib=L)


def test_effects_writer_is_byte_exact(L, golden, tmp_path):
    m = pb.Mesh_lepl1110(LEPL, lib=L)
    ez = pb.Ez_lepl1110(m, PROBLEM, lib=L)
    u = golden["lepl8/u"]
    eff = np.ctypeslib.as_array(ez.c_ez.instance.effects, shape=(u.size,))
    eff[:] = u
    out = tmp_path / "U.txt"
    ez.write(str(out), 0)
    assert out.read_bytes() == golden["lepl8/U_txt"].tobytes()


def test_dense_api_solves_like_numpy(L):
    rng = np.random.default_rng(3)
    m = 30
    A = rng.normal(size=(m, m))
    A = A @ A.T + m * np.eye(m)
    y = rng.normal(size=m)
    mat = F.bfm_matrix_t()
    assert L.bfm_matrix_full_create(C.byref(mat), C.byref(L.state), 0, m) == 0
    for i in range(m):
        for j in range(m):
            assert L.bfm_matrix_set(C.byref(mat), i, j, A[i, j]) == 0
    vec = F.bfm_vec_t()
    assert L.bfm_vec_create(C.byref(vec), C.byref(L.state), m) == 0
    for i in range(m):
        vec.data[i] = y[i]
    assert L.bfm_matrix_solve(C.byref(mat), C.byref(vec)) == 0
    x = np.ctypeslib.as_array(vec.data, shape=(m,)).copy()
    assert np.linalg.norm(x - np.linalg.solve(A, y)) <= 1e-12 * np.linalg.norm(x)
    assert np.isnan(L.bfm_matrix_get(C.byref(mat), m, 0))
    L.bfm_matrix_destroy(C.byref(mat))
    L.bfm_vec_destroy(C.byref(vec))


def test_dense_rcm_and_band_match_oracle(L):
    """bfm_perm_rcm + renumbering + band LU on a hand-built dense system,
    against the oracle's restatement of the reference (bit-exact)."""
    mesh = meshgen.beam(10, 3, kind=4, jitter=0.2, seed=8)
    kinds, values, flags = meshgen.beam_conditions(mesh)
    u_o, K, b, perm_o, k_o, rc = O.solve_dense(mesh, True, 2e11, 0.3, 0.0, kinds, values, flags)
    n = mesh.n_dof
    sysm = F.bfm_system_t()
    assert L.bfm_system_create(C.byref(sysm), C.byref(L.state), n) == 0
    data = np.ctypeslib.as_array(sysm.A.full.data, shape=(n * n,))
    data[:] = K.reshape(-1)
    bv = np.ctypeslib.as_array(sysm.b.data, shape=(n,))
    bv[:] = b
    assert L.bfm_system_renumber(C.byref(sysm)) == 0
    perm = np.ctypeslib.as_array(sysm.perm.perm, shape=(n,)).copy()
    assert np.array_equal(perm, perm_o)
    assert sysm.A.kind == F.MATRIX_KIND_BAND and sysm.A.band.k == k_o
    assert L.bfm_matrix_solve(C.byref(sysm.A), C.byref(sysm.b)) == 0
    assert L.bfm_perm_perm_vec(C.byref(sysm.perm), C.byref(sysm.b), True) == 0
    u = np.ctypeslib.as_array(sysm.b.data, shape=(n,)).copy()
    assert np.array_equal(u, u_o)
    L.bfm_system_destroy(C.byref(sysm))


def test_band_rejects_out_of_band_values(L):
    mat = F.bfm_matrix_t()
    assert L.bfm_matrix_band_create(C.byref(mat), C.byref(L.state), 0, 10, 2) == 0
    assert L.bfm_matrix_set(C.byref(mat), 0, 5, 1.0) == -1
    assert L.bfm_matrix_set(C.byref(mat), 0, 5, 1e-30) == 0
    assert L.bfm_matrix_get(C.byref(mat), 0, 5) == 0
    assert L.bfm_matrix_bandwidth(C.byref(mat)) == 2
    L.bfm_matrix_destroy(C.byref(mat))


def test_allocator_failure_is_reported(L):
    state = F.bfm_state_t()
    assert L.bfm_state_create(C.byref(state)) == 0
    failing = F.ALLOC(lambda size: None)
    assert L.bfm_set_alloc(C.byref(state), C.cast(failing, C.c_void_p)) == 0
    vec = F.bfm_vec_t()
    assert L.bfm_vec_create(C.byref(vec), C.byref(state), 100) == -1
    assert state.err.has and b"allocate" in state.err.msg


def test_shapes_rules_forces(L):
    rule = pb.Rule_gauss_legendre(2, 4, lib=L)
    r = rule.c_rule
    assert r.n_points == 4 and [r.weights[i] for i in range(4)] == [1.0] * 4
    assert r.points[0][0] == -0.577350269189626
    phi = (C.c_double * 4)()
    pt = (C.c_double * 2)(0.1, -0.3)
    fn = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))(r.shape.phi)
    assert fn(C.addressof(r.shape), pt, phi) == 0
    assert abs(sum(phi) - 1.0) < 1e-15
    with pytest.raises(AssertionError):
        pb.Rule_gauss_legendre(2, 6, lib=L)
    f = pb.Force_linear((0.0, -9.81), lib=L)
    out = F.bfm_vec_t()
    pos = F.bfm_vec_t()
    assert L.bfm_vec_create(C.byref(out), C.byref(L.state), 2) == 0
    assert L.bfm_vec_create(C.byref(pos), C.byref(L.state), 2) == 0
    assert L.bfm_force_eval(C.byref(f.c_force), C.byref(pos), C.byref(out)) == 0
    assert out.data[1] == -9.81


def test_sim_run_without_gpu_fails_loudly(L):
    if D.device_count() > 0:
        pytest.skip("a GPU is visible")
    m = pb.Mesh_lepl1110(LEPL, lib=L)
    ez = pb.Ez_lepl1110(m, PROBLEM, lib=L)
    with pytest.raises(AssertionError):
        ez.sim.run()
    assert b"GPU" in L.state.err.msg
