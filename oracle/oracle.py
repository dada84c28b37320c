"""ctypes wrapper of the CPU oracle (liboracle.so built from bfm_oracle.c).

TEST INFRASTRUCTURE ONLY: loaded by tests, the smoke check and the
benchmark's CPU-baseline leg, never by the bfm_amd library. See
bfm_oracle.c for which reference function each entry restates.
"""

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libbfm_ref.so")

_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class _Opt:
    """ndpointer that also accepts None (NULL)."""

    def __init__(self, base):
        self.base = base

    def from_param(self, obj):
        if obj is None:
            return None
        return self.base.from_param(obj)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built (run make -C oracle)")
        L = C.CDLL(LIB_PATH)
        sz = C.c_size_t
        L.orc_dconst.argtypes = [C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_assemble_dense.restype = C.c_int
        L.orc_assemble_dense.argtypes = [C.c_int, sz, _dp, sz, _up, C.c_int, C.c_double, C.c_double,
                                         C.c_double, _Opt(_dp), sz, _Opt(_dp), _dp, _dp]
        L.orc_apply_conditions_dense.restype = C.c_int
        L.orc_apply_conditions_dense.argtypes = [sz, _dp, sz, _up, _u8, sz, _i32, _dp, _u8, _dp, _dp]
        L.orc_rcm.restype = C.c_int
        L.orc_rcm.argtypes = [sz, _dp, _up, _up]
        L.orc_bandwidth.restype = sz
        L.orc_bandwidth.argtypes = [sz, _dp]
        L.orc_band_from_dense.restype = C.c_int
        L.orc_band_from_dense.argtypes = [sz, sz, _dp, _dp]
        L.orc_band_lu.restype = C.c_int
        L.orc_band_lu.argtypes = [sz, sz, _dp]
        L.orc_band_lu_solve.restype = C.c_int
        L.orc_band_lu_solve.argtypes = [sz, sz, _dp, _dp]
        L.orc_solve_dense.restype = C.c_int
        L.orc_solve_dense.argtypes = [C.c_int, sz, _dp, sz, _up, C.c_int, C.c_double, C.c_double,
                                      C.c_double, _Opt(_dp), sz, _Opt(_dp), sz, _up, _u8, sz, _i32, _dp,
                                      _u8, _dp, _Opt(_dp), _Opt(_dp), _Opt(_up), C.POINTER(sz)]
        L.orc_csr_pattern.restype = C.c_int64
        L.orc_csr_pattern.argtypes = [C.c_int, sz, sz, _up, _Opt(_i64), _Opt(_i32)]
        L.orc_csr_assemble.restype = C.c_int
        L.orc_csr_assemble.argtypes = [C.c_int, sz, _dp, sz, _up, C.c_int, C.c_double, C.c_double,
                                       C.c_double, _Opt(_dp), sz, _Opt(_dp), _i64, _i32, _dp, _dp]
        L.orc_csr_apply_conditions.restype = C.c_int
        L.orc_csr_apply_conditions.argtypes = [sz, _dp, sz, _up, _u8, sz, _i32, _dp, _u8, C.c_int, _i64, _i32,
                                               _dp, _dp]
        L.orc_apply_conditions_dense_axi.restype = C.c_int
        L.orc_apply_conditions_dense_axi.argtypes = [sz, _dp, sz, _up, _u8, sz, _i32, _dp, _u8, _dp, _dp]
        L.orc_csr_spmv.argtypes = [sz, _i64, _i32, _dp, _dp, _dp]
        L.orc_csr_pcg.restype = C.c_int64
        L.orc_csr_pcg.argtypes = [sz, _i64, _i32, _dp, _dp, _dp, C.c_double, C.c_int64,
                                  C.POINTER(C.c_double)]
        L.orc_rcm_csr.restype = C.c_int
        L.orc_rcm_csr.argtypes = [sz, _i64, _i32, _dp, _up, _up]
        L.orc_bandwidth_csr.restype = sz
        L.orc_bandwidth_csr.argtypes = [sz, _i64, _i32, _dp, _up]
        L.orc_band_window_csr.restype = C.c_int
        L.orc_band_window_csr.argtypes = [sz, _i64, _i32, _dp, _up, _up, sz, sz, sz, _dp]
        L.orc_band_lu_steps.restype = C.c_int
        L.orc_band_lu_steps.argtypes = [sz, sz, _dp, sz]
        L.orc_num_threads.restype = C.c_int
        L.orc_set_num_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def _forces(forces):
    if forces is None or len(forces) == 0:
        return 0, None
    f = np.ascontiguousarray(np.asarray(forces, dtype=np.float64).reshape(-1))
    return f.size // 2, f


def _edges(mesh):
    en = np.ascontiguousarray(mesh.edges.astype(np.uint64).reshape(-1))
    eb = np.ones(mesh.edges.shape[0], dtype=np.uint8)
    return mesh.edges.shape[0], en, eb


def dconst(stress, E, nu):
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    lib().orc_dconst(int(stress), E, nu, C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def assemble_dense(mesh, stress, E, nu, rho, forces=None):
    m = mesh.n_dof
    K = np.zeros(m * m)
    b = np.zeros(m)
    nf, f = _forces(forces)
    rc = lib().orc_assemble_dense(mesh.kind, mesh.n_nodes, np.ascontiguousarray(mesh.coords.reshape(-1)),
                                  mesh.n_elems, np.ascontiguousarray(mesh.elems.reshape(-1)), int(stress), E, nu,
                                  rho, mesh.elem_E, nf, f, K, b)
    assert rc == 0
    return K.reshape(m, m), b


def apply_conditions_dense(mesh, K, b, kinds, values, flags, edge_boundary=None, axisymmetric=False):
    ne, en, eb = _edges(mesh)
    if edge_boundary is not None:
        eb = np.ascontiguousarray(edge_boundary.astype(np.uint8))
    K = np.ascontiguousarray(K.reshape(-1).copy())
    b = b.copy()
    fn = lib().orc_apply_conditions_dense_axi if axisymmetric else lib().orc_apply_conditions_dense
    rc = fn(mesh.n_nodes, np.ascontiguousarray(mesh.coords.reshape(-1)), ne, en, eb,
            len(kinds), np.asarray(kinds, np.int32), np.asarray(values, np.float64),
            np.ascontiguousarray(np.asarray(flags, np.uint8).reshape(-1)), K, b)
    assert rc == 0
    return K.reshape(b.size, b.size), b


def solve_dense(mesh, stress, E, nu, rho, kinds, values, flags, forces=None, edge_boundary=None):
    """run_elasticity restated: returns (u, K_bc, b_bc, perm, k, rc). stress: 0
    plane strain, 1 plane stress, 2 axisymmetric (the reference's formulas)."""
    m = mesh.n_dof
    u = np.zeros(m)
    K = np.zeros(m * m)
    b = np.zeros(m)
    perm = np.zeros(m, dtype=np.uint64)
    k = C.c_size_t()
    nf, f = _forces(forces)
    ne, en, eb = _edges(mesh)
    if edge_boundary is not None:
        eb = np.ascontiguousarray(edge_boundary.astype(np.uint8))
    rc = lib().orc_solve_dense(mesh.kind, mesh.n_nodes, np.ascontiguousarray(mesh.coords.reshape(-1)), mesh.n_elems,
                               np.ascontiguousarray(mesh.elems.reshape(-1)), int(stress), E, nu, rho, mesh.elem_E,
                               nf, f, ne, en, eb, len(kinds), np.asarray(kinds, np.int32),
                               np.asarray(values, np.float64),
                               np.ascontiguousarray(np.asarray(flags, np.uint8).reshape(-1)), u, K, b, perm,
                               C.byref(k))
    return u, K.reshape(m, m), b, perm, k.value, rc


class CSR:
    def __init__(self, rowptr, col, val):
        self.rowptr, self.col, self.val = rowptr, col, val

    @property
    def m(self):
        return self.rowptr.size - 1

    def matvec(self, x):
        y = np.empty(self.m)
        lib().orc_csr_spmv(self.m, self.rowptr, self.col, self.val, np.ascontiguousarray(x, np.float64), y)
        return y

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.m, self.m))


def assemble_csr(mesh, stress, E, nu, rho, forces=None, kinds=None, values=None, flags=None):
    """Sparse assembly (+ ordered conditions when kinds is given): (CSR, b)."""
    L = lib()
    el = np.ascontiguousarray(mesh.elems.reshape(-1))
    nnz = L.orc_csr_pattern(mesh.kind, mesh.n_nodes, mesh.n_elems, el, None, None)
    assert nnz > 0
    rowptr = np.zeros(mesh.n_dof + 1, dtype=np.int64)
    col = np.zeros(nnz, dtype=np.int32)
    L.orc_csr_pattern(mesh.kind, mesh.n_nodes, mesh.n_elems, el, rowptr, col)
    val = np.zeros(nnz)
    b = np.zeros(mesh.n_dof)
    nf, f = _forces(forces)
    xy = np.ascontiguousarray(mesh.coords.reshape(-1))
    rc = L.orc_csr_assemble(mesh.kind, mesh.n_nodes, xy, mesh.n_elems, el, int(stress), E, nu, rho, mesh.elem_E, nf,
                            f, rowptr, col, val, b)
    assert rc == 0
    if kinds is not None:
        ne, en, eb = _edges(mesh)
        rc = L.orc_csr_apply_conditions(mesh.n_nodes, xy, ne, en, eb, len(kinds), np.asarray(kinds, np.int32),
                                        np.asarray(values, np.float64),
                                        np.ascontiguousarray(np.asarray(flags, np.uint8).reshape(-1)),
                                        int(stress == 2), rowptr, col, val, b)
        assert rc == 0
    return CSR(rowptr, col, val), b


def pcg_csr(A, b, rtol=1e-8, max_iter=None, threads=None):
    """Jacobi-PCG from zero: (x, iterations, recursive relative residual)."""
    L = lib()
    if threads is not None:
        L.orc_set_num_threads(int(threads))
    x = np.zeros(A.m)
    rr = C.c_double()
    it = L.orc_csr_pcg(A.m, A.rowptr, A.col, A.val, np.ascontiguousarray(b, np.float64), x, rtol,
                       int(max_iter if max_iter else 50 * A.m), C.byref(rr))
    return x, int(it), rr.value


def band_solve(A_dense, y):
    """Band LU of an already banded-ordered dense matrix, as libbfm runs it."""
    L = lib()
    m = A_dense.shape[0]
    Af = np.ascontiguousarray(A_dense.reshape(-1), np.float64)
    k = L.orc_bandwidth(m, Af)
    band = np.zeros(m * (2 * k + 1))
    assert L.orc_band_from_dense(m, k, Af, band) == 0
    yy = np.ascontiguousarray(y, np.float64).copy()
    rc = L.orc_band_lu(m, k, band)
    rc2 = L.orc_band_lu_solve(m, k, band, yy) if rc == 0 else -1
    return yy, rc, rc2


def rcm(A_dense):
    m = A_dense.shape[0]
    perm = np.zeros(m, dtype=np.uint64)
    inv = np.zeros(m, dtype=np.uint64)
    assert lib().orc_rcm(m, np.ascontiguousarray(A_dense.reshape(-1), np.float64), perm, inv) == 0
    return perm, inv


def rcm_csr(A):
    """RCM of a CSR matrix by the reference's rules (orc_rcm_csr): perm, inv."""
    perm = np.zeros(A.m, dtype=np.uint64)
    inv = np.zeros(A.m, dtype=np.uint64)
    assert lib().orc_rcm_csr(A.m, A.rowptr, A.col, A.val, perm, inv) == 0
    return perm, inv


def bandwidth_csr(A, perm):
    return int(lib().orc_bandwidth_csr(A.m, A.rowptr, A.col, A.val, perm))


def band_window(A, perm, inv, k, r0, nrows):
    """Rows r0 .. r0 + nrows - 1 of P A P^T in band storage (window-local)."""
    band = np.zeros(nrows * (2 * k + 1))
    assert lib().orc_band_window_csr(A.m, A.rowptr, A.col, A.val, perm, inv, k, r0, nrows, band) == 0
    return band


def band_lu_steps(band, m, k, steps):
    return int(lib().orc_band_lu_steps(m, k, band, steps))


def band_lu_solve(band, m, k, y):
    return int(lib().orc_band_lu_solve(m, k, band, y))
