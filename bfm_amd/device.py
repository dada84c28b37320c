"""ctypes wrapper of libbfmhip.so, the GPU planar-elasticity path (include/bfmx.h).

There is no CPU fallback: when the library is missing or no GPU is visible,
constructing a System raises.
"""

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libbfmhip.so")

PRECOND_JACOBI, PRECOND_BLOCK_JACOBI, PRECOND_AMG, PRECOND_LEARNED, PRECOND_MINRES, PRECOND_AMG_LEARNED = 0, 1, 2, 3, 4, 5
PLANE_STRAIN, PLANE_STRESS, AXISYMMETRIC = 0, 1, 2
FP64, MIXED = 0, 1
KERNEL_CG_SPMV64, KERNEL_CG_SPMV32, KERNEL_FCG_SPMV, KERNEL_AMG_SMOOTH = 0, 1, 2, 3
DIRICHLET_X, DIRICHLET_Y, NEUMANN_X, NEUMANN_Y = 0, 1, 2, 3
NEUMANN_NORMAL, NEUMANN_TANGENT, DIRICHLET_NORMAL, DIRICHLET_TANGENT = 4, 5, 6, 7

ERRORS = {-1: "BFMX_E_ARG", -2: "BFMX_E_HIP", -3: "BFMX_E_NOMEM", -4: "BFMX_E_BREAKDOWN",
          -5: "BFMX_E_MAXITER", -6: "BFMX_E_NODEV", -7: "BFMX_E_UNSUPPORTED", -8: "BFMX_E_STAGNATION"}

# every entry point of include/bfmx.h
EXPORTS = [
    "bfmx_last_error", "bfmx_device_count", "bfmx_system_create", "bfmx_system_destroy",
    "bfmx_system_set_edges", "bfmx_system_assemble", "bfmx_system_apply_condition",
    "bfmx_system_solve", "bfmx_system_solve_noalloc", "bfmx_system_info", "bfmx_system_node_perm", "bfmx_system_get_bsr",
    "bfmx_system_get_rhs", "bfmx_system_get_u", "bfmx_system_get_entry", "bfmx_system_time_spmv",
    "bfmx_system_time_kernel", "bfmx_model_n_params", "bfmx_model_create", "bfmx_model_destroy",
    "bfmx_model_get_params", "bfmx_model_set_params", "bfmx_trainset_create", "bfmx_trainset_destroy",
    "bfmx_trainset_add", "bfmx_trainset_rebase", "bfmx_trainset_info", "bfmx_trainset_get", "bfmx_model_loss_grad", "bfmx_model_train",
    "bfmx_system_set_model", "bfmx_system_apply_model", "bfmx_batch_create", "bfmx_batch_destroy",
    "bfmx_batch_autotune", "bfmx_batch_solve", "bfmx_batch_get_u", "bfmx_batch_create_planar",
    "bfmx_batch_system",
]


class Stats(C.Structure):
    _fields_ = [("status", C.c_int32), ("precond", C.c_int32), ("precision", C.c_int32),
                ("restarts", C.c_int32), ("iterations", C.c_int64), ("n_dof", C.c_int64),
                ("rtol", C.c_double), ("rel_residual", C.c_double), ("solve_ms", C.c_double),
                ("b_norm", C.c_double), ("fallback", C.c_int32), ("fallback_iterations", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Instance(C.Structure):
    """bfmx_instance_t: one planar problem of a batch built from meshes."""
    _fields_ = [("kind", C.c_int32), ("stress", C.c_int32), ("n_nodes", C.c_size_t), ("coords", C.c_void_p),
                ("n_elems", C.c_size_t), ("elems", C.c_void_p), ("n_edges", C.c_size_t),
                ("edge_nodes", C.c_void_p), ("edge_boundary", C.c_void_p), ("E", C.c_double), ("nu", C.c_double),
                ("rho", C.c_double), ("elem_E", C.c_void_p), ("gx", C.c_double), ("gy", C.c_double),
                ("n_conditions", C.c_size_t), ("cond_kind", C.c_void_p), ("cond_value", C.c_void_p),
                ("cond_flags", C.c_void_p)]


class BfmxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64 = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class _Opt:
    def __init__(self, base):
        self.base = base

    def from_param(self, obj):
        return None if obj is None else self.base.from_param(obj)


def lib():
    """Loads libbfmhip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        sz, vp = C.c_size_t, C.c_void_p
        L.bfmx_last_error.restype = C.c_char_p
        L.bfmx_device_count.argtypes = [C.POINTER(C.c_int)]
        L.bfmx_system_create.argtypes = [C.POINTER(vp), C.c_int, sz, _dp, sz, _up, C.c_int]
        L.bfmx_system_destroy.argtypes = [vp]
        L.bfmx_system_set_edges.argtypes = [vp, sz, _up, _Opt(_u8)]
        L.bfmx_system_assemble.argtypes = [vp, C.c_int, C.c_double, C.c_double, C.c_double, _Opt(_dp), _Opt(_dp),
                                           C.c_double, C.c_double]
        L.bfmx_system_apply_condition.argtypes = [vp, C.c_int, C.c_double, _u8]
        L.bfmx_system_solve.argtypes = [vp, _Opt(_dp), _Opt(_dp), C.c_int, C.c_int, C.c_double, sz,
                                        C.POINTER(Stats)]
        L.bfmx_system_solve_noalloc.argtypes = [vp, _Opt(_dp), _Opt(_dp), C.c_int, C.c_double, sz, C.POINTER(Stats)]
        L.bfmx_system_info.argtypes = [vp, C.POINTER(sz), C.POINTER(sz), C.POINTER(sz), C.POINTER(sz)]
        L.bfmx_system_node_perm.argtypes = [vp, _i64]
        L.bfmx_system_get_bsr.argtypes = [vp, _Opt(_i64), _Opt(_i32), _Opt(_dp)]
        L.bfmx_system_get_rhs.argtypes = [vp, _dp]
        L.bfmx_system_get_u.argtypes = [vp, _dp]
        L.bfmx_system_get_entry.argtypes = [vp, sz, sz, C.POINTER(C.c_double)]
        L.bfmx_system_time_spmv.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.bfmx_system_time_kernel.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
        L.bfmx_model_n_params.argtypes = [C.POINTER(sz)]
        L.bfmx_model_create.argtypes = [C.POINTER(vp), C.c_uint64]
        L.bfmx_model_destroy.argtypes = [vp]
        L.bfmx_model_get_params.argtypes = [vp, f32p]
        L.bfmx_model_set_params.argtypes = [vp, f32p]
        L.bfmx_trainset_create.argtypes = [C.POINTER(vp)]
        L.bfmx_trainset_destroy.argtypes = [vp]
        L.bfmx_trainset_add.argtypes = [vp, vp, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_int)]
        L.bfmx_trainset_rebase.argtypes = [vp]
        L.bfmx_trainset_info.argtypes = [vp, C.c_int, C.POINTER(sz), C.POINTER(C.c_int)]
        L.bfmx_trainset_get.argtypes = [vp, C.c_int, _Opt(f32p), _Opt(f32p), _Opt(f32p), _Opt(f32p)]
        L.bfmx_model_loss_grad.argtypes = [vp, vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double), _Opt(f32p)]
        L.bfmx_model_train.argtypes = [vp, vp, C.c_int, C.c_int, C.c_float, C.c_uint64, _Opt(_dp)]
        L.bfmx_system_set_model.argtypes = [vp, vp]
        L.bfmx_system_apply_model.argtypes = [vp, _dp, _dp]
        L.bfmx_batch_create.argtypes = [C.POINTER(vp), sz, C.POINTER(vp)]
        L.bfmx_batch_destroy.argtypes = [vp]
        L.bfmx_batch_autotune.argtypes = [vp, C.c_int, C.c_int, C.c_int, _i32, _Opt(_dp)]
        L.bfmx_batch_solve.argtypes = [vp, C.c_int, C.c_double, sz, _Opt(_i32), C.c_void_p,
                                       C.POINTER(C.c_double)]
        L.bfmx_batch_get_u.argtypes = [vp, sz, _dp]
        L.bfmx_batch_create_planar.argtypes = [C.POINTER(vp), sz, C.POINTER(Instance), C.c_int]
        L.bfmx_batch_system.argtypes = [vp, sz, C.POINTER(vp)]
        for name in EXPORTS[1:]:
            getattr(L, name).restype = C.c_int
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise BfmxError(rc, lib().bfmx_last_error().decode())


def device_count():
    n = C.c_int()
    _check(lib().bfmx_device_count(C.byref(n)))
    return n.value


class System:
    """One planar system on the GPU: mesh upload + renumbering at creation."""

    def __init__(self, kind, coords, elems, renumber=True, edges=None, edge_boundary=None):
        L = lib()
        coords = np.ascontiguousarray(np.asarray(coords, np.float64).reshape(-1))
        elems = np.ascontiguousarray(np.asarray(elems, np.uint64).reshape(-1))
        self.n_nodes = coords.size // 2
        self.h = C.c_void_p()
        _check(L.bfmx_system_create(C.byref(self.h), kind, self.n_nodes, coords, elems.size // kind, elems,
                                    int(bool(renumber))))
        if edges is not None:
            self.set_edges(edges, edge_boundary)

    @classmethod
    def from_mesh(cls, mesh, renumber=True):
        return cls(mesh.kind, mesh.coords, mesh.elems, renumber, edges=mesh.edges)

    def close(self):
        if self.h:
            lib().bfmx_system_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_edges(self, edges, boundary=None):
        e = np.ascontiguousarray(np.asarray(edges, np.uint64).reshape(-1))
        b = None if boundary is None else np.ascontiguousarray(np.asarray(boundary, np.uint8))
        _check(lib().bfmx_system_set_edges(self.h, e.size // 2, e, b))

    def assemble(self, stress, E, nu, rho, elem_E=None, node_force=None, g=(0.0, 0.0)):
        eE = None if elem_E is None else np.ascontiguousarray(elem_E, np.float64)
        nf = None if node_force is None else np.ascontiguousarray(np.asarray(node_force, np.float64).reshape(-1))
        _check(lib().bfmx_system_assemble(self.h, int(stress), E, nu, rho, eE, nf, g[0], g[1]))

    def apply_condition(self, kind, value, flags):
        _check(lib().bfmx_system_apply_condition(self.h, int(kind), float(value),
                                                 np.ascontiguousarray(np.asarray(flags).astype(np.uint8))))

    def apply_conditions(self, kinds, values, flags):
        for k, v, f in zip(kinds, values, flags):
            self.apply_condition(k, v, f)

    def solve(self, b=None, precond=PRECOND_JACOBI, precision=FP64, rtol=1e-8, max_iter=0, want_u=True):
        u = np.empty(2 * self.n_nodes) if want_u else None
        st = Stats()
        bb = None if b is None else np.ascontiguousarray(b, np.float64)
        rc = lib().bfmx_system_solve(self.h, bb, u, precond, precision, rtol, max_iter, C.byref(st))
        _check(rc)
        return u, st.as_dict()

    def solve_noalloc(self, b=None, precond=PRECOND_JACOBI, rtol=1e-8, max_iter=0):
        """bfmx_system_solve_noalloc: point / block Jacobi PCG in the library's
        fixed scratch array (no device allocation)."""
        u = np.empty(2 * self.n_nodes)
        st = Stats()
        bb = None if b is None else np.ascontiguousarray(b, np.float64)
        _check(lib().bfmx_system_solve_noalloc(self.h, bb, u, precond, rtol, max_iter, C.byref(st)))
        return u, st.as_dict()

    def info(self):
        v = [C.c_size_t() for _ in range(4)]
        _check(lib().bfmx_system_info(self.h, *[C.byref(x) for x in v]))
        return dict(zip(["n_nodes", "nnzb", "nnzb_padded", "n_colors"], [x.value for x in v]))

    def node_perm(self):
        p = np.empty(self.n_nodes, np.int64)
        _check(lib().bfmx_system_node_perm(self.h, p))
        return p

    def get_bsr(self):
        """(rowptr, col, val (nnzb, 2, 2)) in the internal node order."""
        info = self.info()
        rowptr = np.empty(self.n_nodes + 1, np.int64)
        col = np.empty(info["nnzb"], np.int32)
        val = np.empty(4 * info["nnzb"])
        _check(lib().bfmx_system_get_bsr(self.h, rowptr, col, val))
        return rowptr, col, val.reshape(-1, 2, 2)

    def get_dense(self):
        """The condition-applied matrix as dense (small meshes), original DOF order."""
        rowptr, col, val = self.get_bsr()
        perm = self.node_perm()
        inv = np.empty_like(perm)
        inv[perm] = np.arange(perm.size)
        m = 2 * self.n_nodes
        K = np.zeros((m, m))
        for i in range(self.n_nodes):
            oi = inv[i]
            for t in range(rowptr[i], rowptr[i + 1]):
                oj = inv[col[t]]
                K[2 * oi:2 * oi + 2, 2 * oj:2 * oj + 2] = val[t]
        return K

    def get_csr_scipy(self):
        """The condition-applied matrix as scipy CSR in original DOF order."""
        import scipy.sparse as sp
        rowptr, col, val = self.get_bsr()
        perm = self.node_perm()
        inv = np.empty_like(perm)
        inv[perm] = np.arange(perm.size)
        counts = np.diff(rowptr)
        rows_new = np.repeat(np.arange(self.n_nodes), counts)
        bi = inv[rows_new]
        bj = inv[col]
        r = np.concatenate([2 * bi, 2 * bi, 2 * bi + 1, 2 * bi + 1])
        c = np.concatenate([2 * bj, 2 * bj + 1, 2 * bj, 2 * bj + 1])
        v = np.concatenate([val[:, 0, 0], val[:, 0, 1], val[:, 1, 0], val[:, 1, 1]])
        m = 2 * self.n_nodes
        return sp.csr_matrix((v, (r, c)), shape=(m, m))

    def get_rhs(self):
        b = np.empty(2 * self.n_nodes)
        _check(lib().bfmx_system_get_rhs(self.h, b))
        return b

    def get_u(self):
        u = np.empty(2 * self.n_nodes)
        _check(lib().bfmx_system_get_u(self.h, u))
        return u

    def get_entry(self, i, j):
        v = C.c_double()
        _check(lib().bfmx_system_get_entry(self.h, i, j, C.byref(v)))
        return v.value

    def time_spmv(self, precision=FP64, reps=50):
        ms, by = C.c_double(), C.c_double()
        _check(lib().bfmx_system_time_spmv(self.h, precision, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def set_model(self, model):
        _check(lib().bfmx_system_set_model(self.h, None if model is None else model.h))
        self._model = model   # keep it alive while attached

    def apply_model(self, r):
        """The learned preconditioner applied once, internal node order (2 per node)."""
        r = np.ascontiguousarray(r, np.float64).reshape(-1)
        z = np.empty_like(r)
        _check(lib().bfmx_system_apply_model(self.h, r, z))
        return z

    def time_kernel(self, kernel, reps=50):
        """(mean ms per launch, algorithmic bytes per launch) of one solver kernel."""
        ms, by = C.c_double(), C.c_double()
        _check(lib().bfmx_system_time_kernel(self.h, kernel, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value


class Model:
    """The learned preconditioner's parameters on the GPU (include/bfmx.h)."""

    def __init__(self, seed=0, params=None):
        self.h = C.c_void_p()
        _check(lib().bfmx_model_create(C.byref(self.h), int(seed)))
        if params is not None:
            self.set_params(params)

    @staticmethod
    def n_params():
        n = C.c_size_t()
        _check(lib().bfmx_model_n_params(C.byref(n)))
        return n.value

    def get_params(self):
        p = np.empty(self.n_params(), np.float32)
        _check(lib().bfmx_model_get_params(self.h, p))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params()
        _check(lib().bfmx_model_set_params(self.h, p))

    def save(self, path):
        self.get_params().tofile(path)

    @classmethod
    def load(cls, path):
        return cls(params=np.fromfile(path, dtype=np.float32))

    def loss_grad(self, ts, entry, p0, batch):
        loss = C.c_double()
        g = np.empty(self.n_params(), np.float32)
        _check(lib().bfmx_model_loss_grad(self.h, ts.h, entry, p0, batch, C.byref(loss), g))
        return loss.value, g

    def train(self, ts, steps, batch=8, lr=1e-3, seed=0):
        losses = np.zeros(steps)
        _check(lib().bfmx_model_train(self.h, ts.h, steps, batch, lr, seed, losses))
        return losses

    def close(self):
        if self.h:
            lib().bfmx_model_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TrainSet:
    """(residual, error) pairs recorded from plain PCG runs (include/bfmx.h)."""

    def __init__(self):
        self.h = C.c_void_p()
        _check(lib().bfmx_trainset_create(C.byref(self.h)))
        self.systems = []   # the systems must outlive the set

    def add(self, system, max_pairs=64, every=8, rtol=1e-10):
        n = C.c_int()
        _check(lib().bfmx_trainset_add(self.h, system.h, max_pairs, every, rtol, C.byref(n)))
        self.systems.append(system)
        return n.value

    def rebase(self):
        """bfmx_trainset_rebase: the pairs scored for the multigrid-based
        preconditioner (PRECOND_AMG_LEARNED); once, after the last add."""
        _check(lib().bfmx_trainset_rebase(self.h))

    def __len__(self):
        return len(self.systems)

    def info(self, entry):
        n, p = C.c_size_t(), C.c_int()
        _check(lib().bfmx_trainset_info(self.h, entry, C.byref(n), C.byref(p)))
        return n.value, p.value

    def get(self, entry):
        """(feat (n, 8), r (P, n, 2), e (P, n, 2), s (P,), den (P,)) in the internal node order."""
        n, P = self.info(entry)
        feat = np.empty(8 * n, np.float32)
        r = np.empty(2 * n * P, np.float32)
        e = np.empty(2 * n * P, np.float32)
        sd = np.empty(2 * P, np.float32)
        _check(lib().bfmx_trainset_get(self.h, entry, feat, r, e, sd))
        return feat.reshape(n, 8), r.reshape(P, n, 2), e.reshape(P, n, 2), sd[:P], sd[P:]

    def close(self):
        if self.h:
            lib().bfmx_trainset_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_CLASSES, BATCH_CONFIGS = 3, 5


def planar_instance(mesh, stress, material, kinds=(), values=(), flags=(), g=(0.0, 0.0), elem_E=None):
    """The arrays of one problem of a batch built from meshes (Batch.from_instances);
    stress: PLANE_STRAIN (or False), PLANE_STRESS (or True) or AXISYMMETRIC."""
    n = mesh.n_nodes
    kinds = np.ascontiguousarray(kinds, np.int32)
    return {
        "kind": int(mesh.kind), "stress": int(stress), "n_nodes": n,
        "coords": np.ascontiguousarray(np.asarray(mesh.coords, np.float64).reshape(-1)),
        "elems": np.ascontiguousarray(np.asarray(mesh.elems, np.uint64).reshape(-1)),
        "edges": np.ascontiguousarray(np.asarray(mesh.edges, np.uint64).reshape(-1)),
        "E": float(material["E"]), "nu": float(material["nu"]), "rho": float(material["rho"]),
        "elem_E": None if elem_E is None else np.ascontiguousarray(elem_E, np.float64),
        "g": (float(g[0]), float(g[1])), "kinds": kinds,
        "values": np.ascontiguousarray(values, np.float64),
        "flags": np.ascontiguousarray(np.asarray(flags, np.uint8).reshape(len(kinds), n)),
    }


def batch_bytes(nnzb, n, iterations, checks, precond=PRECOND_JACOBI):
    """Algorithmic HBM bytes of the batched PCG of one system: per iteration
    the SpMV pass (36 B per stored block, 4 B row length per row, the
    gathered z and previous direction read once, new direction and A p
    written) and the update pass (x, p, r, A p and the Jacobi inverse read;
    x, r, z written); per true-residual check one residual pass."""
    dinv = 16 if precond == PRECOND_JACOBI else 32
    it = 36 * nnzb + 4 * n + 64 * n + (64 + dinv + 48) * n
    check = 36 * nnzb + 4 * n + 16 * n + 16 * n + dinv * n + 32 * n
    return iterations * it + (checks + 1) * check


def panel_instance(mesh, traction, material=None):
    """A holed panel of the batched workload as a batch instance (as panel_system)."""
    from . import meshgen
    kinds, values, flags = meshgen.panel_conditions(mesh, traction)
    return planar_instance(mesh, False, meshgen.PANEL_MATERIAL if material is None else material, kinds, values,
                           flags)


class _BatchSystem(System):
    """A system owned by a batch built from meshes (bfmx_batch_system): the
    getters and solo solves of System; never destroyed on its own."""

    def __init__(self, batch, h, n_nodes):
        self.batch = batch   # keeps the batch alive
        self.h = h
        self.n_nodes = n_nodes

    def close(self):
        self.h = C.c_void_p()


class Batch:
    """Systems solved together by one persistent launch per size class (include/bfmx.h)."""

    def __init__(self, systems):
        self.systems = list(systems)   # kept alive with the batch
        self.sizes = [s.n_nodes for s in self.systems]
        arr = (C.c_void_p * len(self.systems))(*[s.h.value for s in self.systems])
        self.h = C.c_void_p()
        _check(lib().bfmx_batch_create(C.byref(self.h), len(self.systems), arr))

    @classmethod
    def from_instances(cls, instances, renumber=True):
        """A batch built from meshes (dicts of planar_instance): every solve
        launch assembles and solves them."""
        self = cls.__new__(cls)
        self.systems = []
        self.sizes = [d["n_nodes"] for d in instances]
        recs = (Instance * len(instances))()
        ptr = lambda a: None if a is None or a.size == 0 else a.ctypes.data
        for r, d in zip(recs, instances):
            r.kind, r.stress, r.n_nodes = d["kind"], d["stress"], d["n_nodes"]
            r.coords, r.n_elems, r.elems = ptr(d["coords"]), d["elems"].size // d["kind"], ptr(d["elems"])
            r.n_edges, r.edge_nodes, r.edge_boundary = d["edges"].size // 2, ptr(d["edges"]), None
            r.E, r.nu, r.rho, r.elem_E = d["E"], d["nu"], d["rho"], ptr(d["elem_E"])
            r.gx, r.gy = d["g"]
            r.n_conditions = d["kinds"].size
            r.cond_kind, r.cond_value, r.cond_flags = ptr(d["kinds"]), ptr(d["values"]), ptr(d["flags"])
        self.h = C.c_void_p()
        _check(lib().bfmx_batch_create_planar(C.byref(self.h), len(instances), recs, int(bool(renumber))))
        return self

    def system(self, k):
        """System k: the batch's own for a batch built from meshes (available
        after a solve), else the System the batch was made of."""
        if self.systems:
            return self.systems[k]
        h = C.c_void_p()
        _check(lib().bfmx_batch_system(self.h, k, C.byref(h)))
        return _BatchSystem(self, h, self.sizes[k])

    def autotune(self, precond=PRECOND_JACOBI, sample=0, iters=32):
        """(picks (3,), times_ms (3, 5)): the fastest configuration of each size
        class, timed on `sample` systems of the class (0: all)."""
        picks = np.zeros(BATCH_CLASSES, np.int32)
        times = np.zeros(BATCH_CLASSES * BATCH_CONFIGS)
        _check(lib().bfmx_batch_autotune(self.h, precond, sample, iters, picks, times))
        return picks, times.reshape(BATCH_CLASSES, BATCH_CONFIGS)

    def solve(self, precond=PRECOND_JACOBI, rtol=1e-8, max_iter=0, picks=None, check=True):
        """(list of stats dicts, device ms of the launch)."""
        st = (Stats * len(self.sizes))()
        ms = C.c_double()
        p = None if picks is None else np.ascontiguousarray(picks, np.int32)
        rc = lib().bfmx_batch_solve(self.h, precond, rtol, max_iter, p, C.cast(st, C.c_void_p), C.byref(ms))
        stats = [x.as_dict() for x in st]
        if check:
            _check(rc)
        return stats, ms.value

    def get_u(self, k):
        u = np.empty(self.sizes[k] * 2)
        _check(lib().bfmx_batch_get_u(self.h, k, u))
        return u

    def close(self):
        if self.h:
            lib().bfmx_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def panel_system(mesh, traction, material=None):
    """A holed panel of the batched workload: plane strain, traction on top,
    bottom corners pinned."""
    from . import meshgen
    mat = dict(meshgen.PANEL_MATERIAL if material is None else material)
    s = System.from_mesh(mesh)
    s.assemble(0, mat["E"], mat["nu"], mat["rho"])
    kinds, values, flags = meshgen.panel_conditions(mesh, traction)
    s.apply_conditions(kinds, values, flags)
    return s


def beam_system(mesh, stress=True, material=None, traction=None, gravity=None, renumber=True):
    """Builds, assembles and applies the beam problem's conditions."""
    from . import meshgen
    mat = dict(meshgen.BEAM_MATERIAL if material is None else material)
    s = System.from_mesh(mesh, renumber=renumber)
    s.assemble(stress, mat["E"], mat["nu"], mat["rho"], elem_E=mesh.elem_E,
               g=gravity if gravity is not None else (0.0, 0.0))
    kinds, values, flags = meshgen.beam_conditions(mesh, meshgen.BEAM_TRACTION if traction is None else traction)
    s.apply_conditions(kinds, values, flags)
    return s
