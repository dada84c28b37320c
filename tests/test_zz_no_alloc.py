"""bfmx_system_solve_noalloc (include/bfmx.h), called through the C ABI: the
solve that allocates nothing, its buffers carved from one fixed device array
per GPU. Two calls in a row, the second on a smaller system (the array keeps
the first call's contents), each against bfmx_system_solve with the same
preconditioner and tolerance (the same kernels in the same order: equal to
rounding, 1e-12 relative, and the same iteration count), with the assembled
load and with a given right-hand side; a system larger than the array is
refused with BFMX_E_ARG."""

import ctypes as C

import numpy as np
import pytest

from bfm_amd import device as D
from bfm_amd import meshgen

pytestmark = pytest.mark.gpu

MAX_NODES = 408880   # the 64 MiB array's capacity (the header's stated limit)


def noalloc(s, b=None, precond=D.PRECOND_JACOBI, rtol=1e-9):
    L = D.lib()
    u = np.empty(2 * s.n_nodes)
    st = D.Stats()
    bb = None if b is None else np.ascontiguousarray(b, np.float64)
    rc = L.bfmx_system_solve_noalloc(s.h, bb, u, precond, rtol, 0, C.byref(st))
    return rc, u, st.as_dict()


def close(u, ref):
    return np.abs(u - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("precond", [D.PRECOND_JACOBI, D.PRECOND_BLOCK_JACOBI])
def test_noalloc_matches_the_allocating_solve(precond):
    big = meshgen.beam(240, 60, kind=4, jitter=0.2, seed=31)
    big.elem_E = meshgen.loguniform_E(big.n_elems, seed=31)
    small = meshgen.beam(60, 15, kind=3, jitter=0.2, seed=32)
    rng = np.random.default_rng(33)
    for mesh in (big, small):           # the second call on a smaller system
        s = D.beam_system(mesh)
        rc, u, st = noalloc(s, precond=precond)
        u_ref, st_ref = s.solve(precond=precond, rtol=1e-9)
        print(f"{mesh.n_dof} DOFs: no-alloc {st['iterations']} iterations, residual {st['rel_residual']:.2e}; "
              f"solve {st_ref['iterations']}, max |diff| {np.abs(u - u_ref).max():.2e}")
        assert rc == 0 and st["status"] == 0
        assert st["iterations"] == st_ref["iterations"]
        assert close(u, u_ref)
        # a given right-hand side (original DOF order)
        b = rng.standard_normal(mesh.n_dof) * np.abs(s.get_rhs()).max()
        fixed = np.zeros((mesh.n_nodes, 2), bool)
        fixed[mesh.node_sets["left"]] = True
        b[fixed.reshape(-1)] = 0.0
        rc, u2, st2 = noalloc(s, b=b, precond=precond)
        u2_ref, st2_ref = s.solve(b=b, precond=precond, rtol=1e-9)
        assert rc == 0 and st2["iterations"] == st2_ref["iterations"]
        assert close(u2, u2_ref)
        s.close()


def test_noalloc_refuses_a_system_beyond_its_array():
    mesh = meshgen.beam(1000, 410, kind=4)
    assert mesh.n_nodes > MAX_NODES
    s = D.beam_system(mesh)
    rc, _, _ = noalloc(s)
    assert rc == -1
    msg = D.lib().bfmx_last_error().decode()
    print(msg)
    assert f"at most {MAX_NODES} nodes" in msg
    # and an allowed size right after still works
    small = D.beam_system(meshgen.beam(40, 10, kind=3))
    rc, u, st = noalloc(small, rtol=1e-10)
    u_ref, _ = small.solve(rtol=1e-10)
    assert rc == 0 and close(u, u_ref)
