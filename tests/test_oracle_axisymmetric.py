"""The oracle's axisymmetric formulation (stress == 2 in oracle/bfm_oracle.c)
pinned bit for bit against the reference compiled from its own sources
(fixtures axi_* of tests/golden/make_golden.py): K and b before and after
every condition kind, with a body load, on Q1 and P1 meshes with r = x > 0,
and the displacements of the reference's own sim on an annulus-like beam.

The reference's axisymmetric constants and f_22 (libbfm/src/system.c:241-243,
339) differ from the textbook formulation; the oracle restates them as
written. Two properties follow that the tests also record: K stays symmetric
(to rounding) and is indefinite for nu = 0.3, so conjugate gradients do not
apply to it."""

import numpy as np
import pytest

import oracle as O
from conftest import dense_from_golden, rel_err
from make_golden import TINY_CONDITIONS, node_sets

from bfm_amd import meshgen


def tiny(kind):
    m = meshgen.beam(3, 2, kind=kind, length=3.0, height=2.0, jitter=0.2, seed=11)
    m.coords = m.coords + np.array([1.0, 0.0])
    return m


def conds(mesh, program):
    sets = node_sets(mesh)
    return ([c[0] for c in program], [c[1] for c in program],
            np.stack([sets[c[2]] for c in program]).astype(np.uint8))


@pytest.mark.parametrize("kind,tag", [(4, "q1"), (3, "p1")])
def test_axisymmetric_bit_exact(golden, kind, tag):
    mesh = tiny(kind)
    K0, b0 = O.assemble_dense(mesh, 2, 7.0e10, 0.33, 2700.0, forces=[(0.0, -9.81)])
    assert np.array_equal(K0, golden[f"axi_{tag}/K_pre"])
    assert np.array_equal(b0, golden[f"axi_{tag}/b_pre"])
    kinds, values, flags = conds(mesh, TINY_CONDITIONS)
    u, K, b, perm, k, rc = O.solve_dense(mesh, 2, 7.0e10, 0.33, 2700.0, kinds, values, flags,
                                         forces=[(0.0, -9.81)])
    assert rc == 0
    assert np.array_equal(K, golden[f"axi_{tag}/K"])
    assert np.array_equal(b, golden[f"axi_{tag}/b"])
    assert np.array_equal(u, golden[f"axi_{tag}/u"])


def test_axisymmetric_ring_solution(golden):
    mesh = meshgen.beam(40, 10, kind=4, length=4.0, height=1.0, jitter=0.2, seed=12)
    mesh.coords = mesh.coords + np.array([1.0, 0.0])
    program = [(1, 0.0, "bottom"), (0, 0.0, "left"), (3, -1e6, "top")]
    kinds, values, flags = conds(mesh, program)
    u, K, b, perm, k, rc = O.solve_dense(mesh, 2, 2.0e11, 0.3, 7850.0, kinds, values, flags,
                                         forces=[(0.0, -9.81)])
    assert rc == 0
    assert np.array_equal(K, dense_from_golden(golden, "axi_ring", mesh.n_dof))
    assert np.array_equal(b, golden["axi_ring/b"])
    assert np.array_equal(u, golden["axi_ring/u"])
    # the CSR path assembles the same entries
    A, bc = O.assemble_csr(mesh, 2, 2.0e11, 0.3, 7850.0, forces=[(0.0, -9.81)], kinds=kinds, values=values,
                           flags=flags)
    assert np.array_equal(A.to_scipy().toarray(), K)
    assert np.array_equal(bc, b)


def test_axisymmetric_constants_and_definiteness():
    # the reference multiplies by (1 - 2 nu) where plane strain divides
    a, b, c = O.dconst(2, 1.0, 0.3)
    assert a == pytest.approx(0.7 / 1.3 * 0.4, rel=1e-15)
    assert b == pytest.approx(0.3 / 1.3 / 0.4, rel=1e-15)
    assert c == pytest.approx(1 / 2.6, rel=1e-15)
    a0, b0, c0 = O.dconst(0, 1.0, 0.3)
    assert b == b0 and c == c0 and a != a0
    mesh = tiny(4)
    K, _ = O.assemble_dense(mesh, 2, 7.0e10, 0.33, 0.0)
    assert np.abs(K - K.T).max() <= 1e-15 * np.abs(K).max()
    w = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert w[0] < -1e-3 * w[-1]          # indefinite: a negative eigenvalue of size
    # the textbook constant would make the element matrix positive semi-definite
    assert a < b < a0


def test_axisymmetric_neumann_weighting():
    """Neumann X/Y loads carry r1 + r2 with r formed from one node's x and y;
    normal/tangent loads add nothing."""
    mesh = tiny(4)
    top = node_sets(mesh)["top"].astype(np.uint8)
    m = mesh.n_dof
    K = np.zeros((m, m))
    b = np.zeros(m)
    _, b_y = O.apply_conditions_dense(mesh, K, b, [3], [1.0], top[None], axisymmetric=True)
    _, b_p = O.apply_conditions_dense(mesh, K, b, [3], [1.0], top[None], axisymmetric=False)
    s3 = 1 / np.sqrt(3)
    xy = mesh.coords
    tot = 0.0
    for n1, n2 in mesh.edges.astype(np.int64):
        if top[n1] and top[n2]:
            r1 = xy[n1, 0] * (1 - s3) / 2 + xy[n1, 1] * (1 + s3) / 2
            r2 = xy[n2, 0] * (1 - s3) / 2 + xy[n2, 1] * (1 + s3) / 2
            tot += 2 * (r1 + r2) * np.hypot(*(xy[n1] - xy[n2])) / 2
    assert b_y[1::2].sum() == pytest.approx(tot, rel=1e-14)
    assert b_p[1::2].sum() == pytest.approx(3.0, rel=1e-14)        # the planar total: the edge length
    for kind in (4, 5):
        _, bn = O.apply_conditions_dense(mesh, K, b, [kind], [1.0], top[None], axisymmetric=True)
        assert not bn.any()
