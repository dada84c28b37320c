"""The axisymmetric formulation on the GPU (BFMX_AXISYMMETRIC, replacing
fill_axisymmetric_elem and bfm_system_create_axisymmetric_strain,
libbfm/src/system.c:228-356, 539-622, formulas kept as the reference writes
them): K and b against the reference's own vectors for every condition kind
with a body load, on Q1 and P1 meshes; u from MINRES (the reference's
axisymmetric K is indefinite) against the reference's band-LU solution,
through the device API, through the drop-in libbfm sim (pybfm-style), and in
a batch built from meshes."""

import numpy as np
import pytest

from conftest import dense_from_golden, rel_err
from make_golden import TINY_CONDITIONS, node_sets

from bfm_amd import device as D
from bfm_amd import meshgen
from bfm_amd import pybfm as pb

pytestmark = pytest.mark.gpu

K_TOL = 1e-13


def tiny(kind):
    m = meshgen.beam(3, 2, kind=kind, length=3.0, height=2.0, jitter=0.2, seed=11)
    m.coords = m.coords + np.array([1.0, 0.0])
    return m


def ring():
    m = meshgen.beam(40, 10, kind=4, length=4.0, height=1.0, jitter=0.2, seed=12)
    m.coords = m.coords + np.array([1.0, 0.0])
    return m


RING_CONDS = [(1, 0.0, "bottom"), (0, 0.0, "left"), (3, -1e6, "top")]


def program(mesh, prog):
    sets = node_sets(mesh)
    return ([c[0] for c in prog], [c[1] for c in prog], np.stack([sets[c[2]] for c in prog]).astype(np.uint8))


def close(a, ref, tol=K_TOL):
    assert np.abs(a - ref).max() <= tol * np.abs(ref).max()


@pytest.mark.parametrize("kind,tag", [(4, "q1"), (3, "p1")])
def test_axisymmetric_system_against_reference(golden, kind, tag):
    mesh = tiny(kind)
    s = D.System.from_mesh(mesh)
    s.assemble(D.AXISYMMETRIC, 7.0e10, 0.33, 2700.0, g=(0.0, -9.81))
    close(s.get_dense(), golden[f"axi_{tag}/K_pre"])
    close(s.get_rhs(), golden[f"axi_{tag}/b_pre"])
    s.assemble(D.AXISYMMETRIC, 7.0e10, 0.33, 2700.0, g=(0.0, -9.81))
    s.apply_conditions(*program(mesh, TINY_CONDITIONS))
    close(s.get_dense(), golden[f"axi_{tag}/K"])
    close(s.get_rhs(), golden[f"axi_{tag}/b"])
    u, st = s.solve(precond=D.PRECOND_MINRES, rtol=1e-13)
    print(f"axi {tag}: MINRES {st['iterations']} iterations, residual {st['rel_residual']:.2e}, "
          f"error vs libbfm {rel_err(u, golden[f'axi_{tag}/u']):.2e}")
    assert st["status"] == 0
    assert rel_err(u, golden[f"axi_{tag}/u"]) <= 1e-10


def test_axisymmetric_ring_minres(golden):
    mesh = ring()
    s = D.System.from_mesh(mesh)
    s.assemble(D.AXISYMMETRIC, 2.0e11, 0.3, 7850.0, g=(0.0, -9.81))
    s.apply_conditions(*program(mesh, RING_CONDS))
    K = s.get_dense()
    close(K, dense_from_golden(golden, "axi_ring", mesh.n_dof))
    close(s.get_rhs(), golden["axi_ring/b"])
    u, st = s.solve(precond=D.PRECOND_MINRES, rtol=1e-10)
    err = rel_err(u, golden["axi_ring/u"])
    print(f"axi ring: MINRES {st['iterations']} iterations, residual {st['rel_residual']:.2e}, error vs libbfm {err:.2e}")
    assert st["status"] == 0 and st["rel_residual"] <= 1e-10
    assert err <= 1e-6


def test_axisymmetric_sim_through_the_dropin(golden):
    L = pb.Library()
    mesh = ring()
    cm = pb.Mesh_arrays.from_generated(mesh, lib=L)
    mat = pb.Material("steel", 7850.0, 2.0e11, 0.3, lib=L)
    rule = pb.Rule_gauss_legendre(2, mesh.kind, lib=L)
    obj = pb.Obj(cm, mat, rule, lib=L)
    inst = pb.Instance(obj, lib=L)
    sets = node_sets(mesh)
    conds = []
    for kind, value, where in RING_CONDS:
        c = pb.Condition(cm, kind, value, lib=L)
        c.set_nodes(sets[where])
        inst.add_condition(c)
        conds.append(c)
    sim = pb.Sim(pb.Sim.AXISYMMETRIC_STRAIN, lib=L)
    sim.add_instance(inst)
    f = pb.Force_linear((0.0, -9.81), lib=L)
    sim.add_force(f)
    sim.run()
    err = rel_err(inst.effects, golden["axi_ring/u"])
    print(f"axisymmetric sim: error vs libbfm {err:.2e}")
    assert err <= 1e-6


def test_axisymmetric_batch_instance_matches_the_system():
    mesh = ring()
    kinds, values, flags = program(mesh, RING_CONDS)
    mat = dict(E=2.0e11, nu=0.3, rho=7850.0)
    b = D.Batch.from_instances([D.planar_instance(mesh, D.AXISYMMETRIC, mat, kinds, values, flags, g=(0.0, -9.81))])
    b.solve(rtol=1e-3, max_iter=1, check=False)      # the launch assembles; PCG itself does not apply
    s = D.System.from_mesh(mesh)
    s.assemble(D.AXISYMMETRIC, mat["E"], mat["nu"], mat["rho"], g=(0.0, -9.81))
    s.apply_conditions(kinds, values, flags)
    own = b.system(0)
    assert np.array_equal(own.get_bsr()[2], s.get_bsr()[2])
    assert np.array_equal(own.get_rhs(), s.get_rhs())
    u1, st1 = own.solve(precond=D.PRECOND_MINRES, rtol=1e-9)
    u2, st2 = s.solve(precond=D.PRECOND_MINRES, rtol=1e-9)
    assert np.array_equal(u1, u2)
