"""Multigrid mode of the batched solve (bfmx_batch_solve with
BFMX_PRECOND_AMG, bfm_amd/csrc/hip/batch.hip): the batch solved as one
block-diagonal system by flexible CG with the FP32 multigrid V-cycle, every
load scaled to unit norm and the union's tolerance checks testing each
instance's residual. Every instance must meet rtol on its own true residual
(recomputed here on the host
from the matrix the launch assembled), small instances must match the
oracle's direct solve, and a mesh made of several disconnected parts solves
with one dense coarsest inverse per part."""

import numpy as np
import pytest

from conftest import rel_err

from bfm_amd import device as D
from bfm_amd import meshgen

pytestmark = pytest.mark.gpu

RANGES = [(600, 1500), (1500, 4000), (5000, 12000), (12000, 20000)]


def host_rel_residual(system, u):
    K, b = system.get_csr_scipy(), system.get_rhs()
    return np.linalg.norm(b - K @ u) / np.linalg.norm(b)


def test_multigrid_batch_meets_rtol_per_instance():
    import oracle as O
    meshes = []
    for k in range(12):
        lo, hi = RANGES[k % len(RANGES)]
        meshes.append(meshgen.panel(200 + k, lo, hi))
    pb = D.Batch.from_instances([D.panel_instance(m, t) for m, t in meshes])
    st, ms = pb.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    iters = st[0]["iterations"]
    print(f"{len(meshes)} panels, {sum(m.n_dof for m, _ in meshes)} DOFs: {iters} iterations, {ms:.1f} ms")
    # a union without the instances' coordinates (no rotation modes) took
    # ~100 iterations here
    assert iters <= 60
    mat = meshgen.PANEL_MATERIAL
    for k, (mesh, t) in enumerate(meshes):
        assert st[k]["status"] == 0 and st[k]["iterations"] == iters, k
        assert st[k]["precond"] == D.PRECOND_AMG
        u = pb.get_u(k)
        rel = host_rel_residual(pb.system(k), u)
        assert rel <= 1e-9 and st[k]["rel_residual"] <= 1e-9, (k, rel)
        # the two evaluations of b - K u round differently (up to ~2e-13 of ||b||
        # apart for panels that end far below rtol)
        assert abs(rel - st[k]["rel_residual"]) <= 1e-3 * rel + 1e-12, (k, rel, st[k]["rel_residual"])
        if mesh.n_dof <= 1600:
            kinds, values, flags = meshgen.panel_conditions(mesh, t)
            u_ref, *_ = O.solve_dense(mesh, False, mat["E"], mat["nu"], mat["rho"], kinds, values, flags)
            err = rel_err(u, u_ref)
            print(f"panel {k} ({mesh.n_dof} DOFs): error vs oracle {err:.2e}")
            assert err <= 1e-6, k
    # a batch-owned system solved on its own by the multigrid (its
    # coordinates live on the device only)
    u_solo, st_solo = pb.system(11).solve(precond=D.PRECOND_AMG, rtol=1e-9)
    assert st_solo["status"] == 0 and st_solo["iterations"] <= 40
    assert rel_err(u_solo, pb.get_u(11)) <= 1e-4
    # the same launch again: deterministic
    u1 = [pb.get_u(k) for k in range(len(meshes))]
    pb.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    assert all(np.array_equal(u1[k], pb.get_u(k)) for k in range(len(meshes)))
    # and the persistent PCG launch of the same batch agrees to the tolerance
    stp, _ = pb.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-9)
    for k in range(len(meshes)):
        assert stp[k]["status"] == 0
        assert rel_err(pb.get_u(k), u1[k]) <= 1e-4, k


def test_multigrid_batch_of_systems_with_a_zero_load():
    systems = []
    for k in range(4):
        m, t = meshgen.panel(300 + k, 1500, 6000)
        systems.append(D.panel_system(m, t))
    m, _ = meshgen.panel(310, 1500, 3000)
    mat = meshgen.PANEL_MATERIAL
    s0 = D.System.from_mesh(m)
    s0.assemble(0, mat["E"], mat["nu"], mat["rho"])
    pins = m.node_sets["pins"]
    s0.apply_conditions([0, 1], [0.0, 0.0], [pins, pins])   # no load at all
    systems.append(s0)
    b = D.Batch(systems)
    st, ms = b.solve(precond=D.PRECOND_AMG, rtol=1e-8)
    assert all(x["status"] == 0 for x in st)
    assert np.all(b.get_u(4) == 0) and st[4]["b_norm"] == 0
    for k in range(4):
        rel = host_rel_residual(systems[k], b.get_u(k))
        assert rel <= 1e-8, (k, rel)


def test_multigrid_batch_refuses_axisymmetric_instances():
    mesh = meshgen.beam(8, 4, kind=4, length=2.0, height=1.0)
    kinds, values, flags = meshgen.beam_conditions(mesh)
    inst = D.planar_instance(mesh, D.AXISYMMETRIC, meshgen.BEAM_MATERIAL, kinds, values, flags)
    pb = D.Batch.from_instances([inst])
    with pytest.raises(D.BfmxError) as e:
        pb.solve(precond=D.PRECOND_AMG, rtol=1e-8)
    assert e.value.code == -7


def test_disconnected_mesh_one_dense_inverse_per_part():
    """Three panels in one mesh (one system, RCM numbers each part in one
    range): the multigrid solve coarsens every part to at most 64 nodes and
    inverts each densely; it matches block-Jacobi PCG."""
    coords, elems, edges, top, pins, trac, offs = [], [], [], [], [], [], [0]
    for k in range(3):
        m, t = meshgen.panel(400 + k, 800, 3000)
        o = offs[-1]
        coords.append(m.coords + np.array([0.0, 2.0 * k]))
        elems.append(m.elems.astype(np.int64) + o)
        edges.append(m.edges.astype(np.int64) + o)
        top.append(m.node_sets["top"])
        pins.append(m.node_sets["pins"])
        trac.append(t)
        offs.append(o + m.n_nodes)
    N = offs[-1]
    s = D.System(3, np.concatenate(coords), np.concatenate(elems), True, edges=np.concatenate(edges))
    mat = meshgen.PANEL_MATERIAL
    s.assemble(0, mat["E"], mat["nu"], mat["rho"])
    top, pins = np.concatenate(top), np.concatenate(pins)
    for k in range(3):
        f = np.zeros(N, np.uint8)
        f[offs[k]:offs[k + 1]] = top[offs[k]:offs[k + 1]]
        s.apply_condition(3, trac[k], f)
    s.apply_conditions([0, 1], [0.0, 0.0], [pins, pins])
    u, st = s.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    print(f"{2 * N} DOFs in 3 parts: {st['iterations']} iterations")
    assert st["status"] == 0 and st["iterations"] <= 60
    assert host_rel_residual(s, u) <= 1e-9
    u_ref, _ = s.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-10)
    assert rel_err(u, u_ref) <= 1e-5


def test_multigrid_set_up_threads_give_the_same_hierarchy(monkeypatch):
    """The host build of the transfer operators (BFMX_AMG_BUILD=host) runs one
    thread per chunk of whole segments; one thread and sixteen give bitwise
    the same solve. The device build of the same batch (the default, whose
    finest level is swept on host threads the same way) also meets the
    tolerance on every panel, and gives bitwise the same solve with one
    thread and with sixteen."""
    meshes = [meshgen.panel(500 + k, 1000, 4000) for k in range(20)]
    dev = []
    for threads in ("1", "16"):   # the finest level's sweeps run on host threads
        monkeypatch.setenv("BFMX_AMG_THREADS", threads)
        pb = D.Batch.from_instances([D.panel_instance(m, t) for m, t in meshes])
        st, _ = pb.solve(precond=D.PRECOND_AMG, rtol=1e-8)
        assert all(x["status"] == 0 for x in st)
        dev.append((st[0]["iterations"], [pb.get_u(k) for k in range(len(meshes))]))
        pb.close()
    assert dev[0][0] == dev[1][0]
    assert all(np.array_equal(a, b) for a, b in zip(dev[0][1], dev[1][1]))
    monkeypatch.setenv("BFMX_AMG_BUILD", "host")
    out = []
    for threads in ("1", "16"):
        monkeypatch.setenv("BFMX_AMG_THREADS", threads)
        pb = D.Batch.from_instances([D.panel_instance(m, t) for m, t in meshes])
        st, _ = pb.solve(precond=D.PRECOND_AMG, rtol=1e-8)
        assert all(x["status"] == 0 for x in st)
        out.append((st[0]["iterations"], [pb.get_u(k) for k in range(len(meshes))]))
        pb.close()
    assert out[0][0] == out[1][0]
    assert all(np.array_equal(a, b) for a, b in zip(out[0][1], out[1][1]))


def test_segmented_cg_takes_the_iterations_of_the_hardest_instance(monkeypatch):
    """Every instance runs its own CG scalars: the batch needs about the
    iterations of its hardest instance solved alone (shared scalars needed
    more, growing with the batch: 30 at 16 panels, 51 at 4096). Run twice:
    with the default aggregation (host sweeps over whole segments: every
    instance of the union gets the aggregates it gets alone) and
    with MIS-2 aggregates on the GPU at every level (BFMX_AMG_HOST_AGG=0)."""
    for host_agg in (None, "0"):
        if host_agg is not None:
            monkeypatch.setenv("BFMX_AMG_HOST_AGG", host_agg)
        segmented_vs_alone()


def segmented_vs_alone():
    meshes = [meshgen.panel(600 + k, 2000, 12000) for k in range(8)]
    solo = []
    for m, t in meshes:
        b1 = D.Batch.from_instances([D.panel_instance(m, t)])
        st, _ = b1.solve(precond=D.PRECOND_AMG, rtol=1e-8)
        assert st[0]["status"] == 0
        solo.append(st[0]["iterations"])
        b1.close()
    pb = D.Batch.from_instances([D.panel_instance(m, t) for m, t in meshes])
    st, _ = pb.solve(precond=D.PRECOND_AMG, rtol=1e-8)
    print(f"alone {solo}, together {st[0]['iterations']}")
    assert all(x["status"] == 0 and x["rel_residual"] <= 1e-8 for x in st)
    # one eigenvalue estimate for the whole batch and 4-iteration chunks between
    # tolerance checks allow a few more
    assert st[0]["iterations"] <= max(solo) + 4


def test_small_instance_falls_back_to_shared_scalars():
    """An instance of fewer than 256 nodes: the batch shares one pair of CG
    scalars, and every instance still meets rtol on its own residual."""
    meshes = [meshgen.panel(700, 200, 400)] + [meshgen.panel(701 + k, 2000, 6000) for k in range(3)]
    assert meshes[0][0].n_nodes < 256
    pb = D.Batch.from_instances([D.panel_instance(m, t) for m, t in meshes])
    st, _ = pb.solve(precond=D.PRECOND_AMG, rtol=1e-8)
    for k, (m, t) in enumerate(meshes):
        assert st[k]["status"] == 0
        assert host_rel_residual(pb.system(k), pb.get_u(k)) <= 1e-8, k


def test_multigrid_batch_mixes_element_kinds_and_materials():
    """Q1 beams with per-element E and a body load beside P1 panels, one
    multigrid launch: each instance against its own single-system solve."""
    inst, systems = [], []
    for k, (nx, ny) in enumerate([(48, 12), (96, 24)]):
        mesh = meshgen.beam(nx, ny, kind=4, jitter=0.2, seed=40 + k)
        eE = meshgen.loguniform_E(mesh.n_elems, lo=5e10, hi=2e11, seed=40 + k)
        kinds, values, flags = meshgen.beam_conditions(mesh)
        mat = meshgen.BEAM_MATERIAL
        inst.append(D.planar_instance(mesh, True, mat, kinds, values, flags, g=(0.0, -9.81), elem_E=eE))
        s = D.System.from_mesh(mesh)
        s.assemble(1, mat["E"], mat["nu"], mat["rho"], elem_E=eE, g=(0.0, -9.81))
        s.apply_conditions(kinds, values, flags)
        systems.append(s)
    for k in range(2):
        m, t = meshgen.panel(800 + k, 3000, 9000)
        inst.append(D.panel_instance(m, t))
        systems.append(D.panel_system(m, t))
    pb = D.Batch.from_instances(inst)
    st, _ = pb.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    for k, s in enumerate(systems):
        assert st[k]["status"] == 0 and st[k]["rel_residual"] <= 1e-9, k
        u_ref, _ = s.solve(precond=D.PRECOND_AMG, rtol=1e-10)
        assert rel_err(pb.get_u(k), u_ref) <= 1e-5, k
