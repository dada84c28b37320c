"""Batched solves (bfm_amd/csrc/hip/batch.hip): one persistent launch per size
class solves uneven holed panels; every system's solution matches the
oracle's direct solve or the single-system PCG, does not depend on the rest of
the batch (bitwise), and the autotuner's picks replay bitwise. Batches built
from meshes are assembled by the launch itself: their matrices and loads
match the reference's golden vectors and, bitwise, the single-system path."""

import numpy as np
import pytest

from conftest import rel_err

from bfm_amd import device as D
from bfm_amd import meshgen

pytestmark = pytest.mark.gpu

# panel DOF ranges covering the three size classes (< 2048, < 8192, >= 8192 nodes)
RANGES = [(600, 1500), (1500, 4000), (5000, 12000), (12000, 20000), (20000, 30000)]


@pytest.fixture(scope="module")
def panels():
    out = []
    for k in range(10):
        lo, hi = RANGES[k % len(RANGES)]
        mesh, t = meshgen.panel(100 + k, lo, hi)
        out.append((mesh, t, D.panel_system(mesh, t)))
    return out


def test_batch_matches_oracle_and_solo_pcg(panels):
    import oracle as O
    b = D.Batch([s for _, _, s in panels])
    stats, ms = b.solve(rtol=1e-10)
    print(f"{len(panels)} panels in {ms:.1f} ms; iterations {[st['iterations'] for st in stats]}")
    mat = meshgen.PANEL_MATERIAL
    for k, (mesh, t, s) in enumerate(panels):
        assert stats[k]["status"] == 0 and stats[k]["rel_residual"] <= 1e-10
        u = b.get_u(k)
        if mesh.n_dof <= 1600:
            kinds, values, flags = meshgen.panel_conditions(mesh, t)
            u_ref, *_ = O.solve_dense(mesh, False, mat["E"], mat["nu"], mat["rho"], kinds, values, flags)
            err = rel_err(u, u_ref)
            print(f"panel {k} ({mesh.n_dof} DOFs): error vs oracle {err:.2e}")
            assert err <= 1e-8
        u_solo, st = s.solve(precond=D.PRECOND_JACOBI, rtol=1e-10)
        assert rel_err(u, u_solo) <= 1e-7, k


def test_system_result_does_not_depend_on_the_batch(panels):
    systems = [s for _, _, s in panels]
    full = D.Batch(systems)
    full.solve(rtol=1e-9)
    for k in (0, 3, 9):
        one = D.Batch([systems[k]])
        one.solve(rtol=1e-9)
        assert np.array_equal(full.get_u(k), one.get_u(0)), k


def test_autotune_picks_replay_bitwise(panels):
    systems = [s for _, _, s in panels]
    b = D.Batch(systems)
    picks, times = b.autotune(sample=8, iters=32)
    print(f"picks {picks.tolist()}\ntimes ms\n{np.array2string(times, precision=3)}")
    assert picks.shape == (3,) and np.all((0 <= picks) & (picks < 5))
    assert np.all(times[times > 0] > 0)
    st1, ms1 = b.solve(rtol=1e-9, picks=picks)
    u1 = [b.get_u(k) for k in range(len(systems))]
    st2, ms2 = b.solve(rtol=1e-9, picks=picks)
    assert all(np.array_equal(u1[k], b.get_u(k)) for k in range(len(systems)))
    fresh = D.Batch(systems)
    fresh.solve(rtol=1e-9, picks=picks)
    assert all(np.array_equal(u1[k], fresh.get_u(k)) for k in range(len(systems)))
    print(f"solve with picks: {ms1:.1f} ms, {ms2:.1f} ms")


def test_block_jacobi_batch_and_limits(panels):
    systems = [s for _, _, s in panels]
    b = D.Batch(systems)
    st_j, _ = b.solve(precond=D.PRECOND_JACOBI, rtol=1e-9)
    st_b, _ = b.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-9)
    assert all(s["status"] == 0 for s in st_b)
    print(f"iterations jacobi {sum(s['iterations'] for s in st_j)}, block {sum(s['iterations'] for s in st_b)}")
    st, _ = b.solve(rtol=1e-9, max_iter=5, check=False)
    assert all(s["status"] == -5 and s["iterations"] == 5 for s in st)
    with pytest.raises(D.BfmxError) as e:
        b.solve(rtol=1e-9, max_iter=5)
    assert e.value.code == -5


# ---- batches built from meshes: assembled by the launch itself ----------

def test_planar_batch_against_reference_golden(golden):
    """Tiny Q1 and P1 meshes in plane stress and strain, with every condition
    kind in the reference's order and a body load: the batch-assembled K and
    b against the reference's own (libbfm) vectors, and u against its solve."""
    from make_golden import TINY_CONDITIONS, node_sets
    names, inst = [], []
    for kind, tag in ((4, "q1"), (3, "p1")):
        for stress in (0, 1):
            mesh = meshgen.beam(3, 2, kind=kind, length=3.0, height=2.0, jitter=0.2, seed=11)
            sets = node_sets(mesh)
            flags = np.stack([sets[c[2]] for c in TINY_CONDITIONS]).astype(np.uint8)
            inst.append(D.planar_instance(mesh, stress, dict(E=7.0e10, nu=0.33, rho=2700.0),
                                          [c[0] for c in TINY_CONDITIONS], [c[1] for c in TINY_CONDITIONS],
                                          flags, g=(0.0, -9.81)))
            names.append(f"tiny_{tag}_{'stress' if stress else 'strain'}")
    b = D.Batch.from_instances(inst)
    with pytest.raises(D.BfmxError):
        b.system(0)                      # not assembled before a launch
    st, _ = b.solve(rtol=1e-14)
    for k, name in enumerate(names):
        s = b.system(k)
        K, rhs = s.get_dense(), s.get_rhs()
        scale = np.abs(golden[f"{name}/K"]).max()
        assert np.abs(K - golden[f"{name}/K"]).max() <= 1e-13 * scale, name
        assert np.abs(rhs - golden[f"{name}/b"]).max() <= 1e-13 * np.abs(golden[f"{name}/b"]).max(), name
        assert st[k]["status"] == 0
        assert rel_err(b.get_u(k), golden[f"{name}/u"]) <= 1e-8, name


def test_planar_batch_is_the_solo_path_bitwise(panels):
    """The launch assembles each panel into exactly the matrix and load the
    single-system kernels build from the same mesh; the solve then gives the
    same u as the batch of those systems; a batch-owned system solved on its
    own gives the single system's solo u."""
    inst = [D.panel_instance(m, t) for m, t, _ in panels]
    pb = D.Batch.from_instances(inst)
    st, ms = pb.solve(rtol=1e-10)
    assert all(x["status"] == 0 for x in st)
    sb = D.Batch([s for _, _, s in panels])
    sb.solve(rtol=1e-10)
    for k, (m, t, s) in enumerate(panels):
        own = pb.system(k)
        r1, c1, v1 = own.get_bsr()
        r2, c2, v2 = s.get_bsr()
        assert np.array_equal(r1, r2) and np.array_equal(c1, c2), k
        dv = np.abs(v1 - v2).max() / np.abs(v2).max()
        db = np.abs(own.get_rhs() - s.get_rhs()).max() / np.abs(s.get_rhs()).max()
        assert dv == 0 and db == 0, f"panel {k}: K differs by {dv:.2e}, b by {db:.2e} (relative)"
        assert np.array_equal(pb.get_u(k), sb.get_u(k)), k
    for k in (0, 7):
        u1, s1 = pb.system(k).solve(precond=D.PRECOND_JACOBI, rtol=1e-10)
        u2, s2 = panels[k][2].solve(precond=D.PRECOND_JACOBI, rtol=1e-10)
        assert np.array_equal(u1, u2) and s1["iterations"] == s2["iterations"], k
    with pytest.raises(D.BfmxError):
        pb.system(0).assemble(0, 1e9, 0.3, 0.0)    # batch-owned: not rebuilt on its own


def test_planar_batch_q1_heterogeneous_with_gravity():
    """Q1 beams with per-element E, a body load and the beam's conditions,
    against the single-system path (bitwise K and b; u to the solve tolerance)."""
    inst, systems = [], []
    for k, (nx, ny) in enumerate([(24, 6), (40, 10), (90, 30)]):
        mesh = meshgen.beam(nx, ny, kind=4, jitter=0.2, seed=20 + k)
        eE = meshgen.loguniform_E(mesh.n_elems, seed=20 + k)
        kinds, values, flags = meshgen.beam_conditions(mesh)
        mat = meshgen.BEAM_MATERIAL
        inst.append(D.planar_instance(mesh, True, mat, kinds, values, flags, g=(0.0, -9.81), elem_E=eE))
        s = D.System.from_mesh(mesh)
        s.assemble(1, mat["E"], mat["nu"], mat["rho"], elem_E=eE, g=(0.0, -9.81))
        s.apply_conditions(kinds, values, flags)
        systems.append(s)
    pb = D.Batch.from_instances(inst)
    st, _ = pb.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-9)
    for k, s in enumerate(systems):
        own = pb.system(k)
        assert np.array_equal(own.get_bsr()[2], s.get_bsr()[2]), k
        assert np.array_equal(own.get_rhs(), s.get_rhs()), k
        u, _ = s.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-9)
        assert st[k]["status"] == 0
        assert rel_err(pb.get_u(k), u) <= 1e-6, k
