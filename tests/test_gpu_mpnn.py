"""The learned preconditioner on the GPU (bfm_amd/csrc/hip/mpnn.hip) against
its CPU restatement (oracle/mpnn.py): node features, the forward pass
(MFMA BF16 layers), the loss and its gradient from the hand-written backward
kernels; flexible CG with the untrained network (block Jacobi) against
block-Jacobi PCG; training; the gain of the trained network over the method
without it; the block-Jacobi fallback."""

import numpy as np
import pytest

import mpnn as M
from conftest import rel_err

from bfm_amd import device as D
from bfm_amd import meshgen

pytestmark = pytest.mark.gpu


def beam(nx=40, ny=10, seed=0, hetero=False):
    mesh = meshgen.beam(nx, ny, kind=4, jitter=0.2, seed=seed)
    if hetero:
        mesh.elem_E = meshgen.loguniform_E(mesh.n_elems, seed=seed)
    return mesh, D.beam_system(mesh)


def internal_fixed(mesh, s):
    fixed = np.zeros((mesh.n_nodes, 2), bool)
    kinds, values, flags = meshgen.beam_conditions(mesh)
    for k, f in zip(kinds, flags):
        if k in (0, 1):
            fixed[np.asarray(f, bool), k] = True
    out = np.zeros_like(fixed)
    out[s.node_perm()] = fixed
    return out


def model_with_encoder(seed, scale=0.3):
    """A model whose network is active (the GPU's Glorot layers, a random encoder)."""
    m = D.Model(seed=seed)
    p = m.get_params()
    rng = np.random.default_rng(seed)
    p[M.OFF_W0:M.OFF_L] = scale * rng.standard_normal(M.OFF_L - M.OFF_W0)
    m.set_params(p)
    return m, p


@pytest.fixture(scope="module")
def small():
    mesh, s = beam(40, 10, seed=0, hetero=True)
    ts = D.TrainSet()
    ts.add(s, max_pairs=12, every=4, rtol=1e-9)
    return mesh, s, ts


def test_features_and_pairs(small):
    mesh, s, ts = small
    feat, r, e, sv, den = ts.get(0)
    rowptr, col, val = s.get_bsr()
    f_or = M.features(rowptr, col, val, internal_fixed(mesh, s))
    assert np.allclose(feat[:, :6], f_or[:, :6], rtol=1e-6, atol=0)
    assert np.allclose(feat[:, 6], f_or[:, 6], rtol=1e-5, atol=1e-6)
    assert np.array_equal(feat[:, 7], f_or[:, 7])
    # pair 0 is the start of PCG: r = b, e = x*
    b = s.get_rhs().reshape(-1, 2)
    b_int = np.zeros_like(b)
    b_int[s.node_perm()] = b
    assert np.allclose(r[0], b_int.astype(np.float32), rtol=0, atol=0)
    # e_0 = x*, the Jacobi-PCG solution to the same tolerance (same iterates)
    u, _ = s.solve(precond=D.PRECOND_JACOBI, rtol=1e-9)
    u_int = np.zeros_like(b)
    u_int[s.node_perm()] = u.reshape(-1, 2)
    assert np.array_equal(e[0], u_int.astype(np.float32))
    # s = rms of r, den = e^T A e
    assert np.allclose(sv, np.sqrt((r.astype(np.float64) ** 2).sum(axis=(1, 2)) / r[0].size), rtol=1e-5)
    A = M.Graph(rowptr, col, val, feat).A   # FP32 values
    for p in (0, 7):
        ev = e[p].astype(np.float64).reshape(-1)
        assert den[p] == pytest.approx(ev @ (A @ ev), rel=1e-3)


def test_forward_matches_oracle(small):
    mesh, s, ts = small
    feat, r, e, sv, den = ts.get(0)
    m, p = model_with_encoder(3)
    s.set_model(m)
    G = M.Graph(*s.get_bsr(), feat)
    rng = np.random.default_rng(0)
    free = internal_fixed(mesh, s) == 0
    rand = rng.standard_normal((G.n, 2)) * np.abs(r[3]).max()
    rand[~free] = 0.0   # residuals vanish on constrained DOFs
    for rv in (r[3].astype(np.float64), rand):
        z = s.apply_model(rv).reshape(-1, 2)
        z_or = M.apply(p, G, rv, bf=True)
        y = G.block_jacobi(rv)
        corr_err = rel_err((z - y)[free], (z_or - y)[free])
        print(f"forward: correction rel error {corr_err:.2e}, z rel error {rel_err(z, z_or):.2e}")
        assert corr_err <= 2e-2
        assert rel_err(z, z_or) <= 1e-3
    s.set_model(None)


def test_loss_and_gradient_match_oracle(small):
    mesh, s, ts = small
    feat, r, e, sv, den = ts.get(0)
    m, p = model_with_encoder(4, scale=0.5)
    G = M.Graph(*s.get_bsr(), feat)
    p0, B = 2, 4
    loss, g = m.loss_grad(ts, 0, p0, B)
    loss_o, g_o = M.loss_grad(p, G, r[p0:p0 + B], e[p0:p0 + B], den[p0:p0 + B], bf=True)
    print(f"loss gpu {loss:.6e} oracle {loss_o:.6e}")
    assert abs(loss - loss_o) <= 1e-3 * loss_o
    for name, lo, hi in [("W0", M.OFF_W0, M.OFF_L), ("W1", M.OFF_L, M.OFF_L + M.LAYER),
                         ("W2", M.OFF_L + M.LAYER, M.OFF_L + 2 * M.LAYER),
                         ("W3", M.OFF_L + 2 * M.LAYER, M.N_PARAMS)]:
        err = rel_err(g[lo:hi], g_o[lo:hi])
        print(f"grad {name}: rel error {err:.2e}")
        assert err <= 3e-2, name


def test_untrained_network_is_block_jacobi_in_fcg():
    """Zero encoder (a new model): the learned preconditioner is exactly FP64
    block Jacobi, so flexible CG tracks block-Jacobi PCG."""
    mesh, s = beam(60, 15, seed=1)
    m = D.Model(seed=1)
    p = m.get_params()
    p[M.OFF_W0:M.OFF_L] = 0.0
    m.set_params(p)
    s.set_model(m)
    u1, st1 = s.solve(precond=D.PRECOND_LEARNED, rtol=1e-9)
    u2, st2 = s.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-9)
    print(f"fcg(block-jacobi net) {st1['iterations']} vs pcg(block jacobi) {st2['iterations']} iterations")
    assert st1["status"] == 0 and st1["fallback"] == 0
    assert abs(st1["iterations"] - st2["iterations"]) <= 0.05 * st2["iterations"] + 4
    assert rel_err(u1, u2) <= 1e-7


def test_learned_without_model_is_an_error():
    mesh, s = beam(8, 4)
    with pytest.raises(D.BfmxError) as e:
        s.solve(precond=D.PRECOND_LEARNED)
    assert e.value.code == -1


def test_fallback_on_a_broken_network():
    """A network that returns NaN breaks flexible CG down; block-Jacobi PCG
    takes over from the current iterate and the solve still converges."""
    mesh, s = beam(40, 10, seed=2)
    m, p = model_with_encoder(5)
    p[M.OFF_W0 + 3] = np.nan
    m.set_params(p)
    s.set_model(m)
    u, st = s.solve(precond=D.PRECOND_LEARNED, rtol=1e-9)
    u_ref, _ = s.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-9)
    print(f"fallback after breakdown: {st['iterations']} iterations ({st['fallback_iterations']} block Jacobi)")
    assert st["status"] == 0 and st["fallback"] == 1
    assert rel_err(u, u_ref) <= 1e-7


def test_training_lowers_the_loss_and_the_iterations():
    """Trained on pairs from a few small heterogeneous beams, the network
    lowers the held-out loss and, inside flexible CG on a held-out beam, the
    iteration count against block-Jacobi PCG (the method without it)."""
    ts = D.TrainSet()
    keep = []
    for seed in (10, 11, 12):
        mesh, s = beam(60, 15, seed=seed, hetero=True)
        ts.add(s, max_pairs=48, every=6, rtol=1e-10)
        keep.append(s)
    mesh, s_test = beam(60, 15, seed=20, hetero=True)
    ts.add(s_test, max_pairs=16, every=20, rtol=1e-10)
    held = len(ts) - 1
    m = D.Model(seed=0)
    l0, _ = m.loss_grad(ts, held, 0, 8)
    train = D.TrainSet()   # the held-out entry is not trained on
    for s in keep:
        train.add(s, max_pairs=48, every=6, rtol=1e-10)
    losses = m.train(train, steps=300, batch=8, lr=2e-3, seed=1)
    l1, _ = m.loss_grad(ts, held, 0, 8)
    print(f"held-out loss {l0:.4f} -> {l1:.4f}; training loss {losses[:10].mean():.4f} -> {losses[-10:].mean():.4f}")
    # the A-norm error of PCG iterates is dominated by smooth modes that no
    # local operator can fix: the loss moves little, the iterations a lot
    assert l1 < l0
    s_test.set_model(m)
    u1, st1 = s_test.solve(precond=D.PRECOND_LEARNED, rtol=1e-9)
    u2, st2 = s_test.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-9)
    print(f"held-out beam: learned FCG {st1['iterations']} iterations ({st1['solve_ms']:.1f} ms, fallback "
          f"{st1['fallback']}) vs block-Jacobi PCG {st2['iterations']} ({st2['solve_ms']:.1f} ms)")
    assert st1["status"] == 0 and st1["fallback"] == 0
    assert rel_err(u1, u2) <= 1e-7
    assert st1["iterations"] < st2["iterations"]


def test_learned_solve_partial_sums_fit():
    """Regression: the decoder writes one FCG partial per 64 nodes (more than
    the 3 per 256 of the CG kernels); at 10k nodes the solver's partial-sum
    buffer must hold them."""
    mesh, s = beam(200, 50, seed=6)
    m = D.Model(seed=2)
    s.set_model(m)
    u, st = s.solve(precond=D.PRECOND_LEARNED, rtol=1e-8)
    u_ref, _ = s.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-8)
    print(f"10k nodes: learned {st['iterations']} iterations (fallback {st['fallback']})")
    assert st["status"] == 0
    assert rel_err(u, u_ref) <= 1e-6


def test_model_swap_replays_no_stale_graph():
    """Regression: captured learned chunks were looked up by the model's host
    address, so a model created at the address of a destroyed one replayed
    graphs that read the destroyed model's freed parameters. Models A, B, C
    are attached in turn with learned solves (two with A: chunks are captured
    from a system's second solve with a preconditioner on), A is destroyed
    before C exists; C's solve (through captured chunks) must equal C's solve
    on a fresh system (launched directly) bitwise, and its preconditioner the
    oracle's with C's weights."""
    mesh, s = beam(40, 10, seed=7, hetero=True)
    models = {}
    for name, seed, solves in (("A", 11, 2), ("B", 12, 1)):
        models[name], _ = model_with_encoder(seed)
        s.set_model(models[name])
        for _ in range(solves):
            _, st = s.solve(precond=D.PRECOND_LEARNED, rtol=1e-9)
            assert st["status"] == 0
    models["A"].close()
    models["C"], p_c = model_with_encoder(13)
    s.set_model(models["C"])
    u, st = s.solve(precond=D.PRECOND_LEARNED, rtol=1e-9)
    _, fresh = beam(40, 10, seed=7, hetero=True)
    fresh.set_model(models["C"])
    u_fresh, st_fresh = fresh.solve(precond=D.PRECOND_LEARNED, rtol=1e-9)
    print(f"model C: {st['iterations']} iterations, fresh system {st_fresh['iterations']}")
    assert st["status"] == 0 and st["fallback"] == 0
    assert st["iterations"] == st_fresh["iterations"]
    assert np.array_equal(u, u_fresh)
    feat = M.features(*s.get_bsr(), internal_fixed(mesh, s))
    G = M.Graph(*s.get_bsr(), feat)
    rng = np.random.default_rng(3)
    rv = rng.standard_normal((G.n, 2))
    rv[internal_fixed(mesh, s)] = 0.0
    z = s.apply_model(rv).reshape(-1, 2)
    assert rel_err(z, M.apply(p_c, G, rv, bf=True)) <= 1e-3
    s.set_model(None)
    fresh.set_model(None)


# ------------------------------------------------------- multigrid base (PRECOND_AMG_LEARNED)


def test_multigrid_base_with_a_zero_network_is_the_multigrid():
    """z = V(r) + (network term): with a zero encoder the term vanishes, so
    flexible CG follows the multigrid FCG (the same z; only the z . q sums
    are formed in another kernel)."""
    mesh, s = beam(120, 30, seed=21, hetero=True)
    m = D.Model(seed=1)
    p = m.get_params()
    p[M.OFF_W0:M.OFF_L] = 0.0
    m.set_params(p)
    s.set_model(m)
    u1, st1 = s.solve(precond=D.PRECOND_AMG_LEARNED, rtol=1e-9)
    u2, st2 = s.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    print(f"zero network on the multigrid: {st1['iterations']} vs multigrid FCG {st2['iterations']} iterations")
    assert st1["status"] == 0 and st1["fallback"] == 0
    assert abs(st1["iterations"] - st2["iterations"]) <= 2
    assert rel_err(u1, u2) <= 1e-7


def test_multigrid_base_training_and_held_out_solve():
    """Pairs rebased on the multigrid (e -> e - V(r) + D^-1 r): with a zero
    network the loss is then the V-cycle's own A-norm error on the pairs,
    well below block Jacobi's; training lowers it on held-out pairs; the
    trained correction inside flexible CG on a held-out heterogeneous beam
    converges without fallback, in fewer iterations than block-Jacobi PCG
    (the numbers against the multigrid FCG alone are printed)."""
    ts = D.TrainSet()
    keep = []
    for seed in (40, 41, 42):
        mesh, s = beam(100, 25, seed=seed, hetero=True)
        ts.add(s, max_pairs=48, every=20, rtol=1e-10)
        keep.append(s)
    mesh, s_test = beam(100, 25, seed=50, hetero=True)
    ts.add(s_test, max_pairs=16, every=60, rtol=1e-10)
    held = len(ts) - 1
    zero = D.Model(seed=0)
    pz = zero.get_params()
    pz[M.OFF_W0:M.OFF_L] = 0.0
    zero.set_params(pz)
    l_jac, _ = zero.loss_grad(ts, held, 0, 8)
    ts.rebase()
    l_mg, _ = zero.loss_grad(ts, held, 0, 8)
    m = D.Model(seed=3)
    l0, _ = m.loss_grad(ts, held, 0, 8)
    train = D.TrainSet()
    for s in keep:
        train.add(s, max_pairs=48, every=20, rtol=1e-10)
    train.rebase()
    m.train(train, steps=300, batch=8, lr=2e-3, seed=2)
    l1, _ = m.loss_grad(ts, held, 0, 8)
    print(f"held-out loss: block Jacobi {l_jac:.4f}, V-cycle {l_mg:.4f}; network {l0:.4f} -> {l1:.4f}")
    assert l_mg < 0.5 * l_jac
    assert l1 < l0
    s_test.set_model(m)
    u1, st1 = s_test.solve(precond=D.PRECOND_AMG_LEARNED, rtol=1e-9)
    u2, st2 = s_test.solve(precond=D.PRECOND_BLOCK_JACOBI, rtol=1e-9)
    u3, st3 = s_test.solve(precond=D.PRECOND_AMG, rtol=1e-9)
    print(f"held-out beam: multigrid + network {st1['iterations']} iterations ({st1['solve_ms']:.1f} ms, fallback "
          f"{st1['fallback']}), multigrid {st3['iterations']} ({st3['solve_ms']:.1f} ms), block-Jacobi PCG "
          f"{st2['iterations']} ({st2['solve_ms']:.1f} ms)")
    assert st1["status"] == 0 and st1["fallback"] == 0
    assert st1["iterations"] < st2["iterations"]
    assert rel_err(u1, u2) <= 1e-6 and rel_err(u1, u3) <= 1e-6
