"""The batched workload's holed panels (bfm_amd.meshgen.panel): valid,
seeded, within the DOF range, holes really cut, boundary conditions in place."""

import numpy as np

from bfm_amd import meshgen


def tri_area(m):
    c = m.coords[m.elems.astype(np.int64)]
    return 0.5 * ((c[:, 1, 0] - c[:, 0, 0]) * (c[:, 2, 1] - c[:, 0, 1])
                  - (c[:, 2, 0] - c[:, 0, 0]) * (c[:, 1, 1] - c[:, 0, 1]))


def test_panels_are_valid_and_seeded():
    sizes = []
    for seed in range(12):
        m, t = meshgen.panel(seed)
        m2, t2 = meshgen.panel(seed)
        assert np.array_equal(m.coords, m2.coords) and np.array_equal(m.elems, m2.elems) and t == t2
        assert m.kind == 3
        a = tri_area(m)
        assert np.all(a > 0)                              # counter-clockwise, none degenerate
        used = np.zeros(m.n_nodes, bool)
        used[m.elems.astype(np.int64).ravel()] = True
        assert used.all()                                 # no isolated node
        assert 1e5 <= -t <= 1e7
        assert m.node_sets["pins"].sum() == 2
        top = m.node_sets["top"]
        assert np.allclose(m.coords[top, 1], m.coords[:, 1].max())
        assert len(m.edges) == top.sum() - 1
        sizes.append(m.n_dof)
    assert 900 <= min(sizes) and max(sizes) <= 61000


def test_holes_remove_area():
    # a panel with holes has less area than its rectangle, by about the holes' area
    for seed in range(30):
        m, _ = meshgen.panel(seed, 3000, 6000)
        W, H = m.coords[:, 0].max(), m.coords[:, 1].max()
        missing = W * H - tri_area(m).sum()
        assert missing >= -1e-9 * W * H
        if missing > 1e-3 * W * H:
            break
    else:
        raise AssertionError("no panel with holes among 30 seeds")
