"""Seeded synthetic meshes for the planar-elasticity workloads.

Beam: the rectangle [0, L] x [0, H] on an nx x ny grid, nodes numbered
x-fastest (node(i, j) = j * (nx + 1) + i).
  * kind 4 (Q1): one quad per cell, nodes listed counter-clockwise from the
    top-right corner, matching the Q1 shape functions' node order
    (libbfm/src/shape.c:19-26 in the reference tree).
  * kind 3 (P1): each cell split along its (i, j)-(i+1, j+1) diagonal into
    [(i,j), (i+1,j), (i+1,j+1)] and [(i,j), (i+1,j+1), (i,j+1)].
Jitter moves interior nodes by U(-f h, +f h) per axis and boundary nodes only
along their edge, so the boundary stays straight; corners never move.
Edges are the boundary edges, counter-clockwise, as a .lepl1110 mesh lists
them (elems[1] = -1).
"""

from dataclasses import dataclass, field

import numpy as np


@dataclass
class Mesh:
    kind: int                      # 3 = P1 triangles, 4 = Q1 quads
    coords: np.ndarray             # (n_nodes, 2) float64
    elems: np.ndarray              # (n_elems, kind) uint64
    edges: np.ndarray              # (n_edges, 2) uint64, node pairs
    edge_elem: np.ndarray          # (n_edges,) int64, adjacent element
    node_sets: dict = field(default_factory=dict)   # name -> bool (n_nodes,)
    elem_E: np.ndarray = None      # optional per-element Young's modulus

    @property
    def n_nodes(self):
        return self.coords.shape[0]

    @property
    def n_elems(self):
        return self.elems.shape[0]

    @property
    def n_dof(self):
        return 2 * self.coords.shape[0]


def beam(nx, ny, kind=4, length=4.0, height=1.0, jitter=0.0, seed=0):
    if kind not in (3, 4):
        raise ValueError("kind must be 3 (P1) or 4 (Q1)")
    hx, hy = length / nx, height / ny
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), indexing="xy")
    ii, jj = ii.ravel(), jj.ravel()
    x = ii * hx
    y = jj * hy
    if jitter > 0:
        rng = np.random.default_rng(seed)
        dx = rng.uniform(-jitter * hx, jitter * hx, size=x.shape)
        dy = rng.uniform(-jitter * hy, jitter * hy, size=y.shape)
        on_x = (ii == 0) | (ii == nx)   # left/right edges: move along y only
        on_y = (jj == 0) | (jj == ny)   # bottom/top edges: move along x only
        dx[on_x] = 0.0
        dy[on_y] = 0.0
        x = x + dx
        y = y + dy
    coords = np.stack([x, y], axis=1).astype(np.float64)

    def node(i, j):
        return j * (nx + 1) + i

    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="xy")
    ci, cj = ci.ravel(), cj.ravel()
    if kind == 4:
        elems = np.stack([node(ci + 1, cj + 1), node(ci, cj + 1), node(ci, cj), node(ci + 1, cj)], axis=1)
    else:
        t0 = np.stack([node(ci, cj), node(ci + 1, cj), node(ci + 1, cj + 1)], axis=1)
        t1 = np.stack([node(ci, cj), node(ci + 1, cj + 1), node(ci, cj + 1)], axis=1)
        elems = np.empty((2 * ci.size, 3), dtype=np.int64)
        elems[0::2] = t0
        elems[1::2] = t1
    elems = elems.astype(np.uint64)

    def cell(i, j):
        return j * nx + i

    def tri_of(c, second):
        return 2 * c + second if kind == 3 else c

    ex = np.arange(nx)
    ey = np.arange(ny)
    bottom = np.stack([node(ex, 0), node(ex + 1, 0)], 1)
    bottom_e = tri_of(cell(ex, 0), 0)
    right = np.stack([node(nx, ey), node(nx, ey + 1)], 1)
    right_e = tri_of(cell(nx - 1, ey), 0)
    top = np.stack([node(ex[::-1] + 1, ny), node(ex[::-1], ny)], 1)
    top_e = tri_of(cell(ex[::-1], ny - 1), 1)
    left = np.stack([node(0, ey[::-1] + 1), node(0, ey[::-1])], 1)
    left_e = tri_of(cell(0, ey[::-1]), 1)
    edges = np.concatenate([bottom, right, top, left]).astype(np.uint64)
    edge_elem = np.concatenate([bottom_e, right_e, top_e, left_e]).astype(np.int64)

    sets = {
        "left": ii == 0,
        "right": ii == nx,
        "bottom": jj == 0,
        "top": jj == ny,
    }
    return Mesh(kind, coords, elems, edges, edge_elem, sets)


def loguniform_E(n_elems, lo=1e9, hi=1e11, seed=0):
    """Per-element Young's modulus drawn log-uniformly from [lo, hi]."""
    rng = np.random.default_rng(seed + 7919)
    return 10.0 ** rng.uniform(np.log10(lo), np.log10(hi), size=n_elems)


# The beam problem of the planar workloads: E = 211 GPa, nu = 0.3,
# rho = 7850, left edge clamped, uniform traction -1e6 N/m on the top edge.
# The Neumann condition is listed before the two Dirichlet ones, so the
# clamped top-left corner keeps its prescribed zero (a Neumann listed after
# a Dirichlet would add its load to the constrained DOF).
BEAM_MATERIAL = dict(E=211e9, nu=0.3, rho=7850.0)
BEAM_TRACTION = -1e6


def beam_conditions(mesh, traction=BEAM_TRACTION):
    """(kinds, values, flags) of the beam problem; flags is (n_cond, n_nodes) uint8."""
    kinds = np.array([3, 0, 1], dtype=np.int32)           # Neumann-Y, Dirichlet-X, Dirichlet-Y
    values = np.array([traction, 0.0, 0.0])
    flags = np.stack([mesh.node_sets["top"], mesh.node_sets["left"], mesh.node_sets["left"]]).astype(np.uint8)
    return kinds, values, flags


# ---------------------------------------------------------------- holed panels

PANEL_MATERIAL = dict(E=211e9, nu=0.3, rho=7850.0)


def panel(seed, dofs_lo=1000, dofs_hi=60000, height=1.0):
    """A bridge-like truss panel of the batched workload (P1, plane strain).

    The rectangle [0, W] x [0, H] (W / H uniform in [3, 6]) on a structured
    grid of about U(dofs_lo, dofs_hi) DOFs, each cell split along its
    (i, j)-(i+1, j+1) diagonal, with 0-12 seeded circular holes of radius
    2-8% of H: triangles whose centroid lies in a hole are removed, nodes no
    longer used are dropped. Holes keep two cells of material to every side
    and to each other (a drawn hole that would not is drawn again, at most 50
    times), so the panel stays one piece, every kept triangle shares an edge
    with another, and the top edge and the bottom corners are kept.
    Returns (mesh, traction): traction is U(1e5, 1e7) N/m, downward.
    Node sets: "top" (the top edge), "pins" (the two bottom corners)."""
    rng = np.random.default_rng(1_000_003 + seed)
    H = height
    W = H * rng.uniform(3.0, 6.0)
    dofs = rng.uniform(dofs_lo, dofs_hi)
    # (nx + 1)(ny + 1) nodes, nx / ny = W / H, before the holes remove a few
    ny = max(2, int(round(np.sqrt(dofs / 2.0 / (W / H)))))
    nx = max(2, int(round(ny * W / H)))
    gap = 2.0 * max(W / nx, H / ny)
    holes = []
    for _ in range(int(rng.integers(0, 13))):
        for _try in range(50):
            r = H * rng.uniform(0.02, 0.08)
            m = r + gap
            if 2 * m >= H:
                continue
            hx, hy = rng.uniform(m, W - m), rng.uniform(m, H - m)
            if all((hx - x) ** 2 + (hy - y) ** 2 > (r + q + gap) ** 2 for x, y, q in holes):
                holes.append((hx, hy, r))
                break
    ii, jj = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1))
    coords = np.stack([ii.ravel() * W / nx, jj.ravel() * H / ny], 1)

    def node(i, j):
        return j * (nx + 1) + i

    ci, cj = np.meshgrid(np.arange(nx), np.arange(ny))
    ci, cj = ci.ravel(), cj.ravel()
    t1 = np.stack([node(ci, cj), node(ci + 1, cj), node(ci + 1, cj + 1)], 1)
    t2 = np.stack([node(ci, cj), node(ci + 1, cj + 1), node(ci, cj + 1)], 1)
    tris = np.concatenate([t1, t2])
    cen = coords[tris].mean(axis=1)
    keep = np.ones(len(tris), bool)
    for (hx, hy, r) in holes:
        keep &= (cen[:, 0] - hx) ** 2 + (cen[:, 1] - hy) ** 2 > r * r
    tris = tris[keep]
    used = np.zeros(len(coords), bool)
    used[tris.ravel()] = True
    new = -np.ones(len(coords), np.int64)
    new[used] = np.arange(used.sum())
    top_old = np.array([node(i, ny) for i in range(nx, -1, -1)])   # right to left: counter-clockwise
    assert used[top_old].all() and used[node(0, 0)] and used[node(nx, 0)]
    coords = coords[used]
    tris = new[tris]
    top = new[top_old]
    edges = np.stack([top[:-1], top[1:]], 1).astype(np.uint64)
    # the element of each top edge: the upper-left triangle t2 of cell (i, ny - 1)
    sets = {"top": np.zeros(len(coords), bool), "pins": np.zeros(len(coords), bool)}
    sets["top"][top] = True
    sets["pins"][new[node(0, 0)]] = True
    sets["pins"][new[node(nx, 0)]] = True
    mesh = Mesh(3, coords, tris.astype(np.uint64), edges, -np.ones(len(edges), np.int64), sets)
    traction = -rng.uniform(1e5, 1e7)
    return mesh, traction


def panel_conditions(mesh, traction):
    """(kinds, values, flags): traction on the top edge, bottom corners pinned."""
    kinds = np.array([3, 0, 1], dtype=np.int32)
    values = np.array([traction, 0.0, 0.0])
    flags = np.stack([mesh.node_sets["top"], mesh.node_sets["pins"], mesh.node_sets["pins"]]).astype(np.uint8)
    return kinds, values, flags


def workload(name, seed=0):
    """Named workloads: (mesh, stress flag).

    c1        40 x 10 P1 beam, 902 DOFs, plane stress (CPU-runnable case)
    q1_502k   1000 x 250 Q1 beam, 20% jitter, 502,502 DOFs, plane stress
    q1_2m     2000 x 500 Q1 beam, 20% jitter, 2,005,002 DOFs, plane stress
    q1_2m_het q1_2m with E log-uniform in [1e9, 1e11] per element
    q1_WxH    any W x H Q1 beam with 20% jitter (e.g. q1_64x16)
    """
    if name == "c1":
        return beam(40, 10, kind=3), True
    if name == "q1_502k":
        return beam(1000, 250, kind=4, jitter=0.2, seed=seed), True
    if name == "q1_2m":
        return beam(2000, 500, kind=4, jitter=0.2, seed=seed), True
    if name == "q1_2m_het":
        m = beam(2000, 500, kind=4, jitter=0.2, seed=seed)
        m.elem_E = loguniform_E(m.n_elems, seed=seed)
        return m, True
    if name.startswith("q1_") and "x" in name:
        w, h = name[3:].split("x")
        return beam(int(w), int(h), kind=4, jitter=0.2, seed=seed), True
    if name.startswith("p1_") and "x" in name:
        w, h = name[3:].split("x")
        return beam(int(w), int(h), kind=3, jitter=0.2, seed=seed), True
    raise KeyError(name)
