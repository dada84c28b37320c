"""Writes the golden fixtures of tests/golden/ from the reference libbfm.

Run in a container that holds the reference tree:
    make -C oracle ref          # compiles the reference C sources into oracle/_ref
    python tests/golden/make_golden.py

The reference is driven through the same ctypes ABI declarations
(bfm_amd/_ffi.py) as this implementation's libbfm.so. Outputs:
    ref_data/               data files copied from the reference tree (mesh,
                            problem, expected U/V, one .obj mesh)
    golden.npz              arrays named <case>/<quantity>
Every case records the full dense K and b only when it has at most 64 DOFs;
larger cases record K as (row, col, value) triplets of its non-zeros.
"""

import ctypes as C
import os
import shutil
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from bfm_amd import _ffi as F  # noqa: E402
from bfm_amd import meshgen  # noqa: E402
from bfm_amd import pybfm as pb  # noqa: E402

REF_TREE = "/root/reference"
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libbfm_ref.so")
DATA_FILES = ["meshes/8.lepl1110", "problems/problem.txt", "data/U.txt", "data/V.txt",
              "meshes/bridge-nothing.obj"]

# Condition program of the tiny cases: every kind once, a Neumann after a
# Dirichlet on a shared node, a DOF constrained twice with different values.
TINY_CONDITIONS = [
    (3, -1e6, "top"),                 # Neumann-Y
    (0, 1e-4, "left"),                # Dirichlet-X (non-zero: lifts into other rows)
    (1, 0.0, "bottom"),               # Dirichlet-Y
    (2, 3e5, "right"),                # Neumann-X
    (4, 2e5, "right"),                # Neumann-normal
    (5, 1e5, "top"),                  # Neumann-tangent, adds to the top-left DOF fixed above
    (6, 2e-5, "corner_br"),           # Dirichlet-normal on the bottom-right corner
    (7, -1e-5, "corner_tr"),          # Dirichlet-tangent on the top-right corner
    (1, 1e-5, "bottom"),              # Dirichlet-Y again with another value
]


def node_sets(m):
    s = dict(m.node_sets)
    br = s["bottom"] & s["right"]
    tr = s["top"] & s["right"]
    s["corner_br"], s["corner_tr"] = br, tr
    return s


def ref_system(L, mesh, stress, E, nu, rho, conditions, gravity, solve=True):
    """bfm_system_create_planar_* (stress 0 / 1) or
    bfm_system_create_axisymmetric_strain (stress 2) on the reference: (K
    dense, b) after conditions, and u from its sim."""
    cm = pb.Mesh_arrays.from_generated(mesh, lib=L)
    mat = pb.Material("steel", rho, E, nu, lib=L)
    rule = pb.Rule_gauss_legendre(2, mesh.kind, lib=L)
    obj = pb.Obj(cm, mat, rule, lib=L)
    inst = pb.Instance(obj, lib=L)
    sets = node_sets(mesh)
    conds = []
    for kind, value, where in conditions:
        c = pb.Condition(cm, kind, value, lib=L)
        c.set_nodes(sets[where])
        inst.add_condition(c)
        conds.append(c)
    forces = []
    if gravity:
        forces.append(pb.Force_linear((0.0, -9.81), lib=L))
    arr = (C.POINTER(F.bfm_force_t) * max(1, len(forces)))()
    for i, f in enumerate(forces):
        arr[i] = C.pointer(f.c_force)
    sysm = F.bfm_system_t()
    fn = (L.bfm_system_create_axisymmetric_strain if stress == 2 else
          L.bfm_system_create_planar_stress if stress else L.bfm_system_create_planar_strain)
    assert fn(C.byref(sysm), C.byref(inst.c_instance), len(forces), arr) == 0
    m = mesh.n_dof
    K = np.ctypeslib.as_array(sysm.A.full.data, shape=(m * m,)).reshape(m, m).copy()
    b = np.ctypeslib.as_array(sysm.b.data, shape=(m,)).copy()
    L.bfm_system_destroy(C.byref(sysm))

    if not solve:
        return K, b, None
    sim = pb.Sim(pb.Sim.AXISYMMETRIC_STRAIN if stress == 2 else
                 pb.Sim.PLANAR_STRESS if stress else pb.Sim.PLANAR_STRAIN, lib=L)
    sim.add_instance(inst)
    for f in forces:
        sim.add_force(f)
    sim.run()
    u = inst.effects
    del sim
    return K, b, u


def add_case(out, name, K, b, u, dense_limit=64):
    m = b.size
    if m <= dense_limit:
        out[f"{name}/K"] = K
    else:
        r, c = np.nonzero(K)
        out[f"{name}/K_row"] = r.astype(np.int32)
        out[f"{name}/K_col"] = c.astype(np.int32)
        out[f"{name}/K_val"] = K[r, c]
    out[f"{name}/b"] = b
    out[f"{name}/u"] = u


def main():
    if not os.path.exists(REF_LIB):
        raise SystemExit(f"{REF_LIB} missing: run `make -C oracle ref` first")
    L = pb.Library(REF_LIB)
    out = {}

    # C1 beam: 40 x 10 P1, plane stress, traction on top, left clamped.
    beam = meshgen.beam(40, 10, kind=3)
    mat = meshgen.BEAM_MATERIAL
    conds = [(3, meshgen.BEAM_TRACTION, "top"), (0, 0.0, "left"), (1, 0.0, "left")]
    K, b, u = ref_system(L, beam, True, mat["E"], mat["nu"], mat["rho"], conds, False)
    add_case(out, "c1", K, b, u)
    K, b, u = ref_system(L, beam, True, mat["E"], mat["nu"], mat["rho"], conds, True)
    add_case(out, "c1_gravity", K, b, u)
    # Neumann listed after the Dirichlet pair: the clamped top-left DOFs pick up the load.
    K, b, u = ref_system(L, beam, True, mat["E"], mat["nu"], mat["rho"], conds[1:] + conds[:1], False)
    add_case(out, "c1_neumann_last", K, b, u)

    # Tiny jittered meshes with every condition kind; pre-condition K and b too.
    for kind, tag in ((4, "q1"), (3, "p1")):
        tiny = meshgen.beam(3, 2, kind=kind, length=3.0, height=2.0, jitter=0.2, seed=11)
        for stress in (0, 1):
            name = f"tiny_{tag}_{'stress' if stress else 'strain'}"
            K, b, u = ref_system(L, tiny, stress, 7.0e10, 0.33, 2700.0, TINY_CONDITIONS, True)
            add_case(out, name, K, b, u)
            K0, b0, _ = ref_system(L, tiny, stress, 7.0e10, 0.33, 2700.0, [], True, solve=False)
            out[f"{name}/K_pre"] = K0
            out[f"{name}/b_pre"] = b0

    # Axisymmetric (r = x + 1 > 0): tiny meshes with every condition kind and
    # a body load, before and after the conditions, and an annulus-like beam.
    for kind, tag in ((4, "q1"), (3, "p1")):
        tiny = meshgen.beam(3, 2, kind=kind, length=3.0, height=2.0, jitter=0.2, seed=11)
        tiny.coords = tiny.coords + np.array([1.0, 0.0])
        name = f"axi_{tag}"
        K, b, u = ref_system(L, tiny, 2, 7.0e10, 0.33, 2700.0, TINY_CONDITIONS, True)
        add_case(out, name, K, b, u)
        K0, b0, _ = ref_system(L, tiny, 2, 7.0e10, 0.33, 2700.0, [], True, solve=False)
        out[f"{name}/K_pre"] = K0
        out[f"{name}/b_pre"] = b0
    ring = meshgen.beam(40, 10, kind=4, length=4.0, height=1.0, jitter=0.2, seed=12)
    ring.coords = ring.coords + np.array([1.0, 0.0])
    conds = [(1, 0.0, "bottom"), (0, 0.0, "left"), (3, -1e6, "top")]
    K, b, u = ref_system(L, ring, 2, 2.0e11, 0.3, 7850.0, conds, True)
    add_case(out, "axi_ring", K, b, u)

    # Single elements: K of a one-element mesh is the element matrix.
    rng = np.random.default_rng(5)
    for kind, tag in ((4, "q1"), (3, "p1")):
        for t in range(3):
            m1 = meshgen.beam(1, 1, kind=kind)
            if kind == 3:
                m1.elems = m1.elems[:1]
            m1.coords = m1.coords + rng.uniform(-0.15, 0.15, size=m1.coords.shape)
            K0, b0, _ = ref_system(L, m1, t % 2, 2.0e11, 0.25 + 0.05 * t, 7800.0, [], True, solve=False)
            out[f"elem_{tag}_{t}/coords"] = m1.coords
            out[f"elem_{tag}_{t}/elems"] = m1.elems.astype(np.int64)
            out[f"elem_{tag}_{t}/stress"] = np.array(t % 2)
            out[f"elem_{tag}_{t}/nu"] = np.array(0.25 + 0.05 * t)
            out[f"elem_{tag}_{t}/K"] = K0
            out[f"elem_{tag}_{t}/b"] = b0

    # LEPL1110 mesh + problem through the ez front end, and the parsed mesh.
    data_dir = os.path.join(HERE, "ref_data")
    os.makedirs(data_dir, exist_ok=True)
    for rel in DATA_FILES:
        shutil.copy(os.path.join(REF_TREE, rel), os.path.join(data_dir, os.path.basename(rel)))
    mesh = pb.Mesh_lepl1110(os.path.join(data_dir, "8.lepl1110"), lib=L)
    ez = pb.Ez_lepl1110(mesh, os.path.join(data_dir, "problem.txt"), lib=L)
    ez.sim.run()
    out["lepl8/u"] = ez.effects
    out["lepl8/coords"] = mesh.coords
    out["lepl8/elems"] = mesh.elems.astype(np.int64)
    en, ee = mesh.edges
    out["lepl8/edge_nodes"] = en.astype(np.int64)
    out["lepl8/edge_elems"] = ee
    out["lepl8/sim_kind"] = np.array(ez.c_ez.sim.kind)
    out["lepl8/E_nu_rho"] = np.array([ez.c_ez.material.E, ez.c_ez.material.nu, ez.c_ez.material.rho])
    out["lepl8/n_conditions"] = np.array(ez.c_ez.n_conditions)
    flags = []
    for i in range(ez.c_ez.n_conditions):
        c = ez.c_ez.conditions[i]
        flags.append(np.ctypeslib.as_array(C.cast(c.nodes, C.POINTER(C.c_uint8)), shape=(mesh.n_nodes,)).copy())
        out[f"lepl8/cond{i}_kind_value"] = np.array([c.kind, c.value])
    out["lepl8/cond_flags"] = np.stack(flags)
    ez.write(os.path.join("/tmp", "golden_U.txt"), 0)
    with open("/tmp/golden_U.txt") as f:
        out["lepl8/U_txt"] = np.frombuffer(f.read().encode(), dtype=np.uint8)
    del ez

    # Wavefront reader with its edge computation.
    obj = pb.Mesh_wavefront(os.path.join(data_dir, "bridge-nothing.obj"), lib=L)
    out["obj/coords"] = obj.coords
    out["obj/elems"] = obj.elems.astype(np.int64)
    en, ee = obj.edges
    out["obj/edge_nodes"] = en.astype(np.int64)
    out["obj/edge_elems"] = ee

    np.savez_compressed(os.path.join(HERE, "golden.npz"), **out)
    print("wrote", os.path.join(HERE, "golden.npz"), "with", len(out), "arrays")


if __name__ == "__main__":
    main()
