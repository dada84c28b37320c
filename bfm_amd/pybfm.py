"""Headless counterpart of pybfm's wrapper classes, over ctypes.

Mirrors pybfm/bfm/{state,vec,mesh,condition,force,material,rule,obj,
instance,sim,ez}.py in the reference tree: the same class names, the same
constructor arguments, and the same error behaviour (every non-zero libbfm
return raises AssertionError, as pybfm's `assert not lib.bfm_...` does).
The OpenGL parts of pybfm (Bfm, shaders, VBOs) are not reproduced.

Every class takes an optional `lib=` (a Library); the default is this
package's libbfm.so. Library(path) can also wrap any other build of the
libbfm ABI.
"""

import ctypes as C
import os

import numpy as np

from . import _ffi as F

_HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_LIB = os.path.join(_HERE, "lib", "libbfm.so")


class Library:
    """A loaded libbfm with its own bfm_state_t (pybfm/bfm/state.py:1-4)."""

    def __init__(self, path=DEFAULT_LIB):
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built; run __graft_entry__.build()")
        self.path = path
        self.lib = F.bind(C.CDLL(path))
        self.state = F.bfm_state_t()
        assert not self.lib.bfm_state_create(C.byref(self.state))

    def __getattr__(self, name):
        return getattr(self.lib, name)

    def check(self, rc):
        """pybfm's `assert not lib.bfm_...(...)`, with the library's message."""
        if rc:
            err = self.state.err
            raise AssertionError(err.msg.decode() if err.has and err.msg else f"libbfm call failed ({rc})")


_default = None


def default_library():
    global _default
    if _default is None:
        _default = Library()
    return _default


def _L(lib):
    return lib if lib is not None else default_library()


class Vec:
    """pybfm/bfm/vec.py:4-13."""

    def __init__(self, vec, lib=None):
        self.L = _L(lib)
        self.c_vec = F.bfm_vec_t()
        assert not self.L.bfm_vec_create(C.byref(self.c_vec), C.byref(self.L.state), len(vec))
        for i, v in enumerate(vec):
            self.c_vec.data[i] = v

    def __del__(self):
        self.L.bfm_vec_destroy(C.byref(self.c_vec))


class Mesh:
    """Base mesh; coords and elems as numpy views of the C arrays."""

    SIMPLEX = 3
    QUAD = 4

    def _finish(self):
        self.dim = self.c_mesh.dim
        self.kind = self.c_mesh.kind

    @property
    def n_nodes(self):
        return self.c_mesh.n_nodes

    @property
    def coords(self):
        n = self.c_mesh.n_nodes * self.c_mesh.dim
        return np.ctypeslib.as_array(self.c_mesh.coords, shape=(n,)).copy()

    @property
    def elems(self):
        n = self.c_mesh.n_elems * self.c_mesh.kind
        return np.ctypeslib.as_array(self.c_mesh.elems, shape=(n,)).reshape(-1, self.c_mesh.kind).copy()

    @property
    def edges(self):
        ne = self.c_mesh.n_edges
        nodes = np.array([[self.c_mesh.edges[i].nodes[0], self.c_mesh.edges[i].nodes[1]] for i in range(ne)],
                         dtype=np.uint64).reshape(-1, 2)
        elems = np.array([[self.c_mesh.edges[i].elems[0], self.c_mesh.edges[i].elems[1]] for i in range(ne)],
                         dtype=np.int64).reshape(-1, 2)
        return nodes, elems


class Mesh_lepl1110(Mesh):
    """pybfm/bfm/mesh.py:30-38."""

    def __init__(self, name, lib=None):
        self.L = _L(lib)
        self.c_mesh = F.bfm_mesh_t()
        assert not self.L.bfm_mesh_read_lepl1110(C.byref(self.c_mesh), C.byref(self.L.state), name.encode())
        self._finish()

    def __del__(self):
        self.L.bfm_mesh_destroy(C.byref(self.c_mesh))


class Mesh_wavefront(Mesh):
    """pybfm/bfm/mesh.py:40-48."""

    def __init__(self, name, full=False, lib=None):
        self.L = _L(lib)
        self.c_mesh = F.bfm_mesh_t()
        assert not self.L.bfm_mesh_read_wavefront(C.byref(self.c_mesh), C.byref(self.L.state), name.encode(),
                                                   bool(full))
        self._finish()

    def __del__(self):
        self.L.bfm_mesh_destroy(C.byref(self.c_mesh))


class Mesh_arrays(Mesh):
    """A mesh from numpy arrays (coords (n, 2), elems (E, kind), boundary
    edges (m, 2)). The C struct points at memory owned by this object, so it
    is never passed to bfm_mesh_destroy (the generic pybfm Mesh class calls a
    bfm_mesh_create_generic that libbfm does not export, pybfm/bfm/mesh.py:12)."""

    def __init__(self, coords, elems, edges=None, edge_elem=None, lib=None):
        self.L = _L(lib)
        coords = np.ascontiguousarray(np.asarray(coords, np.float64).reshape(-1, 2))
        elems = np.ascontiguousarray(np.asarray(elems, np.uint64))
        kind = elems.shape[1]
        self.c_mesh = F.bfm_mesh_t()
        assert not self.L.bfm_mesh_create(C.byref(self.c_mesh), C.byref(self.L.state), 2, kind)
        self._coords = coords
        self._elems = elems
        self.c_mesh.n_nodes = coords.shape[0]
        self.c_mesh.n_elems = elems.shape[0]
        self.c_mesh.coords = coords.ctypes.data_as(C.POINTER(C.c_double))
        self.c_mesh.elems = elems.ctypes.data_as(C.POINTER(C.c_size_t))
        if edges is not None:
            edges = np.asarray(edges, np.uint64).reshape(-1, 2)
            arr = (F.bfm_edge_t * max(1, edges.shape[0]))()
            for i, (a, b) in enumerate(edges):
                arr[i].nodes[0], arr[i].nodes[1] = int(a), int(b)
                arr[i].elems[0] = int(edge_elem[i]) if edge_elem is not None else 0
                arr[i].elems[1] = -1
            self._edges = arr
            self.c_mesh.n_edges = edges.shape[0]
            self.c_mesh.edges = C.cast(arr, C.POINTER(F.bfm_edge_t))
        self._finish()

    @classmethod
    def from_generated(cls, m, lib=None):
        return cls(m.coords, m.elems, m.edges, m.edge_elem, lib=lib)


class Condition:
    """pybfm/bfm/condition.py:4-29."""

    DIRICHLET_X = 0
    DIRICHLET_Y = 1
    NEUMANN_X = 2
    NEUMANN_Y = 3
    NEUMANN_NORMAL = 4
    NEUMANN_TANGENT = 5
    DIRICHLET_NORMAL = 6
    DIRICHLET_TANGENT = 7

    def __init__(self, mesh, kind, value=0.0, lib=None):
        self.L = _L(lib)
        self.c_condition = F.bfm_condition_t()
        assert not self.L.bfm_condition_create(C.byref(self.c_condition), C.byref(self.L.state),
                                               C.byref(mesh.c_mesh), kind)
        self.c_condition.value = value
        self.mesh = mesh

    def __del__(self):
        self.L.bfm_condition_destroy(C.byref(self.c_condition))

    def populate(self, discriminator):
        """discriminator(mesh, coord) -> bool, per node (condition.py:26-29)."""
        xy = self.mesh.coords.reshape(-1, self.mesh.dim)
        for i in range(self.mesh.c_mesh.n_nodes):
            self.c_condition.nodes[i] = bool(discriminator(self.mesh, tuple(xy[i])))

    def set_nodes(self, flags):
        flags = np.asarray(flags, dtype=bool)
        buf = np.ctypeslib.as_array(C.cast(self.c_condition.nodes, C.POINTER(C.c_uint8)), shape=(flags.size,))
        buf[:] = flags.astype(np.uint8)


class Force:
    """pybfm/bfm/force.py:7-13."""

    def __init__(self, dim, lib=None):
        self.L = _L(lib)
        self.c_force = F.bfm_force_t()
        assert not self.L.bfm_force_create(C.byref(self.c_force), C.byref(self.L.state), dim)

    def __del__(self):
        self.L.bfm_force_destroy(C.byref(self.c_force))


class Force_none(Force):
    def __init__(self, dim, lib=None):
        super().__init__(dim, lib)
        assert not self.L.bfm_force_set_none(C.byref(self.c_force))


class Force_linear(Force):
    """pybfm/bfm/force.py:22-35; EARTH_GRAVITY_2D is (0, -9.81)."""

    EARTH_GRAVITY_2D = (0.0, -9.81)
    EARTH_GRAVITY = (0.0, -9.81, 0.0)

    def __init__(self, vec, lib=None):
        super().__init__(len(vec), lib)
        self._vec = Vec(vec, lib=self.L)
        assert not self.L.bfm_force_set_linear(C.byref(self.c_force), C.byref(self._vec.c_vec))


class Material:
    """pybfm/bfm/material.py:16-29; AA7075 = (rho 2810, E 71.7 GPa, nu 0.33)."""

    AA7075 = ("AA7075", 2.81e3, 71.7e9, 0.33, (0.667, 0.439, 0.459, 1))

    def __init__(self, name, rho, E, nu, colour=(1, 1, 1, 1), lib=None):
        self.L = _L(lib)
        self.c_material = F.bfm_material_t()
        assert not self.L.bfm_material_create(C.byref(self.c_material), C.byref(self.L.state), name.encode(), rho, E,
                                              nu)
        assert not self.L.bfm_material_set_colour(C.byref(self.c_material), *colour)

    def __del__(self):
        self.L.bfm_material_destroy(C.byref(self.c_material))


class Rule_gauss_legendre:
    """pybfm/bfm/rule.py:18-23."""

    def __init__(self, dim, kind, lib=None):
        self.L = _L(lib)
        self.c_rule = F.bfm_rule_t()
        assert not self.L.bfm_rule_create_gauss_legendre(C.byref(self.c_rule), C.byref(self.L.state), dim, kind)

    def __del__(self):
        self.L.bfm_rule_destroy(C.byref(self.c_rule))


class Obj:
    """pybfm/bfm/obj.py:233-238 (without the GL buffers)."""

    def __init__(self, mesh, material, rule, lib=None):
        self.L = _L(lib)
        self.mesh, self.material, self.rule = mesh, material, rule
        self.c_obj = F.bfm_obj_t()
        assert not self.L.bfm_obj_create(C.byref(self.c_obj), C.byref(self.L.state), C.byref(mesh.c_mesh),
                                         C.byref(material.c_material), C.byref(rule.c_rule))


class Instance:
    """pybfm/bfm/instance.py:141-146, 21-22, 89-102 (effects without GL)."""

    def __init__(self, obj, lib=None):
        self.L = _L(lib)
        self.obj = obj
        self.conditions = []
        self.c_instance = F.bfm_instance_t()
        assert not self.L.bfm_instance_create(C.byref(self.c_instance), C.byref(self.L.state), C.byref(obj.c_obj))

    def __del__(self):
        self.L.bfm_instance_destroy(C.byref(self.c_instance))

    def add_condition(self, condition):
        assert not self.L.bfm_instance_add_condition(C.byref(self.c_instance), C.byref(condition.c_condition))
        self.conditions.append(condition)

    @property
    def effects(self):
        n = self.c_instance.n_effects
        return np.ctypeslib.as_array(self.c_instance.effects, shape=(n,)).copy()


class Sim:
    """pybfm/bfm/sim.py:4-72 (headless)."""

    NONE = 0
    PLANAR_STRAIN = 1
    PLANAR_STRESS = 2
    AXISYMMETRIC_STRAIN = 3

    def __init__(self, kind, lib=None):
        self.L = _L(lib)
        self.c_sim = F.bfm_sim_t()
        self.instances = []
        self.forces = []
        assert not self.L.bfm_sim_create(C.byref(self.c_sim), C.byref(self.L.state), kind)

    def __del__(self):
        self.L.bfm_sim_destroy(C.byref(self.c_sim))

    def add_instance(self, instance):
        assert not self.L.bfm_sim_add_instance(C.byref(self.c_sim), C.byref(instance.c_instance))
        self.instances.append(instance)

    def add_force(self, force):
        assert not self.L.bfm_sim_add_force(C.byref(self.c_sim), C.byref(force.c_force))
        self.forces.append(force)

    def run(self):
        self.L.check(self.L.bfm_sim_run(C.byref(self.c_sim)))


class _SimView:
    """The sim embedded in an Ez_lepl1110 (pybfm/bfm/ez.py:22)."""

    def __init__(self, L, c_sim):
        self.L, self.c_sim = L, c_sim

    def run(self):
        self.L.check(self.L.bfm_sim_run(C.byref(self.c_sim)))


class Ez_lepl1110:
    """pybfm/bfm/ez.py:10-26."""

    def __init__(self, mesh, name, lib=None):
        self.L = _L(lib)
        self.mesh = mesh
        self.c_ez = F.bfm_ez_lepl1110_t()
        assert not self.L.bfm_ez_lepl1110_create(C.byref(self.c_ez), C.byref(self.L.state), C.byref(mesh.c_mesh),
                                                 name.encode())
        self.sim = _SimView(self.L, self.c_ez.sim)

    def __del__(self):
        self.L.bfm_ez_lepl1110_destroy(C.byref(self.c_ez))

    def write(self, filename, shift):
        assert not self.L.bfm_ez_lepl1110_write(C.byref(self.c_ez), shift, filename.encode())

    @property
    def effects(self):
        n = self.c_ez.instance.n_effects
        return np.ctypeslib.as_array(self.c_ez.instance.effects, shape=(n,)).copy()


def read_effects_file(path):
    """Parses a U.txt/V.txt written by bfm_ez_lepl1110_write: a header line,
    then %14.7e fields (fixed width 14, three per line)."""
    with open(path) as f:
        head = f.readline().split()
        n = int(head[-1])
        vals = []
        for line in f:
            line = line.rstrip("\n")
            for k in range(0, len(line), 14):
                tok = line[k:k + 14].strip()
                if tok:
                    vals.append(float(tok))
    assert len(vals) == n, (len(vals), n)
    return np.array(vals)
