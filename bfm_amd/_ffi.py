"""ctypes declarations of the libbfm C ABI (include/bfm/*.h).

The struct layouts are those of libbfm's public headers
(libbfm/src/bfm/*.h in the reference tree), which pybfm cdefs verbatim
(pybfm/bfm/gen_libbfm.py:11-40) and reads field by field, so any library
exporting this ABI can be driven by the same declarations: this
implementation's libbfm.so, or a libbfm built from the reference sources.
cffi is not installed in this image, hence ctypes.
"""

import ctypes as C

c_size_t = C.c_size_t
c_ssize_t = C.c_ssize_t


class bfm_err_t(C.Structure):
    _fields_ = [("has", C.c_bool), ("msg", C.c_char_p), ("file", C.c_char_p),
                ("func", C.c_char_p), ("line", c_size_t)]


ALLOC = C.CFUNCTYPE(C.c_void_p, c_size_t)
REALLOC = C.CFUNCTYPE(C.c_void_p, C.c_void_p, c_size_t)
FREE = C.CFUNCTYPE(None, C.c_void_p)


class bfm_state_t(C.Structure):
    _fields_ = [("err", bfm_err_t), ("alloc", C.c_void_p), ("realloc", C.c_void_p),
                ("free", C.c_void_p)]


class bfm_vec_t(C.Structure):
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("n", c_size_t),
                ("data", C.POINTER(C.c_double))]


class _full(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_double))]


class _band(C.Structure):
    _fields_ = [("k", c_size_t), ("data", C.POINTER(C.c_double))]


class _bsr(C.Structure):
    _fields_ = [("nnzb", c_size_t), ("handle", C.c_void_p)]


class _matrix_u(C.Union):
    _fields_ = [("full", _full), ("band", _band), ("bsr", _bsr)]


MATRIX_KIND_FULL, MATRIX_KIND_BAND, MATRIX_KIND_BSR = 0, 1, 2
MAJOR_ROW, MAJOR_COLUMN = 0, 1


class bfm_matrix_t(C.Structure):
    _anonymous_ = ("u",)
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("kind", C.c_int), ("major", C.c_int),
                ("m", c_size_t), ("u", _matrix_u)]


class bfm_edge_t(C.Structure):
    _fields_ = [("nodes", c_size_t * 2), ("elems", c_ssize_t * 2)]


class bfm_domain_t(C.Structure):
    _fields_ = [("name", C.c_char * 50), ("n_elements", c_size_t),
                ("elements", C.POINTER(c_size_t))]


class bfm_mesh_t(C.Structure):
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("dim", c_size_t), ("kind", C.c_int),
                ("n_elems", c_size_t), ("n_nodes", c_size_t), ("n_edges", c_size_t),
                ("coords", C.POINTER(C.c_double)), ("elems", C.POINTER(c_size_t)),
                ("edges", C.POINTER(bfm_edge_t)), ("n_domains", c_size_t),
                ("domains", C.POINTER(bfm_domain_t))]


class bfm_condition_t(C.Structure):
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("mesh", C.POINTER(bfm_mesh_t)),
                ("kind", C.c_int), ("value", C.c_double), ("nodes", C.POINTER(C.c_bool))]


class bfm_force_t(C.Structure):
    pass


FUNKY = C.CFUNCTYPE(C.c_int, C.POINTER(bfm_force_t), C.POINTER(bfm_vec_t),
                    C.POINTER(bfm_vec_t), C.c_void_p)


class _linear(C.Structure):
    _fields_ = [("force", bfm_vec_t)]


class _funky(C.Structure):
    _fields_ = [("func", C.c_void_p), ("data", C.c_void_p)]


class _force_u(C.Union):
    _fields_ = [("linear", _linear), ("funky", _funky)]


bfm_force_t._anonymous_ = ("u",)
bfm_force_t._fields_ = [("state", C.POINTER(bfm_state_t)), ("kind", C.c_int),
                        ("dim", c_size_t), ("u", _force_u)]


class bfm_colour_t(C.Structure):
    _fields_ = [("r", C.c_double), ("g", C.c_double), ("b", C.c_double), ("a", C.c_double)]


class bfm_material_t(C.Structure):
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("name", C.c_char_p),
                ("colour", bfm_colour_t), ("rho", C.c_double), ("E", C.c_double),
                ("nu", C.c_double)]


class bfm_shape_t(C.Structure):
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("dim", c_size_t), ("kind", C.c_int),
                ("phi", C.c_void_p), ("dphi", C.c_void_p)]


class bfm_rule_t(C.Structure):
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("dim", c_size_t), ("kind", C.c_int),
                ("n_points", c_size_t), ("weights", C.POINTER(C.c_double)),
                ("points", C.POINTER(C.POINTER(C.c_double))), ("shape", bfm_shape_t)]


class bfm_obj_t(C.Structure):
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("mesh", C.POINTER(bfm_mesh_t)),
                ("material", C.POINTER(bfm_material_t)), ("rule", C.POINTER(bfm_rule_t))]


class bfm_instance_t(C.Structure):
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("obj", C.POINTER(bfm_obj_t)),
                ("n_effects", c_size_t), ("effects", C.POINTER(C.c_double)),
                ("n_conditions", c_size_t),
                ("conditions", C.POINTER(C.POINTER(bfm_condition_t)))]


class bfm_sim_t(C.Structure):
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("kind", C.c_int),
                ("n_instances", c_size_t), ("instances", C.POINTER(C.POINTER(bfm_instance_t))),
                ("n_forces", c_size_t), ("forces", C.POINTER(C.POINTER(bfm_force_t)))]


class bfm_perm_t(C.Structure):
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("m", c_size_t), ("has_perm", C.c_bool),
                ("perm", C.POINTER(c_size_t)), ("inv_perm", C.POINTER(c_size_t))]


class bfm_system_t(C.Structure):
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("n", c_size_t), ("perm", bfm_perm_t),
                ("A", bfm_matrix_t), ("b", bfm_vec_t)]


class bfm_ez_lepl1110_t(C.Structure):
    _fields_ = [("state", C.POINTER(bfm_state_t)), ("mesh", C.POINTER(bfm_mesh_t)),
                ("n_conditions", c_size_t), ("conditions", C.POINTER(bfm_condition_t)),
                ("gravity", bfm_force_t), ("material", bfm_material_t), ("rule", bfm_rule_t),
                ("obj", bfm_obj_t), ("instance", bfm_instance_t), ("sim", bfm_sim_t)]


P = C.POINTER
_I = C.c_int
# name -> (restype, argtypes); the 67 entry points of the 15 public headers.
SIGNATURES = {
    # bfm.h
    "bfm_state_create": (_I, [P(bfm_state_t)]),
    "bfm_state_destroy": (_I, [P(bfm_state_t)]),
    "bfm_set_alloc": (_I, [P(bfm_state_t), C.c_void_p]),
    "bfm_set_realloc": (_I, [P(bfm_state_t), C.c_void_p]),
    "bfm_set_free": (_I, [P(bfm_state_t), C.c_void_p]),
    "bfm_err_print": (_I, [P(bfm_state_t)]),
    # math.h
    "bfm_vec_create": (_I, [P(bfm_vec_t), P(bfm_state_t), c_size_t]),
    "bfm_vec_copy": (_I, [P(bfm_vec_t), P(bfm_vec_t)]),
    "bfm_vec_destroy": (_I, [P(bfm_vec_t)]),
    # matrix.h
    "bfm_matrix_full_create": (_I, [P(bfm_matrix_t), P(bfm_state_t), _I, c_size_t]),
    "bfm_matrix_band_create": (_I, [P(bfm_matrix_t), P(bfm_state_t), _I, c_size_t, c_size_t]),
    "bfm_matrix_copy": (_I, [P(bfm_matrix_t), P(bfm_matrix_t)]),
    "bfm_matrix_destroy": (_I, [P(bfm_matrix_t)]),
    "bfm_matrix_get": (C.c_double, [P(bfm_matrix_t), c_size_t, c_size_t]),
    "bfm_matrix_set": (_I, [P(bfm_matrix_t), c_size_t, c_size_t, C.c_double]),
    "bfm_matrix_add": (_I, [P(bfm_matrix_t), c_size_t, c_size_t, C.c_double]),
    "bfm_matrix_bandwidth": (c_size_t, [P(bfm_matrix_t)]),
    "bfm_matrix_lu": (_I, [P(bfm_matrix_t)]),
    "bfm_matrix_lu_solve": (_I, [P(bfm_matrix_t), P(bfm_vec_t)]),
    "bfm_matrix_solve": (_I, [P(bfm_matrix_t), P(bfm_vec_t)]),
    # mesh.h
    "bfm_mesh_create": (_I, [P(bfm_mesh_t), P(bfm_state_t), c_size_t, _I]),
    "bfm_mesh_destroy": (_I, [P(bfm_mesh_t)]),
    "bfm_mesh_read_lepl1110": (_I, [P(bfm_mesh_t), P(bfm_state_t), C.c_char_p]),
    "bfm_mesh_read_wavefront": (_I, [P(bfm_mesh_t), P(bfm_state_t), C.c_char_p, C.c_bool]),
    # condition.h
    "bfm_condition_create": (_I, [P(bfm_condition_t), P(bfm_state_t), P(bfm_mesh_t), _I]),
    "bfm_condition_destroy": (_I, [P(bfm_condition_t)]),
    # force.h
    "bfm_force_create": (_I, [P(bfm_force_t), P(bfm_state_t), c_size_t]),
    "bfm_force_destroy": (_I, [P(bfm_force_t)]),
    "bfm_force_set_none": (_I, [P(bfm_force_t)]),
    "bfm_force_set_linear": (_I, [P(bfm_force_t), P(bfm_vec_t)]),
    "bfm_force_set_funky": (_I, [P(bfm_force_t), C.c_void_p, C.c_void_p]),
    "bfm_force_eval": (_I, [P(bfm_force_t), P(bfm_vec_t), P(bfm_vec_t)]),
    # material.h
    "bfm_material_create": (_I, [P(bfm_material_t), P(bfm_state_t), C.c_char_p, C.c_double,
                                 C.c_double, C.c_double]),
    "bfm_material_destroy": (_I, [P(bfm_material_t)]),
    "bfm_material_set_colour": (_I, [P(bfm_material_t), C.c_double, C.c_double, C.c_double,
                                     C.c_double]),
    # shape.h / rule.h
    "bfm_shape_create": (_I, [P(bfm_shape_t), P(bfm_state_t), c_size_t, _I]),
    "bfm_shape_destroy": (_I, [P(bfm_shape_t)]),
    "bfm_rule_create": (_I, [P(bfm_rule_t), P(bfm_state_t), c_size_t, _I, c_size_t]),
    "bfm_rule_destroy": (_I, [P(bfm_rule_t)]),
    "bfm_rule_create_gauss_legendre": (_I, [P(bfm_rule_t), P(bfm_state_t), c_size_t, _I]),
    # obj.h / instance.h
    "bfm_obj_create": (_I, [P(bfm_obj_t), P(bfm_state_t), P(bfm_mesh_t), P(bfm_material_t),
                            P(bfm_rule_t)]),
    "bfm_obj_destroy": (_I, [P(bfm_obj_t)]),
    "bfm_instance_create": (_I, [P(bfm_instance_t), P(bfm_state_t), P(bfm_obj_t)]),
    "bfm_instance_destroy": (_I, [P(bfm_instance_t)]),
    "bfm_instance_set_n_conditions": (_I, [P(bfm_instance_t), c_size_t]),
    "bfm_instance_add_condition": (_I, [P(bfm_instance_t), P(bfm_condition_t)]),
    # sim.h
    "bfm_sim_create": (_I, [P(bfm_sim_t), P(bfm_state_t), _I]),
    "bfm_sim_destroy": (_I, [P(bfm_sim_t)]),
    "bfm_sim_set_n_instances": (_I, [P(bfm_sim_t), c_size_t]),
    "bfm_sim_add_instance": (_I, [P(bfm_sim_t), P(bfm_instance_t)]),
    "bfm_sim_set_n_forces": (_I, [P(bfm_sim_t), c_size_t]),
    "bfm_sim_add_force": (_I, [P(bfm_sim_t), P(bfm_force_t)]),
    "bfm_sim_run": (_I, [P(bfm_sim_t)]),
    # perm.h
    "bfm_perm_create": (_I, [P(bfm_perm_t), P(bfm_state_t), c_size_t]),
    "bfm_perm_destroy": (_I, [P(bfm_perm_t)]),
    "bfm_perm_perm_matrix": (_I, [P(bfm_perm_t), P(bfm_matrix_t), C.c_bool]),
    "bfm_perm_perm_vec": (_I, [P(bfm_perm_t), P(bfm_vec_t), C.c_bool]),
    "bfm_perm_rcm": (_I, [P(bfm_perm_t), P(bfm_matrix_t)]),
    # system.h
    "bfm_system_create": (_I, [P(bfm_system_t), P(bfm_state_t), c_size_t]),
    "bfm_system_destroy": (_I, [P(bfm_system_t)]),
    "bfm_system_renumber": (_I, [P(bfm_system_t)]),
    "bfm_system_create_planar_strain": (_I, [P(bfm_system_t), P(bfm_instance_t), c_size_t,
                                             P(P(bfm_force_t))]),
    "bfm_system_create_planar_stress": (_I, [P(bfm_system_t), P(bfm_instance_t), c_size_t,
                                             P(P(bfm_force_t))]),
    "bfm_system_create_axisymmetric_strain": (_I, [P(bfm_system_t), P(bfm_instance_t), c_size_t,
                                                   P(P(bfm_force_t))]),
    # ez.h
    "bfm_ez_lepl1110_create": (_I, [P(bfm_ez_lepl1110_t), P(bfm_state_t), P(bfm_mesh_t),
                                    C.c_char_p]),
    "bfm_ez_lepl1110_destroy": (_I, [P(bfm_ez_lepl1110_t)]),
    "bfm_ez_lepl1110_write": (_I, [P(bfm_ez_lepl1110_t), c_size_t, C.c_char_p]),
}


class bfm_solver_t(C.Structure):
    """include/bfm/gpu.h (this implementation's extension; no reference counterpart)."""
    _fields_ = [("precond", C.c_int), ("precision", C.c_int), ("rtol", C.c_double), ("max_iter", c_size_t)]


class bfm_solve_stats_t(C.Structure):
    _fields_ = [("status", C.c_int), ("iterations", c_size_t), ("rel_residual", C.c_double),
                ("solve_ms", C.c_double), ("total_ms", C.c_double)]


# include/bfm/gpu.h: bound only where the library exports them (a libbfm
# built from the reference sources does not)
EXTENSIONS = {
    "bfm_sim_set_solver": (_I, [P(bfm_sim_t), P(bfm_solver_t)]),
    "bfm_sim_get_solver": (_I, [P(bfm_sim_t), P(bfm_solver_t)]),
    "bfm_sim_set_model": (_I, [P(bfm_sim_t), P(C.c_float), c_size_t]),
    "bfm_sim_get_stats": (_I, [P(bfm_sim_t), c_size_t, P(bfm_solve_stats_t)]),
}


def bind(lib):
    """Sets restype/argtypes of every libbfm entry point on a loaded CDLL."""
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    for name, (res, args) in EXTENSIONS.items():
        fn = getattr(lib, name, None)
        if fn is not None:
            fn.restype = res
            fn.argtypes = args
    return lib
