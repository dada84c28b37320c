/* Prints the size of every public libbfm struct and the offset of each of
 * its fields, one "struct.field offset" line each, for the ABI layout test.
 * Compiled against this implementation's include/ and, where present, the
 * reference headers: the two must print the same table. */
#include <stddef.h>
#include <stdio.h>

#include <bfm/bfm.h>
#include <bfm/condition.h>
#include <bfm/ez.h>
#include <bfm/force.h>
#include <bfm/instance.h>
#include <bfm/material.h>
#include <bfm/math.h>
#include <bfm/matrix.h>
#include <bfm/mesh.h>
#include <bfm/obj.h>
#include <bfm/perm.h>
#include <bfm/rule.h>
#include <bfm/shape.h>
#include <bfm/sim.h>
#include <bfm/system.h>

#define S(T) printf("%s.sizeof %zu\n", #T, sizeof(T))
#define F(T, f) printf("%s.%s %zu\n", #T, #f, offsetof(T, f))

int main(void) {
	S(bfm_err_t); F(bfm_err_t, has); F(bfm_err_t, msg); F(bfm_err_t, file); F(bfm_err_t, func); F(bfm_err_t, line);
	S(bfm_state_t); F(bfm_state_t, err); F(bfm_state_t, alloc); F(bfm_state_t, realloc); F(bfm_state_t, free);
	S(bfm_vec_t); F(bfm_vec_t, state); F(bfm_vec_t, n); F(bfm_vec_t, data);
	S(bfm_matrix_t); F(bfm_matrix_t, state); F(bfm_matrix_t, kind); F(bfm_matrix_t, major); F(bfm_matrix_t, m);
	F(bfm_matrix_t, full); F(bfm_matrix_t, band);
	S(bfm_edge_t); F(bfm_edge_t, nodes); F(bfm_edge_t, elems);
	S(bfm_domain_t); F(bfm_domain_t, name); F(bfm_domain_t, n_elements); F(bfm_domain_t, elements);
	S(bfm_mesh_t); F(bfm_mesh_t, state); F(bfm_mesh_t, dim); F(bfm_mesh_t, kind); F(bfm_mesh_t, n_elems);
	F(bfm_mesh_t, n_nodes); F(bfm_mesh_t, n_edges); F(bfm_mesh_t, coords); F(bfm_mesh_t, elems);
	F(bfm_mesh_t, edges); F(bfm_mesh_t, n_domains); F(bfm_mesh_t, domains);
	S(bfm_condition_t); F(bfm_condition_t, state); F(bfm_condition_t, mesh); F(bfm_condition_t, kind);
	F(bfm_condition_t, value); F(bfm_condition_t, nodes);
	S(bfm_force_t); F(bfm_force_t, state); F(bfm_force_t, kind); F(bfm_force_t, dim); F(bfm_force_t, linear);
	F(bfm_force_t, funky);
	S(bfm_material_t); F(bfm_material_t, state); F(bfm_material_t, name); F(bfm_material_t, colour);
	F(bfm_material_t, rho); F(bfm_material_t, E); F(bfm_material_t, nu);
	S(bfm_shape_t); F(bfm_shape_t, state); F(bfm_shape_t, dim); F(bfm_shape_t, kind); F(bfm_shape_t, phi);
	F(bfm_shape_t, dphi);
	S(bfm_rule_t); F(bfm_rule_t, state); F(bfm_rule_t, dim); F(bfm_rule_t, kind); F(bfm_rule_t, n_points);
	F(bfm_rule_t, weights); F(bfm_rule_t, points); F(bfm_rule_t, shape);
	S(bfm_obj_t); F(bfm_obj_t, state); F(bfm_obj_t, mesh); F(bfm_obj_t, material); F(bfm_obj_t, rule);
	S(bfm_instance_t); F(bfm_instance_t, state); F(bfm_instance_t, obj); F(bfm_instance_t, n_effects);
	F(bfm_instance_t, effects); F(bfm_instance_t, n_conditions); F(bfm_instance_t, conditions);
	S(bfm_sim_t); F(bfm_sim_t, state); F(bfm_sim_t, kind); F(bfm_sim_t, n_instances); F(bfm_sim_t, instances);
	F(bfm_sim_t, n_forces); F(bfm_sim_t, forces);
	S(bfm_perm_t); F(bfm_perm_t, state); F(bfm_perm_t, m); F(bfm_perm_t, has_perm); F(bfm_perm_t, perm);
	F(bfm_perm_t, inv_perm);
	S(bfm_system_t); F(bfm_system_t, state); F(bfm_system_t, n); F(bfm_system_t, perm); F(bfm_system_t, A);
	F(bfm_system_t, b);
	S(bfm_ez_lepl1110_t); F(bfm_ez_lepl1110_t, state); F(bfm_ez_lepl1110_t, mesh);
	F(bfm_ez_lepl1110_t, n_conditions); F(bfm_ez_lepl1110_t, conditions); F(bfm_ez_lepl1110_t, gravity);
	F(bfm_ez_lepl1110_t, material); F(bfm_ez_lepl1110_t, rule); F(bfm_ez_lepl1110_t, obj);
	F(bfm_ez_lepl1110_t, instance); F(bfm_ez_lepl1110_t, sim);
	printf("enum.BFM_MATRIX_KIND_BAND %d\n", (int) BFM_MATRIX_KIND_BAND);
	printf("enum.BFM_CONDITION_KIND_DIRICHLET_TANGENT %d\n", (int) BFM_CONDITION_KIND_DIRICHLET_TANGENT);
	printf("enum.BFM_SIM_KIND_AXISYMMETRIC_STRAIN %d\n", (int) BFM_SIM_KIND_AXISYMMETRIC_STRAIN);
	printf("enum.BFM_ELEM_KIND_QUADRATIC_TRIANGLE %d\n", (int) BFM_ELEM_KIND_QUADRATIC_TRIANGLE);
	printf("enum.BFM_FORCE_KIND_FUNKY %d\n", (int) BFM_FORCE_KIND_FUNKY);
	return 0;
}
