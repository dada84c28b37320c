#pragma once

/* Linear systems A u = b (libbfm/src/bfm/system.h:9-28).
 *
 * bfm_system_create_planar_strain/stress assemble on the GPU: A comes back
 * as a BFM_MATRIX_KIND_BSR matrix resident in device memory and b as a host
 * vector in the original DOF order (2 * node + component). The reference
 * assembles into a dense host matrix instead (libbfm/src/system.c:5-34). */

#include <bfm/force.h>
#include <bfm/instance.h>
#include <bfm/math.h>
#include <bfm/matrix.h>
#include <bfm/perm.h>

typedef struct {
	bfm_state_t* state;

	size_t n;

	bfm_perm_t perm;
	bfm_matrix_t A;
	bfm_vec_t b;
} bfm_system_t;

int bfm_system_create(bfm_system_t* system, bfm_state_t* state, size_t n);
int bfm_system_destroy(bfm_system_t* system);

int bfm_system_renumber(bfm_system_t* system);

int bfm_system_create_planar_strain(bfm_system_t* system, bfm_instance_t* instance, size_t n_forces, bfm_force_t** forces);
int bfm_system_create_planar_stress(bfm_system_t* system, bfm_instance_t* instance, size_t n_forces, bfm_force_t** forces);
int bfm_system_create_axisymmetric_strain(bfm_system_t* system, bfm_instance_t* instance, size_t n_forces, bfm_force_t** forces);
