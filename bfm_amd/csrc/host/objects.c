/* Materials, conditions, objects and instances (libbfm/src/material.c,
 * condition.c, obj.c, instance.c). Objects hold borrowed pointers; buffers
 * come from state->alloc. */

#include <string.h>

#include "internal.h"

int bfm_material_create(bfm_material_t* material, bfm_state_t* state, char* name, double rho, double E, double nu) {
	memset(material, 0, sizeof *material);
	material->state = state;
	size_t const len = strlen(name);
	material->name = state->alloc(len + 1);
	if (!material->name) {
		BFM_ERROR(state, "cannot allocate the material name");
		return -1;
	}
	memcpy(material->name, name, len + 1);
	material->rho = rho;
	material->E = E;
	material->nu = nu;
	return 0;
}

int bfm_material_destroy(bfm_material_t* material) {
	material->state->free(material->name);
	material->name = NULL;
	return 0;
}

int bfm_material_set_colour(bfm_material_t* material, double r, double g, double b, double a) {
	material->colour = (bfm_colour_t){r, g, b, a};
	return 0;
}

int bfm_condition_create(bfm_condition_t* condition, bfm_state_t* state, bfm_mesh_t* mesh, bfm_condition_kind_t kind) {
	condition->state = state;
	condition->mesh = mesh;
	condition->kind = kind;
	condition->value = 0;
	condition->nodes = state->alloc(mesh->n_nodes ? mesh->n_nodes : 1);
	if (!condition->nodes) {
		BFM_ERROR(state, "cannot allocate %zu node flags", mesh->n_nodes);
		return -1;
	}
	memset(condition->nodes, 0, mesh->n_nodes * sizeof *condition->nodes);
	return 0;
}

int bfm_condition_destroy(bfm_condition_t* condition) {
	condition->state->free(condition->nodes);
	condition->nodes = NULL;
	return 0;
}

int bfm_obj_create(bfm_obj_t* obj, bfm_state_t* state, bfm_mesh_t* mesh, bfm_material_t* material, bfm_rule_t* rule) {
	obj->state = state;
	obj->mesh = mesh;
	obj->material = material;
	obj->rule = rule;
	return 0;
}

int bfm_obj_destroy(bfm_obj_t* obj) {
	(void) obj;
	return 0;
}

int bfm_instance_create(bfm_instance_t* instance, bfm_state_t* state, bfm_obj_t* obj) {
	memset(instance, 0, sizeof *instance);
	instance->state = state;
	instance->obj = obj;
	instance->n_effects = obj->mesh->n_nodes * obj->mesh->dim;
	size_t const size = instance->n_effects * sizeof *instance->effects;
	instance->effects = state->alloc(size ? size : 1);
	if (!instance->effects) {
		BFM_ERROR(state, "cannot allocate %zu effects", instance->n_effects);
		return -1;
	}
	memset(instance->effects, 0, size);
	return 0;
}

int bfm_instance_destroy(bfm_instance_t* instance) {
	bfm_state_t* const state = instance->state;
	if (instance->effects) {
		state->free(instance->effects);
	}
	if (instance->conditions) {
		state->free(instance->conditions);
	}
	instance->effects = NULL;
	instance->conditions = NULL;
	return 0;
}

int bfm_instance_set_n_conditions(bfm_instance_t* instance, size_t n_conditions) {
	bfm_state_t* const state = instance->state;
	if (instance->conditions) {
		state->free(instance->conditions);
	}
	instance->n_conditions = n_conditions;
	size_t const size = n_conditions * sizeof *instance->conditions;
	instance->conditions = state->alloc(size ? size : 1);
	if (!instance->conditions) {
		BFM_ERROR(state, "cannot allocate %zu condition slots", n_conditions);
		return -1;
	}
	memset(instance->conditions, 0, size);
	return 0;
}

int bfm_instance_add_condition(bfm_instance_t* instance, bfm_condition_t* condition) {
	bfm_state_t* const state = instance->state;
	bfm_condition_t** grown = state->realloc(instance->conditions, (instance->n_conditions + 1) * sizeof *grown);
	if (!grown) {
		BFM_ERROR(state, "cannot grow the condition list");
		return -1;
	}
	instance->conditions = grown;
	instance->conditions[instance->n_conditions++] = condition;
	return 0;
}
