/* LEPL1110 problem files (libbfm/src/ez.c).
 *
 * Keys are matched on the first 19 characters of a line, case-insensitively,
 * as the reference's scanner does (ez.c:50-177):
 *   Type of problem     :  Planar strains | Planar stresses | Axi-symetric problem
 *                           (first 13 characters compared)
 *   Young modulus / Poisson ratio / Mass density :  <value>
 *   Gravity             :  g      -> linear body force (0, -g), added per line
 *   Boundary condition  :  <Kind> = <value> : <domain name>
 * A condition's nodes are both ends of every edge listed by the named domain
 * (the first domain whose name matches on 25 characters, trailing spaces
 * included). Conditions are added to the instance in file order. An unknown
 * condition kind fails; other unknown lines are skipped. Blank lines are
 * skipped (the reference's scanner would re-read its previous key there). */

#include <stdio.h>
#include <string.h>
#include <strings.h>

#include "internal.h"

static char* skip_ws(char* s) {
	while (*s == ' ' || *s == '\t') {
		s++;
	}
	return s;
}

/* The text after the key's ':' (or NULL). */
static char* value_of(char* line) {
	size_t const len = strlen(line);
	char* colon = strchr(line + (len < 19 ? len : 19), ':');
	return colon ? skip_ws(colon + 1) : NULL;
}

static int add_condition(bfm_ez_lepl1110_t* ez, bfm_condition_kind_t kind, double value, char const* domain) {
	bfm_state_t* const state = ez->state;
	bfm_mesh_t* const mesh = ez->mesh;
	bfm_condition_t* grown = state->realloc(ez->conditions, (ez->n_conditions + 1) * sizeof *grown);
	if (!grown) {
		BFM_ERROR(state, "out of memory");
		return -1;
	}
	ez->conditions = grown;
	bfm_condition_t* const c = &ez->conditions[ez->n_conditions];
	if (bfm_condition_create(c, state, mesh, kind) < 0) {
		return -1;
	}
	ez->n_conditions++;
	c->value = value;
	for (size_t d = 0; d < mesh->n_domains; d++) {
		if (strncasecmp(mesh->domains[d].name, domain, 25) != 0) {
			continue;
		}
		bfm_domain_t const* const dom = &mesh->domains[d];
		for (size_t j = 0; j < dom->n_elements; j++) {
			size_t const e = dom->elements[j];
			if (e >= mesh->n_edges) {
				BFM_ERROR(state, "domain %s lists edge %zu of %zu", dom->name, e, mesh->n_edges);
				return -1;
			}
			c->nodes[mesh->edges[e].nodes[0]] = true;
			c->nodes[mesh->edges[e].nodes[1]] = true;
		}
		break;
	}
	return 0;
}

static int condition_kind(char const* s, bfm_condition_kind_t* kind) {
	static struct {
		char const* name;
		bfm_condition_kind_t kind;
	} const kinds[] = {
	    {"Dirichlet-X", BFM_CONDITION_KIND_DIRICHLET_X},
	    {"Dirichlet-Y", BFM_CONDITION_KIND_DIRICHLET_Y},
	    {"Neumann-X", BFM_CONDITION_KIND_NEUMANN_X},
	    {"Neumann-Y", BFM_CONDITION_KIND_NEUMANN_Y},
	    {"Neumann-Tangent", BFM_CONDITION_KIND_NEUMANN_TANGENT},
	    {"Neumann-Normal", BFM_CONDITION_KIND_NEUMANN_NORMAL},
	    {"Dirichlet-Normal", BFM_CONDITION_KIND_DIRICHLET_NORMAL},
	    {"Dirichlet-Tangent", BFM_CONDITION_KIND_DIRICHLET_TANGENT},
	};
	for (size_t i = 0; i < sizeof kinds / sizeof *kinds; i++) {
		if (strncasecmp(s, kinds[i].name, 19) == 0) {
			*kind = kinds[i].kind;
			return 0;
		}
	}
	return -1;
}

int bfm_ez_lepl1110_create(bfm_ez_lepl1110_t* ez, bfm_state_t* state, bfm_mesh_t* mesh, char* name) {
	memset(ez, 0, sizeof *ez);
	ez->state = state;
	ez->mesh = mesh;
	if (bfm_sim_create(&ez->sim, state, BFM_SIM_KIND_NONE) < 0 ||
	    bfm_material_create(&ez->material, state, "lepl1110", 0, 0, 0) < 0 ||
	    bfm_rule_create_gauss_legendre(&ez->rule, state, 2, mesh->kind) < 0 ||
	    bfm_obj_create(&ez->obj, state, mesh, &ez->material, &ez->rule) < 0 ||
	    bfm_instance_create(&ez->instance, state, &ez->obj) < 0 ||
	    bfm_sim_add_instance(&ez->sim, &ez->instance) < 0) {
		return -1;
	}
	FILE* fp = fopen(name, "r");
	if (!fp) {
		BFM_ERROR(state, "cannot open %s", name);
		return -1;
	}
	char line[512];
	int rv = 0;
	while (rv == 0 && fgets(line, sizeof line, fp)) {
		line[strcspn(line, "\n")] = 0;
		if (!*skip_ws(line)) {
			continue;
		}
		char* const v = value_of(line);
		if (!v) {
			continue;
		}
		if (strncasecmp(line, "Type of problem     ", 19) == 0) {
			if (strncasecmp(v, "Planar strains", 13) == 0) {
				ez->sim.kind = BFM_SIM_KIND_PLANAR_STRAIN;
			} else if (strncasecmp(v, "Planar stresses", 13) == 0) {
				ez->sim.kind = BFM_SIM_KIND_PLANAR_STRESS;
			} else if (strncasecmp(v, "Axi-symetric problem", 13) == 0) {
				ez->sim.kind = BFM_SIM_KIND_AXISYMMETRIC_STRAIN;
			}
		} else if (strncasecmp(line, "Young modulus       ", 19) == 0) {
			sscanf(v, "%le", &ez->material.E);
		} else if (strncasecmp(line, "Poisson ratio       ", 19) == 0) {
			sscanf(v, "%le", &ez->material.nu);
		} else if (strncasecmp(line, "Mass density        ", 19) == 0) {
			sscanf(v, "%le", &ez->material.rho);
		} else if (strncasecmp(line, "Gravity             ", 19) == 0) {
			double g = 0, vd[2];
			sscanf(v, "%le", &g);
			vd[0] = 0;
			vd[1] = -g;
			bfm_vec_t gv = {state, 2, vd};
			bfm_force_destroy(&ez->gravity);
			if (bfm_force_create(&ez->gravity, state, 2) < 0 || bfm_force_set_linear(&ez->gravity, &gv) < 0 ||
			    bfm_sim_add_force(&ez->sim, &ez->gravity) < 0) {
				rv = -1;
			}
		} else if (strncasecmp(line, "Boundary condition  ", 19) == 0) {
			char kind_s[20] = "", domain[100] = "";
			double val = 0;
			bfm_condition_kind_t kind;
			if (sscanf(v, "%19s = %le : %99[^\n]", kind_s, &val, domain) < 2 || condition_kind(kind_s, &kind) < 0) {
				BFM_ERROR(state, "%s: bad boundary condition '%s'", name, v);
				rv = -1;
			} else {
				rv = add_condition(ez, kind, val, domain);
			}
		}
	}
	fclose(fp);
	/* pointers into ez->conditions are taken once it stops growing (ez.c:181-186) */
	for (size_t i = 0; rv == 0 && i < ez->n_conditions; i++) {
		rv = bfm_instance_add_condition(&ez->instance, &ez->conditions[i]);
	}
	return rv;
}

int bfm_ez_lepl1110_destroy(bfm_ez_lepl1110_t* ez) {
	bfm_state_t* const state = ez->state;
	for (size_t i = 0; i < ez->n_conditions; i++) {
		bfm_condition_destroy(&ez->conditions[i]);
	}
	state->free(ez->conditions);
	ez->conditions = NULL;
	bfm_force_destroy(&ez->gravity);
	bfm_material_destroy(&ez->material);
	bfm_rule_destroy(&ez->rule);
	bfm_obj_destroy(&ez->obj);
	bfm_instance_destroy(&ez->instance);
	bfm_sim_destroy(&ez->sim);
	return 0;
}

/* "Number of nodes N", then %14.7e per node, three per line (ez.c:211-229). */
int bfm_ez_lepl1110_write(bfm_ez_lepl1110_t* ez, size_t shift, char const* filename) {
	FILE* fp = fopen(filename, "w");
	if (!fp) {
		BFM_ERROR(ez->state, "cannot write %s", filename);
		return -1;
	}
	size_t const n = ez->obj.mesh->n_nodes;
	fprintf(fp, "Number of nodes %zu\n", n);
	for (size_t i = 0; i < n; i++) {
		fprintf(fp, "%14.7e", ez->instance.effects[i * 2 + shift]);
		if (i + 1 != ez->instance.n_effects && (i + 1) % 3 == 0) {
			fprintf(fp, "\n");
		}
	}
	fprintf(fp, "\n");
	fclose(fp);
	return 0;
}
