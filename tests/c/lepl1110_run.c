/* A libbfm client written against the public headers only, as
 * lepl1110.py drives the library (lepl1110.py:9-14): read a mesh and a
 * problem file, run the simulation, write U and V. */
#include <stdio.h>

#include <bfm/bfm.h>
#include <bfm/ez.h>
#include <bfm/mesh.h>

int main(int argc, char** argv) {
	if (argc != 5) {
		fprintf(stderr, "usage: %s mesh problem U.txt V.txt\n", argv[0]);
		return 2;
	}
	bfm_state_t state;
	bfm_mesh_t mesh;
	bfm_ez_lepl1110_t ez;
	bfm_state_create(&state);
	if (bfm_mesh_read_lepl1110(&mesh, &state, argv[1]) < 0 || bfm_ez_lepl1110_create(&ez, &state, &mesh, argv[2]) < 0 ||
	    bfm_sim_run(&ez.sim) < 0 || bfm_ez_lepl1110_write(&ez, 0, argv[3]) < 0 ||
	    bfm_ez_lepl1110_write(&ez, 1, argv[4]) < 0) {
		bfm_err_print(&state);
		return 1;
	}
	bfm_ez_lepl1110_destroy(&ez);
	bfm_mesh_destroy(&mesh);
	bfm_state_destroy(&state);
	return 0;
}
