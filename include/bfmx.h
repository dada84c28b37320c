#pragma once

/* C ABI of libbfmhip.so, the GPU (gfx950) implementation of libbfm's
 * planar-elasticity path: mesh upload and node renumbering, element
 * assembly, Dirichlet/Neumann application and the global linear solve.
 *
 * The libbfm host layer (bfm_system_create_planar_*, bfm_matrix_solve,
 * bfm_sim_run) calls these entry points; tests and benchmarks may call them
 * directly. Only plain pointers, sizes and scalars cross this boundary. All
 * host arrays use libbfm's original numbering: node i, DOF 2*i + component.
 * The node renumbering done at creation is internal to the system object.
 *
 * Every entry point returns BFMX_OK (0) on success or a negative BFMX_E_*
 * code; bfmx_last_error() returns a message for the calling thread's last
 * failure. Calls on one system object are not thread-safe; distinct system
 * objects may be used from distinct threads.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * reference tree):
 *   bfmx_system_create       bfm_system_create + get_elem + bfm_system_renumber
 *                            (libbfm/src/system.c:5-34, 44-81, 93-107;
 *                            libbfm/src/perm.c:117-331)
 *   bfmx_system_set_edges    the mesh->edges loops of create_planar
 *                            (libbfm/src/system.c:477-522, 397-420)
 *   bfmx_system_assemble     the element loop of create_planar and
 *                            fill_elasticity_elem (libbfm/src/system.c:109-227,
 *                            454-466)
 *   bfmx_system_apply_condition  one iteration of the condition loop of
 *                            create_planar (libbfm/src/system.c:470-526)
 *   bfmx_system_solve        bfm_matrix_solve + bfm_perm_perm_vec(inv) in
 *                            run_elasticity (libbfm/src/sim.c:122-131;
 *                            libbfm/src/matrix.c:531-541)
 *   bfmx_system_get_entry    bfm_matrix_get (libbfm/src/matrix.c:455-465)
 */

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
	BFMX_OK = 0,
	BFMX_E_ARG = -1,         /* invalid argument or call order */
	BFMX_E_HIP = -2,         /* HIP runtime error */
	BFMX_E_NOMEM = -3,       /* host or device allocation failed */
	BFMX_E_BREAKDOWN = -4,   /* CG breakdown: p.Ap <= 0 or not finite */
	BFMX_E_MAXITER = -5,     /* tolerance not reached within max_iter */
	BFMX_E_NODEV = -6,       /* no usable GPU */
	BFMX_E_UNSUPPORTED = -7, /* element kind or option not supported */
	BFMX_E_STAGNATION = -8,  /* the true residual stopped decreasing above rtol
	                            (rtol below the attainable accuracy) */
};

/* Preconditioners. */
enum {
	BFMX_PRECOND_JACOBI = 0,       /* 1 / diag(A), per DOF */
	BFMX_PRECOND_BLOCK_JACOBI = 1, /* inverse of each node's 2x2 diagonal block */
	BFMX_PRECOND_AMG = 2,          /* flexible CG (FP64) preconditioned by one FP32
	                                  V-cycle of aggregation multigrid built on the
	                                  rigid-body modes; precision is ignored */
	BFMX_PRECOND_LEARNED = 3,      /* flexible CG (FP64) preconditioned by the
	                                  message-passing network attached with
	                                  bfmx_system_set_model (BF16/FP32); falls back
	                                  to block-Jacobi PCG from the current iterate
	                                  on breakdown, the iteration limit or a stall */
	BFMX_PRECOND_MINRES = 4,       /* MINRES (FP64) preconditioned by 1 / |diag(A)|,
	                                  for symmetric indefinite systems (the
	                                  axisymmetric formulation); precision is
	                                  ignored */
	BFMX_PRECOND_AMG_LEARNED = 5,  /* flexible CG (FP64) preconditioned by the FP32
	                                  multigrid V-cycle plus the correction of the
	                                  network attached with bfmx_system_set_model
	                                  (trained on pairs rebased with
	                                  bfmx_trainset_rebase): z = V(r) + the
	                                  network's SPD term; falls back to
	                                  block-Jacobi PCG as BFMX_PRECOND_LEARNED */
};

/* Arithmetic of the solve. */
enum {
	BFMX_PRECISION_FP64 = 0,  /* FP64 matrix, vectors and reductions */
	BFMX_PRECISION_MIXED = 1, /* FP32 matrix and search directions inside FP64
	                             refinement; FP64 solution, residual and dots */
};

/* Formulations (the stress argument of bfmx_system_assemble). */
enum {
	BFMX_PLANE_STRAIN = 0,
	BFMX_PLANE_STRESS = 1,
	BFMX_AXISYMMETRIC = 2,   /* r = x; the reference's axisymmetric formulas as
	                            written (libbfm/src/system.c:228-356, 539-622):
	                            symmetric, indefinite for common nu, so solved by
	                            BFMX_PRECOND_MINRES */
};

/* Condition kinds, numbered as bfm_condition_kind_t (libbfm/src/bfm/condition.h:5-14). */
enum {
	BFMX_DIRICHLET_X = 0,
	BFMX_DIRICHLET_Y = 1,
	BFMX_NEUMANN_X = 2,
	BFMX_NEUMANN_Y = 3,
	BFMX_NEUMANN_NORMAL = 4,
	BFMX_NEUMANN_TANGENT = 5,
	BFMX_DIRICHLET_NORMAL = 6,
	BFMX_DIRICHLET_TANGENT = 7,
};

typedef struct bfmx_system bfmx_system_t;

typedef struct {
	int32_t status;         /* BFMX_OK, BFMX_E_BREAKDOWN or BFMX_E_MAXITER */
	int32_t precond;
	int32_t precision;
	int32_t restarts;       /* true-residual checks that restarted the iteration */
	int64_t iterations;     /* CG iterations (sum over refinement sweeps) */
	int64_t n_dof;
	double rtol;
	double rel_residual;    /* ||b - A u||_2 / ||b||_2, recomputed in FP64 at the end */
	double solve_ms;        /* device time from the first solver kernel to convergence */
	double b_norm;
	int32_t fallback;       /* BFMX_PRECOND_LEARNED: 1 when block-Jacobi PCG finished the solve */
	int32_t fallback_iterations;
} bfmx_stats_t;

const char* bfmx_last_error(void);

/* Number of visible GPUs; 0 with BFMX_OK when none. */
int bfmx_device_count(int* count);

/* Uploads a planar mesh to the current device. kind is 3 (P1 triangles) or
 * 4 (Q1 quads, nodes counter-clockwise from (+1,+1) as in
 * libbfm/src/shape.c:19-26). coords holds 2 doubles per node, elems kind
 * node indices per element. With renumber != 0 the nodes are reordered by
 * reverse Cuthill-McKee on the node graph before anything is stored. */
int bfmx_system_create(bfmx_system_t** out, int kind, size_t n_nodes, const double* coords,
                       size_t n_elems, const size_t* elems, int renumber);
int bfmx_system_destroy(bfmx_system_t* s);

/* Edge list used by Neumann and normal/tangent Dirichlet conditions, in
 * mesh->edges order: edge_nodes holds the two node indices of each edge and
 * edge_boundary is 1 where the edge has a single adjacent element
 * (bfm_edge_t.elems[1] == -1). */
int bfmx_system_set_edges(bfmx_system_t* s, size_t n_edges, const size_t* edge_nodes,
                          const uint8_t* edge_boundary);

/* Assembles K and the body load on the device: stress is a formulation
 * (BFMX_PLANE_STRAIN, BFMX_PLANE_STRESS or BFMX_AXISYMMETRIC; the latter
 * replaces fill_axisymmetric_elem and the condition loop of
 * bfm_system_create_axisymmetric_strain, libbfm/src/system.c:228-356,
 * 539-622); Young's modulus E (or elem_E[e] per element when
 * elem_E is not NULL), Poisson ratio nu, density rho. The body force per unit
 * mass of node j is node_force[2j..2j+1] when node_force is not NULL, else
 * (gx, gy). Clears all boundary conditions. */
int bfmx_system_assemble(bfmx_system_t* s, int stress, double E, double nu, double rho,
                         const double* elem_E, const double* node_force, double gx, double gy);

/* Applies one boundary condition; node_flags has one byte per node (nonzero
 * = in the set). Conditions act in call order with libbfm's semantics: a
 * Dirichlet value overwrites the load of its DOF, a later Neumann load adds
 * to it, and the symmetric elimination lifts the prescribed values into the
 * other rows. */
int bfmx_system_apply_condition(bfmx_system_t* s, int kind, double value, const uint8_t* node_flags);

/* Solves A u = b. b is the assembled, condition-applied load when b_in is
 * NULL, else b_in (2 * n_nodes doubles). u_out (2 * n_nodes doubles, may be
 * NULL) receives the displacements. max_iter = 0 selects 20 * n_dof. stats
 * may be NULL. The solution also stays on the device for bfmx_system_get_u. */
int bfmx_system_solve(bfmx_system_t* s, const double* b_in, double* u_out, int precond,
                      int precision, double rtol, size_t max_iter, bfmx_stats_t* stats);

/* bfmx_system_solve without allocating: no reference counterpart; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). Point or block
 * Jacobi PCG in FP64 (precond BFMX_PRECOND_JACOBI or BFMX_PRECOND_BLOCK_JACOBI)
 * from u = 0, with the same kernels as bfmx_system_solve, launched directly
 * (no graph). Its solver buffers live in one fixed 64 MiB device array per
 * GPU, held until the process ends, so it takes systems of at most 408,880
 * nodes (larger ones are refused with BFMX_E_ARG and a message stating the
 * limit); calls that overlap on one GPU share that array and race. b_in,
 * u_out and stats as in bfmx_system_solve; the solution is not kept for
 * bfmx_system_get_u, and the system's own solver buffers are left as they
 * were. Pending boundary conditions are finalized first (host memory only). */
int bfmx_system_solve_noalloc(bfmx_system_t* s, const double* b_in, double* u_out, int precond, double rtol,
                              size_t max_iter, bfmx_stats_t* stats);

/* Sizes: nodes, stored 2x2 blocks, blocks including the sliced-ELL padding,
 * element colours. Any pointer may be NULL. */
int bfmx_system_info(bfmx_system_t* s, size_t* n_nodes, size_t* nnzb, size_t* nnzb_padded,
                     size_t* n_colors);

/* perm[old] = new node index (n_nodes values). */
int bfmx_system_node_perm(bfmx_system_t* s, int64_t* perm);

/* The matrix after boundary conditions as block-CSR in the internal node
 * order: rowptr (n_nodes + 1), col (nnzb), val (4 * nnzb, row-major 2x2
 * blocks). Any pointer may be NULL. */
int bfmx_system_get_bsr(bfmx_system_t* s, int64_t* rowptr, int32_t* col, double* val);

/* The condition-applied load vector, original DOF order. */
int bfmx_system_get_rhs(bfmx_system_t* s, double* b);

/* Last solution, original DOF order. */
int bfmx_system_get_u(bfmx_system_t* s, double* u);

/* Entry (i, j) of the condition-applied matrix in original DOF numbering;
 * 0 for entries outside the sparsity pattern. */
int bfmx_system_get_entry(bfmx_system_t* s, size_t i, size_t j, double* out);

/* Times reps back-to-back launches of the solver's SpMV kernel (the one the
 * FP64 or mixed CG iteration uses) with HIP events on the system's stream.
 * avg_ms receives the mean duration of one launch; bytes (may be NULL) the
 * kernel's algorithmic bytes per launch. */
int bfmx_system_time_spmv(bfmx_system_t* s, int precision, int reps, double* avg_ms, double* bytes);

/* Kernels that bfmx_system_time_kernel can time. */
enum {
	BFMX_KERNEL_CG_SPMV64 = 0,   /* FP64 CG SpMV (as bfmx_system_time_spmv, FP64) */
	BFMX_KERNEL_CG_SPMV32 = 1,   /* FP32-matrix CG SpMV (as bfmx_system_time_spmv, mixed) */
	BFMX_KERNEL_FCG_SPMV = 2,    /* flexible-CG SpMV: FP32 z, FP64 matrix and directions */
	BFMX_KERNEL_AMG_SMOOTH = 3,  /* the finest level's last post-smoothing step of the
	                                multigrid V-cycle: FP32 SpMV + block-Jacobi
	                                Chebyshev step + partial sums of x . q */
};

/* Times reps back-to-back launches of one solver kernel on the system's
 * stream with HIP events, on the current assembled system (the multigrid
 * kernel sets the hierarchy up first). avg_ms receives the mean duration of
 * one launch; bytes (may be NULL) the kernel's algorithmic bytes per launch.
 * No reference counterpart: the benchmark's roofline figure. */
int bfmx_system_time_kernel(bfmx_system_t* s, int kernel, int reps, double* avg_ms, double* bytes);

/* ---- learned preconditioner (no reference counterpart) ----------------
 * A 3-layer message-passing network over the node graph (hidden width 64,
 * BF16 layer weights and activations, FP32 accumulation) used as the
 * preconditioner of flexible CG, trained on the GPU on (residual, error)
 * pairs recorded from plain PCG runs. */
typedef struct bfmx_model bfmx_model_t;
typedef struct bfmx_trainset bfmx_trainset_t;

/* Number of FP32 parameters of a model. */
int bfmx_model_n_params(size_t* n);
/* New model on the current device (seeded): Glorot-uniform layers and a
 * small encoder, so the untrained preconditioner is block Jacobi plus a small
 * symmetric positive semi-definite term (exactly block Jacobi with a zero
 * encoder). */
int bfmx_model_create(bfmx_model_t** out, uint64_t seed);
int bfmx_model_destroy(bfmx_model_t* m);
int bfmx_model_get_params(bfmx_model_t* m, float* params);
int bfmx_model_set_params(bfmx_model_t* m, const float* params);

int bfmx_trainset_create(bfmx_trainset_t** out);
int bfmx_trainset_destroy(bfmx_trainset_t* ts);
/* Solves the system's current problem with Jacobi PCG to rtol (x*), then
 * reruns it recording up to max_pairs pairs (r_k, x* - x_k), one every
 * `every` iterations from k = 0 (every = 0: spread over the whole run). The matrix values and node features are
 * snapshotted; the system must outlive the set (its graph is shared). */
int bfmx_trainset_add(bfmx_trainset_t* ts, bfmx_system_t* s, int max_pairs, int every, double rtol,
                      int* n_added);
int bfmx_trainset_info(bfmx_trainset_t* ts, int entry, size_t* n_nodes, int* n_pairs);
/* An entry's data in the internal node order: feat (8 per node), r and e
 * (n_pairs x 2 n_nodes), s_den (n_pairs scales s, then n_pairs e^T A e). */
int bfmx_trainset_get(bfmx_trainset_t* ts, int entry, float* feat, float* r, float* e, float* s_den);

/* Prepares every entry's pairs for a model used with
 * BFMX_PRECOND_AMG_LEARNED: with V the FP32 multigrid V-cycle of the entry's
 * system (built here if needed) and D^-1 its block-Jacobi inverse (FP32
 * features), each error e becomes e - V(r) + D^-1 r, so that the loss below,
 * which scores z = D^-1 r + (network term), scores V(r) + (network term)
 * against the original e (the normalizers stay those of the original e).
 * Call once per set, after its last bfmx_trainset_add. */
int bfmx_trainset_rebase(bfmx_trainset_t* ts);

/* Loss (mean over pairs p0 .. p0 + batch - 1 of ||z - e||_A^2 / ||e||_A^2)
 * and its gradient with respect to the parameters (FP32), without an update. */
int bfmx_model_loss_grad(bfmx_model_t* m, bfmx_trainset_t* ts, int entry, int p0, int batch, double* loss,
                         float* grad);
/* steps Adam steps (lr; beta 0.9, 0.999) on random batches of one entry's
 * pairs; losses (may be NULL) receives each step's loss. */
int bfmx_model_train(bfmx_model_t* m, bfmx_trainset_t* ts, int steps, int batch, float lr, uint64_t seed,
                     double* losses);

/* Attaches the model used by BFMX_PRECOND_LEARNED (borrowed; NULL detaches). */
int bfmx_system_set_model(bfmx_system_t* s, bfmx_model_t* m);
/* z = preconditioner(r) once, both in the internal node order (2 per node). */
int bfmx_system_apply_model(bfmx_system_t* s, const double* r, double* z);

/* ---- batched solves (no reference counterpart) -------------------------
 * Many systems (uneven sizes) solved by Jacobi or block-Jacobi PCG, one
 * workgroup per system, in one persistent launch per size class (systems of
 * < 2048, < 8192, >= 8192 nodes), the launches captured in one HIP graph.
 * A batch holds either existing assembled systems (bfmx_batch_create) or
 * meshes that every launch assembles before solving
 * (bfmx_batch_create_planar). A configuration is a workgroup size x rows-per-thread pair:
 * 0 = 256 x 1, 1 = 256 x 2, 2 = 512 x 1, 3 = 512 x 2, 4 = 1024 x 1. A
 * system's solution does not depend on the rest of the batch (bitwise, for
 * the same configuration). The systems must outlive the batch and stay
 * unchanged while it is used. */
typedef struct bfmx_batch bfmx_batch_t;
enum { BFMX_BATCH_CLASSES = 3, BFMX_BATCH_CONFIGS = 5 };

int bfmx_batch_create(bfmx_batch_t** out, size_t n, bfmx_system_t* const* systems);
int bfmx_batch_destroy(bfmx_batch_t* b);
/* Times every configuration on up to `sample` systems of each class (0: all
 * of them, which measures throughput with the GPU full; a small sample
 * measures single-workgroup latency) for `iters` iterations; picks[c]
 * receives the fastest configuration of class c and times_ms (may be NULL)
 * the BFMX_BATCH_CLASSES x BFMX_BATCH_CONFIGS timings. The picks are
 * replayed by passing them to bfmx_batch_solve. */
int bfmx_batch_autotune(bfmx_batch_t* b, int precond, int sample, int iters, int32_t* picks, double* times_ms);
/* Solves every system from u = 0 to rtol (true FP64 residual); picks NULL =
 * the default configurations, max_iter 0 = 20 x the largest DOF count.
 * stats (may be NULL) receives one entry per system, total_ms the launch's
 * device time. Returns BFMX_OK when every system converged.
 * precond BFMX_PRECOND_JACOBI / BFMX_PRECOND_BLOCK_JACOBI: the persistent
 * launch above.
 * precond BFMX_PRECOND_AMG (multigrid mode): the batch is solved as one
 * block-diagonal system (the systems' rows one after another, a second copy
 * of their matrices) by flexible CG with the FP32 multigrid V-cycle, one
 * dense coarsest inverse per system. Each load is scaled to unit norm; every
 * system runs its own CG scalars (alpha, beta) when each has at least 256
 * nodes, else they are shared; the union's tolerance checks test every
 * system's own true residual, and the run ends when each meets rtol. The
 * stats report each system's true residual and the union's iteration count.
 * picks are ignored; the first such call builds the union and its hierarchy
 * on the host. A system's iterates depend on the rest of the batch through
 * the multigrid's smoother and prolongator parameters (one estimate of the
 * largest eigenvalue for the union), so results match a solo solve to the
 * tolerance, not bitwise. Axisymmetric systems are refused
 * (BFMX_E_UNSUPPORTED: indefinite). */
int bfmx_batch_solve(bfmx_batch_t* b, int precond, double rtol, size_t max_iter, const int32_t* picks,
                     bfmx_stats_t* stats, double* total_ms);
/* Solution of system `index`, in its original DOF order. */
int bfmx_batch_get_u(bfmx_batch_t* b, size_t index, double* u);

/* One planar problem of a batch built from meshes: the arguments of
 * bfmx_system_create, bfmx_system_set_edges, bfmx_system_assemble (with a
 * constant body force) and the bfmx_system_apply_condition calls of one
 * system, gathered in one record. */
typedef struct {
	int32_t kind;                  /* 3 (P1 triangles) or 4 (Q1 quads) */
	int32_t stress;                /* 1 plane stress, 0 plane strain */
	size_t n_nodes;
	const double* coords;          /* 2 per node */
	size_t n_elems;
	const size_t* elems;           /* kind per element */
	size_t n_edges;                /* 0: no edges (Neumann loads add nothing) */
	const size_t* edge_nodes;      /* 2 per edge */
	const uint8_t* edge_boundary;  /* 1 per edge, or NULL: every edge on the boundary */
	double E, nu, rho;
	const double* elem_E;          /* per element Young's modulus, or NULL */
	double gx, gy;                 /* body force per unit mass */
	size_t n_conditions;           /* applied in this order */
	const int32_t* cond_kind;
	const double* cond_value;
	const uint8_t* cond_flags;     /* n_conditions x n_nodes bytes */
} bfmx_instance_t;

/* Builds a batch from n meshes. The host prepares every instance as
 * bfmx_system_create would (node renumbering when renumber != 0, block
 * pattern, element colouring; instances are spread over up to 16 threads)
 * and packs all device arrays into a few allocations. Each bfmx_batch_solve
 * launch then assembles K and the load of every instance, applies its
 * conditions and solves it, one workgroup per instance. The input arrays
 * are copied; they need not outlive the call. */
int bfmx_batch_create_planar(bfmx_batch_t** out, size_t n, const bfmx_instance_t* inst, int renumber);

/* Instance `index` of a batch built from meshes, as a system object owned by
 * the batch (valid until bfmx_batch_destroy; not to be destroyed, assembled
 * or given conditions), available once a solve launch has assembled it:
 * the bfmx_system_* getters read the batch-assembled matrix and load, and
 * bfmx_system_solve solves it on its own. For a batch of existing systems,
 * returns that system. */
int bfmx_batch_system(bfmx_batch_t* b, size_t index, bfmx_system_t** s);

#ifdef __cplusplus
}
#endif
