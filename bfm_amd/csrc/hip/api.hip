// extern "C" entry points of libbfmhip.so (declared in include/bfmx.h).

#include <cstring>
#include <exception>
#include <new>

#include "internal.h"

namespace bfmx {

namespace {
thread_local char g_err[1024] = "";
}

void set_error(const char* fmt, ...) {
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof g_err, fmt, ap);
	va_end(ap);
}

const char* last_error_text() {
	return g_err;
}

namespace {

int check(bfmx_system* s) {
	if (!s) {
		set_error("null system");
		return BFMX_E_ARG;
	}
	hipError_t const e = hipSetDevice(s->device);
	if (e != hipSuccess) {
		set_error("hipSetDevice(%d): %s", s->device, hipGetErrorString(e));
		return BFMX_E_HIP;
	}
	return BFMX_OK;
}

// Calls that rebuild a system's topology, matrix or conditions (or free it)
// are refused on the systems a batch built from meshes owns.
int check_owned(bfmx_system* s) {
	BFMX_CHECK(check(s));
	if (s->borrowed) {
		set_error("the system belongs to a batch built from meshes: it cannot be changed or destroyed");
		return BFMX_E_ARG;
	}
	return BFMX_OK;
}

void release(bfmx_system* s) {
	release_solver(s);
	dfree(s->d_coords);
	dfree(s->d_elem);
	dfree(s->d_slots);
	dfree(s->d_elemE);
	dfree(s->d_slice_ptr);
	dfree(s->d_row_len);
	dfree(s->d_diag_blk);
	dfree(s->d_col);
	dfree(s->d_val);
	dfree(s->d_b);
	dfree(s->d_bsolve);
	dfree(s->d_fixed);
	dfree(s->d_vfirst);
	dfree(s->d_flags);
	dfree(s->d_list);
	dfree(s->d_rows);
	dfree(s->d_edge_nodes);
	dfree(s->d_edge_bnd);
	dfree(s->d_inc_ptr);
	dfree(s->d_inc);
	if (s->stream) {
		(void) hipStreamDestroy(s->stream);
		s->stream = nullptr;
	}
}

// Exceptions (std::bad_alloc from host vectors) must not cross the C ABI.
template <typename F>
int guarded(F&& f) {
	try {
		return f();
	} catch (const std::bad_alloc&) {
		set_error("host allocation failed");
		return BFMX_E_NOMEM;
	} catch (const std::exception& ex) {
		set_error("%s", ex.what());
		return BFMX_E_ARG;
	}
}

}  // namespace
}  // namespace bfmx

using namespace bfmx;

#pragma GCC visibility push(default)
extern "C" {

const char* bfmx_last_error(void) {
	return g_err;
}

int bfmx_device_count(int* count) {
	if (!count) {
		set_error("null count");
		return BFMX_E_ARG;
	}
	int n = 0;
	hipError_t const e = hipGetDeviceCount(&n);
	if (e == hipErrorNoDevice || e == hipErrorInsufficientDriver) {
		*count = 0;
		return BFMX_OK;
	}
	if (e != hipSuccess) {
		*count = 0;
		set_error("hipGetDeviceCount: %s", hipGetErrorString(e));
		return BFMX_E_HIP;
	}
	*count = n;
	return BFMX_OK;
}

int bfmx_system_create(bfmx_system_t** out, int kind, size_t n_nodes, const double* coords, size_t n_elems,
                       const size_t* elems, int renumber) {
	return guarded([&]() -> int {
		if (!out || !coords || (n_elems && !elems)) {
			set_error("null argument");
			return BFMX_E_ARG;
		}
		*out = nullptr;
		if (kind != 3 && kind != 4) {
			set_error("element kind %d not supported (3 = P1 triangles, 4 = Q1 quads)", kind);
			return BFMX_E_UNSUPPORTED;
		}
		int ndev = 0;
		if (bfmx_device_count(&ndev) != BFMX_OK || ndev == 0) {
			set_error("no GPU visible to HIP");
			return BFMX_E_NODEV;
		}
		auto* s = new bfmx_system();
		BFMX_HIP(hipGetDevice(&s->device));
		s->kind = kind;
		s->n_nodes = static_cast<int64_t>(n_nodes);
		s->n_elems = static_cast<int64_t>(n_elems);
		hipError_t const e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
		if (e != hipSuccess) {
			set_error("hipStreamCreate: %s", hipGetErrorString(e));
			delete s;
			return BFMX_E_HIP;
		}
		int rc = build_topology(s, coords, elems, renumber);
		if (rc == BFMX_OK) {
			rc = compute_slots(s);
		}
		if (rc == BFMX_OK) {
			hipError_t const e2 = hipStreamSynchronize(s->stream);
			if (e2 != hipSuccess) {
				set_error("topology upload: %s", hipGetErrorString(e2));
				rc = BFMX_E_HIP;
			}
		}
		if (rc != BFMX_OK) {
			release(s);
			delete s;
			return rc;
		}
		*out = s;
		return BFMX_OK;
	});
}

int bfmx_system_destroy(bfmx_system_t* s) {
	if (!s) {
		return BFMX_OK;
	}
	if (s->borrowed) {
		set_error("the system belongs to a batch built from meshes: destroy the batch");
		return BFMX_E_ARG;
	}
	(void) hipSetDevice(s->device);
	if (s->stream) {
		(void) hipStreamSynchronize(s->stream);
	}
	release(s);
	delete s;
	return BFMX_OK;
}

int bfmx_system_set_edges(bfmx_system_t* s, size_t n_edges, const size_t* edge_nodes, const uint8_t* edge_boundary) {
	BFMX_CHECK(check_owned(s));
	if (n_edges && !edge_nodes) {
		set_error("null edge list");
		return BFMX_E_ARG;
	}
	return guarded([&] { return set_edges(s, n_edges, edge_nodes, edge_boundary); });
}

int bfmx_system_assemble(bfmx_system_t* s, int stress, double E, double nu, double rho, const double* elem_E,
                         const double* node_force, double gx, double gy) {
	BFMX_CHECK(check_owned(s));
	if (stress < BFMX_PLANE_STRAIN || stress > BFMX_AXISYMMETRIC) {
		set_error("formulation %d unknown (0 plane strain, 1 plane stress, 2 axisymmetric)", stress);
		return BFMX_E_ARG;
	}
	return guarded([&]() -> int {
		BFMX_CHECK(assemble(s, stress, E, nu, rho, elem_E, node_force, gx, gy));
		BFMX_HIP(hipStreamSynchronize(s->stream));
		return BFMX_OK;
	});
}

int bfmx_system_apply_condition(bfmx_system_t* s, int kind, double value, const uint8_t* node_flags) {
	BFMX_CHECK(check_owned(s));
	if (!node_flags) {
		set_error("null node flags");
		return BFMX_E_ARG;
	}
	return guarded([&] { return apply_condition(s, kind, value, node_flags); });
}

int bfmx_system_solve(bfmx_system_t* s, const double* b_in, double* u_out, int precond, int precision, double rtol,
                      size_t max_iter, bfmx_stats_t* stats) {
	BFMX_CHECK(check(s));
	if (!s->assembled) {
		set_error("solve before assemble");
		return BFMX_E_ARG;
	}
	return guarded([&]() -> int {
		BFMX_CHECK(finalize_conditions(s));
		int const rc = solve(s, b_in, precond, precision, rtol, max_iter, stats);
		if (u_out && (rc == BFMX_OK || rc == BFMX_E_MAXITER)) {
			int const rc2 = bfmx_system_get_u(s, u_out);
			if (rc2 != BFMX_OK) {
				return rc2;
			}
		}
		return rc;
	});
}

int bfmx_system_solve_noalloc(bfmx_system_t* s, const double* b_in, double* u_out, int precond, double rtol,
                              size_t max_iter, bfmx_stats_t* stats) {
	BFMX_CHECK(check(s));
	if (!s->assembled) {
		set_error("solve before assemble");
		return BFMX_E_ARG;
	}
	return guarded([&] { return solve_noalloc(s, b_in, u_out, precond, rtol, max_iter, stats); });
}

int bfmx_system_info(bfmx_system_t* s, size_t* n_nodes, size_t* nnzb, size_t* nnzb_padded, size_t* n_colors) {
	if (!s) {
		set_error("null system");
		return BFMX_E_ARG;
	}
	if (n_nodes) {
		*n_nodes = static_cast<size_t>(s->n_nodes);
	}
	if (nnzb) {
		*nnzb = static_cast<size_t>(s->nnzb);
	}
	if (nnzb_padded) {
		*nnzb_padded = static_cast<size_t>(s->nnzb_pad);
	}
	if (n_colors) {
		*n_colors = static_cast<size_t>(s->n_colors);
	}
	return BFMX_OK;
}

int bfmx_system_node_perm(bfmx_system_t* s, int64_t* perm) {
	if (!s || !perm) {
		set_error("null argument");
		return BFMX_E_ARG;
	}
	for (int64_t i = 0; i < s->n_nodes; i++) {
		perm[i] = s->perm[i];
	}
	return BFMX_OK;
}

int bfmx_system_get_bsr(bfmx_system_t* s, int64_t* rowptr, int32_t* col, double* val) {
	BFMX_CHECK(check(s));
	if (!s->assembled) {
		set_error("matrix not assembled");
		return BFMX_E_ARG;
	}
	return guarded([&]() -> int {
		BFMX_CHECK(finalize_conditions(s));
		int64_t const n = s->n_nodes;
		if (rowptr) {
			for (int64_t i = 0; i <= n; i++) {
				rowptr[i] = s->adj_ptr[i];
			}
		}
		if (col) {
			std::memcpy(col, s->adj.data(), s->adj.size() * sizeof(int32_t));
		}
		if (val) {
			std::vector<double> v(4 * s->nnzb_pad);
			BFMX_HIP(hipMemcpy(v.data(), s->d_val, v.size() * sizeof(double), hipMemcpyDeviceToHost));
			for (int64_t i = 0; i < n; i++) {
				int const r = static_cast<int>(i & 63);
				int64_t const base = s->slice_ptr[i >> 6] + r;
				for (int64_t k = 0; k < s->row_len[i]; k++) {
					int64_t const blk = base + k * kSlice;
					const double* top = v.data() + 4 * blk - 2 * r;
					double* o = val + 4 * (s->adj_ptr[i] + k);
					o[0] = top[0];
					o[1] = top[1];
					o[2] = top[2 * kSlice];
					o[3] = top[2 * kSlice + 1];
				}
			}
		}
		return BFMX_OK;
	});
}

int bfmx_system_get_rhs(bfmx_system_t* s, double* b) {
	BFMX_CHECK(check(s));
	if (!b || !s->assembled) {
		set_error(!b ? "null output" : "not assembled");
		return BFMX_E_ARG;
	}
	return guarded([&]() -> int {
		BFMX_CHECK(finalize_conditions(s));
		int64_t const n = s->n_nodes;
		std::vector<double> v(2 * n);
		BFMX_HIP(hipMemcpy(v.data(), s->d_b, v.size() * sizeof(double), hipMemcpyDeviceToHost));
		for (int64_t i = 0; i < n; i++) {
			b[2 * i] = v[2 * s->perm[i]];
			b[2 * i + 1] = v[2 * s->perm[i] + 1];
		}
		return BFMX_OK;
	});
}

int bfmx_system_get_u(bfmx_system_t* s, double* u) {
	BFMX_CHECK(check(s));
	if (!u || !s->has_u) {
		set_error(!u ? "null output" : "no solution available");
		return BFMX_E_ARG;
	}
	return guarded([&]() -> int {
		int64_t const n = s->n_nodes;
		std::vector<double> v(2 * n);
		BFMX_HIP(hipMemcpy(v.data(), s->d_x, v.size() * sizeof(double), hipMemcpyDeviceToHost));
		for (int64_t i = 0; i < n; i++) {
			u[2 * i] = v[2 * s->perm[i]];
			u[2 * i + 1] = v[2 * s->perm[i] + 1];
		}
		return BFMX_OK;
	});
}

int bfmx_system_get_entry(bfmx_system_t* s, size_t i, size_t j, double* out) {
	BFMX_CHECK(check(s));
	if (!out || !s->assembled) {
		set_error(!out ? "null output" : "not assembled");
		return BFMX_E_ARG;
	}
	size_t const m = 2 * static_cast<size_t>(s->n_nodes);
	if (i >= m || j >= m) {
		set_error("entry (%zu, %zu) outside a %zu x %zu matrix", i, j, m, m);
		return BFMX_E_ARG;
	}
	return guarded([&]() -> int {
		BFMX_CHECK(finalize_conditions(s));
		int32_t const ri = s->perm[i / 2], cj = s->perm[j / 2];
		*out = 0;
		for (int64_t k = 0; k < s->row_len[ri]; k++) {
			if (s->adj[s->adj_ptr[ri] + k] == cj) {
				int const r = ri & 63;
				int64_t const blk = s->slice_ptr[ri >> 6] + r + k * kSlice;
				int64_t const off = 4 * blk - 2 * r + (i % 2) * 2 * kSlice + (j % 2);
				BFMX_HIP(hipMemcpy(out, s->d_val + off, sizeof(double), hipMemcpyDeviceToHost));
				break;
			}
		}
		return BFMX_OK;
	});
}

int bfmx_system_time_spmv(bfmx_system_t* s, int precision, int reps, double* avg_ms, double* bytes) {
	BFMX_CHECK(check(s));
	if (!avg_ms || reps <= 0 || !s->assembled) {
		set_error("bad arguments to time_spmv");
		return BFMX_E_ARG;
	}
	return guarded([&] { return time_spmv(s, precision, reps, avg_ms, bytes); });
}

int bfmx_system_time_kernel(bfmx_system_t* s, int kernel, int reps, double* avg_ms, double* bytes) {
	BFMX_CHECK(check(s));
	if (!avg_ms || reps <= 0 || !s->assembled) {
		set_error("bad arguments to time_kernel");
		return BFMX_E_ARG;
	}
	if (kernel < BFMX_KERNEL_CG_SPMV64 || kernel > BFMX_KERNEL_AMG_SMOOTH) {
		set_error("unknown kernel %d", kernel);
		return BFMX_E_UNSUPPORTED;
	}
	return guarded([&] { return time_kernel(s, kernel, reps, avg_ms, bytes); });
}

// ---- learned preconditioner

int bfmx_model_n_params(size_t* n) {
	if (!n) {
		set_error("null argument");
		return BFMX_E_ARG;
	}
	*n = static_cast<size_t>(mpnn_n_params());
	return BFMX_OK;
}

int bfmx_model_create(bfmx_model_t** out, uint64_t seed) {
	if (!out) {
		set_error("null argument");
		return BFMX_E_ARG;
	}
	int count = 0;
	if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
		set_error("no GPU visible");
		return BFMX_E_NODEV;
	}
	return guarded([&] { return model_create(out, seed); });
}

int bfmx_model_destroy(bfmx_model_t* m) {
	model_destroy(m);
	return BFMX_OK;
}

int bfmx_model_get_params(bfmx_model_t* m, float* params) {
	if (!m || !params) {
		set_error("null argument");
		return BFMX_E_ARG;
	}
	return guarded([&] { return model_get_params(m, params); });
}

int bfmx_model_set_params(bfmx_model_t* m, const float* params) {
	if (!m || !params) {
		set_error("null argument");
		return BFMX_E_ARG;
	}
	return guarded([&] { return model_set_params(m, params); });
}

int bfmx_trainset_create(bfmx_trainset_t** out) {
	if (!out) {
		set_error("null argument");
		return BFMX_E_ARG;
	}
	*out = new (std::nothrow) bfmx_trainset();
	if (!*out) {
		set_error("out of host memory");
		return BFMX_E_NOMEM;
	}
	return BFMX_OK;
}

int bfmx_trainset_destroy(bfmx_trainset_t* ts) {
	trainset_destroy(ts);
	return BFMX_OK;
}

int bfmx_trainset_add(bfmx_trainset_t* ts, bfmx_system_t* s, int max_pairs, int every, double rtol, int* n_added) {
	BFMX_CHECK(check(s));
	if (!ts || max_pairs < 1 || every < 0 || !(rtol > 0)) {
		set_error("bad arguments to trainset_add");
		return BFMX_E_ARG;
	}
	if (!s->assembled) {
		set_error("trainset_add before assemble");
		return BFMX_E_ARG;
	}
	return guarded([&]() -> int {
		size_t const before = ts->entries.size();
		BFMX_CHECK(trainset_add(ts, s, max_pairs, every, rtol));
		if (n_added) {
			*n_added = ts->entries.size() > before ? ts->entries.back().P : 0;
		}
		return BFMX_OK;
	});
}

int bfmx_trainset_info(bfmx_trainset_t* ts, int entry, size_t* n_nodes, int* n_pairs) {
	if (!ts || entry < 0 || entry >= static_cast<int>(ts->entries.size())) {
		set_error("no training entry %d", entry);
		return BFMX_E_ARG;
	}
	if (n_nodes) {
		*n_nodes = static_cast<size_t>(ts->entries[entry].n);
	}
	if (n_pairs) {
		*n_pairs = ts->entries[entry].P;
	}
	return BFMX_OK;
}

int bfmx_trainset_get(bfmx_trainset_t* ts, int entry, float* feat, float* r, float* e, float* s_den) {
	if (!ts) {
		set_error("null training set");
		return BFMX_E_ARG;
	}
	return guarded([&] { return trainset_get(ts, entry, feat, r, e, s_den); });
}

int bfmx_trainset_rebase(bfmx_trainset_t* ts) {
	if (!ts) {
		set_error("null training set");
		return BFMX_E_ARG;
	}
	return guarded([&] { return trainset_rebase(ts); });
}

int bfmx_model_loss_grad(bfmx_model_t* m, bfmx_trainset_t* ts, int entry, int p0, int batch, double* loss,
                         float* grad) {
	if (!m || !ts || entry < 0 || entry >= static_cast<int>(ts->entries.size())) {
		set_error("bad model, training set or entry");
		return BFMX_E_ARG;
	}
	if (batch < 1 || p0 < 0 || p0 + batch > ts->entries[entry].P) {
		set_error("pairs %d .. %d outside the entry's %d", p0, p0 + batch - 1, ts->entries[entry].P);
		return BFMX_E_ARG;
	}
	return guarded([&]() -> int {
		BFMX_CHECK(model_loss_grad(m, ts, entry, p0, batch));
		BFMX_HIP(hipStreamSynchronize(m->stream));
		if (loss) {
			BFMX_HIP(hipMemcpy(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost));
		}
		if (grad) {
			BFMX_HIP(hipMemcpy(grad, m->d_grad, mpnn_n_params() * sizeof(float), hipMemcpyDeviceToHost));
		}
		return BFMX_OK;
	});
}

int bfmx_model_train(bfmx_model_t* m, bfmx_trainset_t* ts, int steps, int batch, float lr, uint64_t seed,
                     double* losses) {
	if (!m || !ts || steps < 0 || batch < 1 || !(lr > 0)) {
		set_error("bad arguments to model_train");
		return BFMX_E_ARG;
	}
	return guarded([&] { return model_train(m, ts, steps, batch, lr, seed, losses); });
}

int bfmx_system_set_model(bfmx_system_t* s, bfmx_model_t* m) {
	BFMX_CHECK(check(s));
	if (m != s->model) {
		// the captured learned chunks read the old model's parameters, which
		// its owner may free once it is detached
		return guarded([&] {
			drop_graphs(s, BFMX_PRECOND_LEARNED);
			drop_graphs(s, BFMX_PRECOND_AMG_LEARNED);
			s->model = m;
			return BFMX_OK;
		});
	}
	return BFMX_OK;
}

int bfmx_system_apply_model(bfmx_system_t* s, const double* r, double* z) {
	BFMX_CHECK(check(s));
	if (!r || !z || !s->assembled) {
		set_error("bad arguments to apply_model");
		return BFMX_E_ARG;
	}
	return guarded([&]() -> int {
		int64_t const n = s->n_nodes;
		hipStream_t const st = s->stream;
		BFMX_CHECK(finalize_conditions(s));
		BFMX_CHECK(ensure_workspace_public(s));
		BFMX_CHECK(mpnn_prepare(s));
		BFMX_CHECK(ensure_f32_public(s));
		BFMX_CHECK(mpnn_setup_numeric(s));
		// S: running, r . r of the given residual
		double rr = 0;
		for (int64_t k = 0; k < 2 * n; k++) {
			rr += r[k] * r[k];
		}
		Scalars init{};
		init.rr = rr;
		BFMX_HIP(hipMemcpyAsync(s->d_scal, &init, sizeof(Scalars), hipMemcpyHostToDevice, st));
		BFMX_HIP(hipMemcpyAsync(s->d_r, r, 2 * n * sizeof(double), hipMemcpyHostToDevice, st));
		BFMX_HIP(hipMemsetAsync(s->d_q, 0, 2 * n * sizeof(double), st));
		int nw = 0;
		precond_block_jacobi(s);
		double* zd = mpnn_apply(s, s->d_r, s->d_q, s->d_partial, s->d_scal, &nw);
		BFMX_HIP(hipGetLastError());
		BFMX_HIP(hipMemcpyAsync(z, zd, 2 * n * sizeof(double), hipMemcpyDeviceToHost, st));
		BFMX_HIP(hipStreamSynchronize(st));
		return BFMX_OK;
	});
}

// ---- batched solves

int bfmx_batch_create(bfmx_batch_t** out, size_t n, bfmx_system_t* const* systems) {
	if (!out || !systems || n == 0) {
		set_error("bad arguments to batch_create");
		return BFMX_E_ARG;
	}
	BFMX_CHECK(check(systems[0]));
	return guarded([&] { return batch_create(out, n, systems); });
}

int bfmx_batch_create_planar(bfmx_batch_t** out, size_t n, const bfmx_instance_t* inst, int renumber) {
	if (!out || !inst || n == 0 || n > (size_t(1) << 30)) {
		set_error("bad arguments to batch_create_planar");
		return BFMX_E_ARG;
	}
	*out = nullptr;
	int ndev = 0;
	if (bfmx_device_count(&ndev) != BFMX_OK || ndev == 0) {
		set_error("no GPU visible to HIP");
		return BFMX_E_NODEV;
	}
	return guarded([&] { return batch_create_planar(out, n, inst, renumber); });
}

int bfmx_batch_system(bfmx_batch_t* b, size_t index, bfmx_system_t** s) {
	if (!b || !s) {
		set_error("null argument");
		return BFMX_E_ARG;
	}
	return guarded([&] { return batch_system(b, index, s); });
}

int bfmx_batch_destroy(bfmx_batch_t* b) {
	batch_destroy(b);
	return BFMX_OK;
}

int bfmx_batch_autotune(bfmx_batch_t* b, int precond, int sample, int iters, int32_t* picks, double* times_ms) {
	if (!b || !picks || sample < 0 || iters < 1) {
		set_error("bad arguments to batch_autotune");
		return BFMX_E_ARG;
	}
	return guarded([&] { return batch_autotune(b, precond, sample, iters, picks, times_ms); });
}

int bfmx_batch_solve(bfmx_batch_t* b, int precond, double rtol, size_t max_iter, const int32_t* picks,
                     bfmx_stats_t* stats, double* total_ms) {
	if (!b || !(rtol > 0)) {
		set_error("bad arguments to batch_solve");
		return BFMX_E_ARG;
	}
	return guarded(
	    [&] { return batch_solve(b, precond, rtol, static_cast<int64_t>(max_iter), picks, stats, total_ms); });
}

int bfmx_batch_get_u(bfmx_batch_t* b, size_t index, double* u) {
	if (!b || !u) {
		set_error("null argument");
		return BFMX_E_ARG;
	}
	return guarded([&] { return batch_get_u(b, index, u); });
}

}  // extern "C"
#pragma GCC visibility pop
