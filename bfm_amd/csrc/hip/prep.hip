// Host-side mesh preparation: validation, node adjacency, reverse
// Cuthill-McKee renumbering, the sliced-ELL block pattern, element colouring
// and the uploads of the resulting topology.
//
// The renumbering follows the rules of libbfm's bfm_perm_rcm
// (libbfm/src/perm.c:117-331 in the reference tree) but on the node graph
// rather than the DOF graph of the condition-applied dense matrix: the start
// of each component is the last unvisited node of minimum degree, neighbours
// are taken in index order and enqueued in a stable sort by degree, and the
// visit order is reversed. Node-level ordering keeps the 2x2 blocks intact.

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <exception>
#include <mutex>
#include <numeric>
#include <thread>

#include "internal.h"

namespace bfmx {

namespace {

constexpr int kPrepThreads = 16;   // host threads of the preparation's parallel passes
thread_local int tl_threads = 0;   // this thread's limit (0: none), see PrepThreads

// f(lo, hi) over [0, n) split in contiguous ranges, one thread each (up to
// kPrepThreads, fewer for small n); exceptions are rethrown here.
template <typename F>
void parallel_ranges(int64_t n, const F& f, int64_t grain = 16384) {
	int hw = std::max(1, static_cast<int>(std::thread::hardware_concurrency()));
	if (tl_threads > 0) {
		hw = std::min(hw, tl_threads);
	}
	int const T = static_cast<int>(std::min<int64_t>({kPrepThreads, hw, std::max<int64_t>(1, n / grain)}));
	if (T <= 1) {
		f(int64_t(0), n);
		return;
	}
	std::vector<std::thread> th;
	std::vector<std::exception_ptr> err(T);
	for (int t = 0; t < T; t++) {
		th.emplace_back([&, t] {
			try {
				f(n * t / T, n * (t + 1) / T);
			} catch (...) {
				err[t] = std::current_exception();
			}
		});
	}
	for (auto& x : th) {
		x.join();
	}
	for (auto& e : err) {
		if (e) {
			std::rethrow_exception(e);
		}
	}
}

// Small rows: insertion sort (a row holds a few dozen entries at most).
void sort_small(int32_t* a, int64_t n) {
	for (int64_t i = 1; i < n; i++) {
		int32_t const v = a[i];
		int64_t j = i;
		for (; j > 0 && a[j - 1] > v; j--) {
			a[j] = a[j - 1];
		}
		a[j] = v;
	}
}

// Sorted, de-duplicated node adjacency (including the node itself) of the
// element graph, in the numbering of `elems`: the node-to-element incidence
// first, then each node's row from its elements' nodes, the rows of each
// thread's range into its own buffer, then copied into place.
void node_adjacency(int64_t n, int kind, int64_t n_elems, const int32_t* elems, std::vector<int64_t>& ptr,
                    std::vector<int32_t>& adj) {
	std::vector<int64_t> iptr(n + 1, 0);
	for (int64_t e = 0; e < n_elems; e++) {
		for (int a = 0; a < kind; a++) {
			iptr[elems[e * kind + a] + 1]++;
		}
	}
	for (int64_t i = 0; i < n; i++) {
		iptr[i + 1] += iptr[i];
	}
	std::vector<int32_t> inc(iptr[n]);
	{
		std::vector<int64_t> fill(iptr.begin(), iptr.end() - 1);
		for (int64_t e = 0; e < n_elems; e++) {
			for (int a = 0; a < kind; a++) {
				inc[fill[elems[e * kind + a]]++] = static_cast<int32_t>(e);
			}
		}
	}
	ptr.assign(n + 1, 0);
	std::mutex m;
	std::vector<std::pair<int64_t, std::vector<int32_t>>> parts;   // (first row, its rows' entries)
	parallel_ranges(n, [&](int64_t lo, int64_t hi) {
		std::vector<int32_t> buf;
		buf.reserve(static_cast<size_t>(hi - lo) * 10);
		for (int64_t i = lo; i < hi; i++) {
			size_t const at = buf.size();
			for (int64_t t = iptr[i]; t < iptr[i + 1]; t++) {
				const int32_t* en = elems + static_cast<int64_t>(inc[t]) * kind;
				buf.insert(buf.end(), en, en + kind);
			}
			if (buf.size() == at) {
				buf.push_back(static_cast<int32_t>(i));   // an isolated node keeps its diagonal
			}
			int32_t* c = buf.data() + at;
			int64_t const len = static_cast<int64_t>(buf.size() - at);
			sort_small(c, len);
			int64_t u = 1;
			for (int64_t k = 1; k < len; k++) {
				if (c[k] != c[u - 1]) {
					c[u++] = c[k];
				}
			}
			buf.resize(at + u);
			ptr[i + 1] = u;
		}
		std::lock_guard<std::mutex> lock(m);
		parts.emplace_back(lo, std::move(buf));
	});
	for (int64_t i = 0; i < n; i++) {
		ptr[i + 1] += ptr[i];
	}
	adj.resize(ptr[n]);
	parallel_ranges(
	    static_cast<int64_t>(parts.size()),
	    [&](int64_t lo, int64_t hi) {
		    for (int64_t q = lo; q < hi; q++) {
			    std::copy(parts[q].second.begin(), parts[q].second.end(), adj.begin() + ptr[parts[q].first]);
		    }
	    },
	    1);
}

// The adjacency of the renumbered graph (node perm[i] takes node i's row,
// its entries renumbered and sorted), rows in parallel.
void renumber_adjacency(int64_t n, const std::vector<int64_t>& ptr0, const std::vector<int32_t>& adj0,
                        const std::vector<int32_t>& perm, const std::vector<int32_t>& iperm,
                        std::vector<int64_t>& ptr, std::vector<int32_t>& adj) {
	ptr.assign(n + 1, 0);
	for (int64_t r = 0; r < n; r++) {
		ptr[r + 1] = ptr[r] + (ptr0[iperm[r] + 1] - ptr0[iperm[r]]);
	}
	adj.resize(ptr[n]);
	parallel_ranges(n, [&](int64_t lo, int64_t hi) {
		for (int64_t r = lo; r < hi; r++) {
			int64_t const o = iperm[r];
			int32_t* out = adj.data() + ptr[r];
			for (int64_t t = ptr0[o]; t < ptr0[o + 1]; t++) {
				*out++ = perm[adj0[t]];
			}
			sort_small(adj.data() + ptr[r], out - (adj.data() + ptr[r]));
		}
	});
}

void rcm(int64_t n, const std::vector<int64_t>& ptr, const std::vector<int32_t>& adj, std::vector<int32_t>& perm,
         std::vector<int32_t>& iperm) {
	std::vector<int32_t> deg(n);
	for (int64_t i = 0; i < n; i++) {
		deg[i] = static_cast<int32_t>(ptr[i + 1] - ptr[i]);
	}
	// Candidate start nodes: ascending degree, descending index, so the first
	// unvisited entry is the last node of minimum degree (a counting sort by
	// degree over the nodes in descending order).
	int32_t dmax = 0;
	for (int64_t i = 0; i < n; i++) {
		dmax = std::max(dmax, deg[i]);
	}
	std::vector<int64_t> dstart(static_cast<size_t>(dmax) + 2, 0);
	for (int64_t i = 0; i < n; i++) {
		dstart[deg[i] + 1]++;
	}
	for (int32_t d = 0; d <= dmax; d++) {
		dstart[d + 1] += dstart[d];
	}
	std::vector<int32_t> order(n);
	for (int64_t i = n - 1; i >= 0; i--) {
		order[dstart[deg[i]]++] = static_cast<int32_t>(i);
	}
	std::vector<uint8_t> visited(n, 0);
	std::vector<int32_t> queue(n);
	std::vector<int32_t> nb;
	iperm.assign(n, 0);
	int64_t qs = 0, qe = 0, p = 0, next_start = 0;
	while (p < n) {
		while (visited[order[next_start]]) {
			next_start++;
		}
		queue[qe++] = order[next_start];
		while (qs != qe) {
			int32_t const cur = queue[qs++];
			visited[cur] = 1;
			nb.clear();
			for (int64_t t = ptr[cur]; t < ptr[cur + 1]; t++) {
				int32_t const j = adj[t];
				if (!visited[j]) {
					visited[j] = 1;
					nb.push_back(j);
				}
			}
			for (size_t u = 1; u < nb.size(); u++) {   // stable insertion sort by degree
				int32_t const v = nb[u];
				size_t w = u;
				for (; w > 0 && deg[nb[w - 1]] > deg[v]; w--) {
					nb[w] = nb[w - 1];
				}
				nb[w] = v;
			}
			for (int32_t j : nb) {
				queue[qe++] = j;
			}
			iperm[n - 1 - p++] = cur;
		}
	}
	perm.assign(n, 0);
	for (int64_t i = 0; i < n; i++) {
		perm[iperm[i]] = static_cast<int32_t>(i);
	}
}

}  // namespace

PrepThreads::PrepThreads(int limit) : prev(tl_threads) {
	tl_threads = limit;
}

PrepThreads::~PrepThreads() {
	tl_threads = prev;
}

// Sliced-ELL layout of the system's renumbered adjacency: slice pointers,
// row lengths, diagonal blocks and the column of every block slot (padding
// slots repeat their row).
int sell_layout(bfmx_system* s, std::vector<int32_t>& col) {
	int64_t const n = s->n_nodes;
	s->n_slices = (n + kSlice - 1) / kSlice;
	s->slice_ptr.assign(s->n_slices + 1, 0);
	s->row_len.assign(n, 0);
	s->diag_blk.assign(n, 0);
	std::vector<int64_t> width(s->n_slices);
	parallel_ranges(s->n_slices, [&](int64_t lo, int64_t hi) {
		for (int64_t sl = lo; sl < hi; sl++) {
			int64_t kmax = 0;
			for (int64_t r = sl * kSlice; r < std::min(n, (sl + 1) * kSlice); r++) {
				kmax = std::max(kmax, s->adj_ptr[r + 1] - s->adj_ptr[r]);
			}
			width[sl] = kmax;
		}
	}, 1024);
	for (int64_t sl = 0; sl < s->n_slices; sl++) {
		int64_t const next = s->slice_ptr[sl] + width[sl] * kSlice;
		if (next >= (int64_t(1) << 31) / 4) {
			set_error("matrix too large for 32-bit block indices");
			return BFMX_E_ARG;
		}
		s->slice_ptr[sl + 1] = static_cast<int32_t>(next);
	}
	s->nnzb = s->adj_ptr[n];
	s->nnzb_pad = s->slice_ptr[s->n_slices];
	col.resize(s->nnzb_pad);
	parallel_ranges(s->n_slices, [&](int64_t lo, int64_t hi) {
		for (int64_t sl = lo; sl < hi; sl++) {
			int64_t const kmax = width[sl];
			for (int r = 0; r < kSlice; r++) {
				int64_t const row = sl * kSlice + r;
				int64_t const len = row < n ? s->adj_ptr[row + 1] - s->adj_ptr[row] : 0;
				if (row < n) {
					s->row_len[row] = static_cast<int32_t>(len);
				}
				for (int64_t k = 0; k < kmax; k++) {
					int64_t const blk = s->slice_ptr[sl] + k * kSlice + r;
					if (k < len) {
						int32_t const c = s->adj[s->adj_ptr[row] + k];
						col[blk] = c;
						if (c == row) {
							s->diag_blk[row] = static_cast<int32_t>(blk);
						}
					} else {
						col[blk] = row < n ? static_cast<int32_t>(row) : 0;
					}
				}
			}
		}
	}, 1024);
	return BFMX_OK;
}

int prepare_topology(bfmx_system* s, const double* coords, const size_t* elems, int renumber,
                     std::vector<int32_t>& col, std::vector<int32_t>& el_col) {
	int64_t const n = s->n_nodes;
	int64_t const ne = s->n_elems;
	int const kind = s->kind;
	if (n <= 0 || n >= (int64_t(1) << 30) || ne < 0 || ne >= (int64_t(1) << 31) / 16) {
		set_error("mesh size out of range: %lld nodes, %lld elements", (long long) n, (long long) ne);
		return BFMX_E_ARG;
	}
	// BFMX_SOLVE_TIMES=1: the preparation's phases on stderr
	auto t_last = std::chrono::steady_clock::now();
	auto mark = [&](const char* what) {
		if (phase_times_on()) {
			auto const now = std::chrono::steady_clock::now();
			std::fprintf(stderr, "prepare phase %-32s %8.3f ms\n", what,
			             std::chrono::duration<double, std::milli>(now - t_last).count());
			t_last = now;
		}
	};
	std::vector<int32_t> el(ne * kind);
	std::atomic<int64_t> bad_elem{-1}, bad_node{-1};
	parallel_ranges(ne * kind, [&](int64_t lo, int64_t hi) {
		for (int64_t t = lo; t < hi; t++) {
			if (elems[t] >= static_cast<size_t>(n)) {
				int64_t expect = -1;
				bad_elem.compare_exchange_strong(expect, t);
				return;
			}
			el[t] = static_cast<int32_t>(elems[t]);
		}
	});
	parallel_ranges(2 * n, [&](int64_t lo, int64_t hi) {
		for (int64_t i = lo; i < hi; i++) {
			if (!std::isfinite(coords[i])) {
				int64_t expect = -1;
				bad_node.compare_exchange_strong(expect, i / 2);
				return;
			}
		}
	});
	if (bad_elem >= 0) {
		int64_t t = 0;   // the first offending entry, as a sequential scan reports it
		while (elems[t] < static_cast<size_t>(n)) {
			t++;
		}
		set_error("element %lld references node %zu of %lld", (long long) (t / kind), elems[t], (long long) n);
		return BFMX_E_ARG;
	}
	if (bad_node >= 0) {
		int64_t i = 0;
		while (std::isfinite(coords[i])) {
			i++;
		}
		set_error("node %lld has a non-finite coordinate", (long long) (i / 2));
		return BFMX_E_ARG;
	}

	mark("validation");
	std::vector<int64_t> ptr0;
	std::vector<int32_t> adj0;
	node_adjacency(n, kind, ne, el.data(), ptr0, adj0);
	mark("node adjacency");
	if (renumber) {
		rcm(n, ptr0, adj0, s->perm, s->iperm);
		mark("RCM");
	} else {
		s->perm.resize(n);
		s->iperm.resize(n);
		std::iota(s->perm.begin(), s->perm.end(), 0);
		std::iota(s->iperm.begin(), s->iperm.end(), 0);
	}
	const std::vector<int32_t>& perm = s->perm;

	// renumbered coordinates and elements, renumbered adjacency
	s->coords.resize(2 * n);
	parallel_ranges(n, [&](int64_t lo, int64_t hi) {
		for (int64_t i = lo; i < hi; i++) {
			s->coords[2 * perm[i] + 0] = coords[2 * i + 0];
			s->coords[2 * perm[i] + 1] = coords[2 * i + 1];
		}
	});
	parallel_ranges(ne * kind, [&](int64_t lo, int64_t hi) {
		for (int64_t t = lo; t < hi; t++) {
			el[t] = perm[el[t]];
		}
	});
	if (renumber) {
		renumber_adjacency(n, ptr0, adj0, perm, s->iperm, s->adj_ptr, s->adj);
	} else {
		s->adj_ptr.swap(ptr0);
		s->adj.swap(adj0);
	}
	mark("renumbered coordinates, elements, adjacency");

	BFMX_CHECK(sell_layout(s, col));
	mark("sliced ELL layout");

	// Greedy colouring in order of each element's lowest node, so a colour's
	// consecutive elements touch nearby rows; 64 colours at most.
	std::vector<int32_t> emin(ne);
	std::vector<int64_t> bucket(n + 1, 0);
	for (int64_t e = 0; e < ne; e++) {
		int32_t m = el[e * kind];
		for (int a = 1; a < kind; a++) {
			m = std::min(m, el[e * kind + a]);
		}
		emin[e] = m;
		bucket[m + 1]++;
	}
	for (int64_t i = 0; i < n; i++) {
		bucket[i + 1] += bucket[i];
	}
	std::vector<int32_t> by_min(ne);
	for (int64_t e = 0; e < ne; e++) {
		by_min[bucket[emin[e]]++] = static_cast<int32_t>(e);
	}
	std::vector<uint64_t> mask(n, 0);
	std::vector<uint8_t> color(ne);
	int ncol = 0;
	for (int64_t t = 0; t < ne; t++) {
		int32_t const e = by_min[t];
		uint64_t forb = 0;
		for (int a = 0; a < kind; a++) {
			forb |= mask[el[e * kind + a]];
		}
		if (forb == ~uint64_t(0)) {
			set_error("element colouring needs more than 64 colours (node valence too high)");
			return BFMX_E_UNSUPPORTED;
		}
		int const c = __builtin_ctzll(~forb);
		color[e] = static_cast<uint8_t>(c);
		ncol = std::max(ncol, c + 1);
		for (int a = 0; a < kind; a++) {
			mask[el[e * kind + a]] |= uint64_t(1) << c;
		}
	}
	mark("element colouring");
	s->n_colors = ncol;
	s->color_ptr.assign(ncol + 1, 0);
	for (int64_t e = 0; e < ne; e++) {
		s->color_ptr[color[e] + 1]++;
	}
	for (int c = 0; c < ncol; c++) {
		s->color_ptr[c + 1] += s->color_ptr[c];
	}
	std::vector<int64_t> cfill(s->color_ptr.begin(), s->color_ptr.end() - 1);
	s->elem_order.assign(ne, 0);
	el_col.assign(ne * kind, 0);
	for (int64_t t = 0; t < ne; t++) {
		int32_t const e = by_min[t];
		int64_t const pos = cfill[color[e]]++;
		s->elem_order[pos] = e;
		std::memcpy(&el_col[pos * kind], &el[e * kind], kind * sizeof(int32_t));
	}
	mark("colour order");
	return BFMX_OK;
}

int build_topology(bfmx_system* s, const double* coords, const size_t* elems, int renumber) {
	std::vector<int32_t> col, el_col;
	BFMX_CHECK(prepare_topology(s, coords, elems, renumber, col, el_col));
	int64_t const n = s->n_nodes;
	int64_t const ne = s->n_elems;
	int const kind = s->kind;
	// uploads
	BFMX_CHECK(dalloc(&s->d_coords, 2 * n));
	BFMX_CHECK(dalloc(&s->d_elem, ne * kind));
	BFMX_CHECK(dalloc(&s->d_slots, ne * kind * kind));
	BFMX_CHECK(dalloc(&s->d_slice_ptr, s->n_slices + 1));
	BFMX_CHECK(dalloc(&s->d_row_len, n));
	BFMX_CHECK(dalloc(&s->d_diag_blk, n));
	BFMX_CHECK(dalloc(&s->d_col, s->nnzb_pad));
	BFMX_CHECK(dalloc(&s->d_val, 4 * s->nnzb_pad));
	BFMX_CHECK(dalloc(&s->d_b, 2 * n));
	BFMX_CHECK(dalloc(&s->d_bsolve, 2 * n));
	BFMX_CHECK(dalloc(&s->d_fixed, 2 * n));
	BFMX_CHECK(dalloc(&s->d_vfirst, 2 * n));
	BFMX_CHECK(dalloc(&s->d_flags, n));
	BFMX_CHECK(dalloc(&s->d_list, n));
	BFMX_CHECK(dalloc(&s->d_rows, n));
	hipStream_t const st = s->stream;
	BFMX_HIP(hipMemcpyAsync(s->d_coords, s->coords.data(), 2 * n * sizeof(double), hipMemcpyHostToDevice, st));
	BFMX_HIP(hipMemcpyAsync(s->d_elem, el_col.data(), el_col.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
	BFMX_HIP(hipMemcpyAsync(s->d_slice_ptr, s->slice_ptr.data(), s->slice_ptr.size() * sizeof(int32_t),
	                        hipMemcpyHostToDevice, st));
	BFMX_HIP(hipMemcpyAsync(s->d_row_len, s->row_len.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
	BFMX_HIP(hipMemcpyAsync(s->d_diag_blk, s->diag_blk.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
	BFMX_HIP(hipMemcpyAsync(s->d_col, col.data(), col.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
	// the pageable sources above must stay alive until the copies ran
	BFMX_HIP(hipStreamSynchronize(st));
	return BFMX_OK;
}

int prepare_edges(bfmx_system* s, size_t n_edges, const size_t* edge_nodes, const uint8_t* edge_boundary) {
	int64_t const n = s->n_nodes;
	s->n_edges = static_cast<int64_t>(n_edges);
	s->edge_nodes.resize(2 * n_edges);
	s->edge_bnd.resize(n_edges);
	for (size_t e = 0; e < n_edges; e++) {
		for (int k = 0; k < 2; k++) {
			size_t const v = edge_nodes[2 * e + k];
			if (v >= static_cast<size_t>(n)) {
				set_error("edge %zu references node %zu of %lld", e, v, (long long) n);
				return BFMX_E_ARG;
			}
			s->edge_nodes[2 * e + k] = s->perm[v];
		}
		s->edge_bnd[e] = edge_boundary ? (edge_boundary[e] ? 1 : 0) : 1;
	}
	// node -> incident edges, in edge order; an edge appears once per endpoint
	// occurrence (twice for a degenerate edge whose endpoints coincide).
	s->inc_ptr.assign(n + 1, 0);
	for (size_t e = 0; e < n_edges; e++) {
		s->inc_ptr[s->edge_nodes[2 * e] + 1]++;
		s->inc_ptr[s->edge_nodes[2 * e + 1] + 1]++;
	}
	for (int64_t i = 0; i < n; i++) {
		s->inc_ptr[i + 1] += s->inc_ptr[i];
	}
	s->inc.assign(s->inc_ptr[n], 0);
	std::vector<int64_t> fill(s->inc_ptr.begin(), s->inc_ptr.end() - 1);
	for (size_t e = 0; e < n_edges; e++) {
		s->inc[fill[s->edge_nodes[2 * e]]++] = static_cast<int32_t>(e);
		s->inc[fill[s->edge_nodes[2 * e + 1]]++] = static_cast<int32_t>(e);
	}
	return BFMX_OK;
}

int set_edges(bfmx_system* s, size_t n_edges, const size_t* edge_nodes, const uint8_t* edge_boundary) {
	BFMX_CHECK(prepare_edges(s, n_edges, edge_nodes, edge_boundary));
	int64_t const n = s->n_nodes;
	dfree(s->d_edge_nodes);
	dfree(s->d_edge_bnd);
	dfree(s->d_inc_ptr);
	dfree(s->d_inc);
	BFMX_CHECK(dalloc(&s->d_edge_nodes, 2 * n_edges));
	BFMX_CHECK(dalloc(&s->d_edge_bnd, n_edges));
	BFMX_CHECK(dalloc(&s->d_inc_ptr, n + 1));
	BFMX_CHECK(dalloc(&s->d_inc, s->inc.size()));
	hipStream_t const st = s->stream;
	if (n_edges) {
		BFMX_HIP(hipMemcpyAsync(s->d_edge_nodes, s->edge_nodes.data(), 2 * n_edges * sizeof(int32_t),
		                        hipMemcpyHostToDevice, st));
		BFMX_HIP(hipMemcpyAsync(s->d_edge_bnd, s->edge_bnd.data(), n_edges, hipMemcpyHostToDevice, st));
		BFMX_HIP(hipMemcpyAsync(s->d_inc, s->inc.data(), s->inc.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
	}
	BFMX_HIP(hipMemcpyAsync(s->d_inc_ptr, s->inc_ptr.data(), (n + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
	BFMX_HIP(hipStreamSynchronize(st));
	return BFMX_OK;
}

}  // namespace bfmx
