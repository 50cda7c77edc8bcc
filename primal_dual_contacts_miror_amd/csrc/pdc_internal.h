// pdc_internal.h — context, device buffers and launch helpers of libpdc.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pdc.h"
#include "pdc_math.h"

namespace pdc {

constexpr int kBlock = 256;            // threads per workgroup of the 1-D kernels
constexpr int kMaxPartialBlocks = 4096;  // workgroups whose partial sums a reduction reads
constexpr int kSortCap = 2048;         // raw sample hits one pair sorts in LDS (more: in global memory)
constexpr int kCellStride = 8;         // body b's sample cells start at floor(soff[b] / 8)

struct HipError : std::runtime_error {
	int code;
	HipError(const std::string& m, int c) : std::runtime_error(m), code(c) {}
};

#define PDC_HIP(call)                                                                   \
	do {                                                                                \
		hipError_t e_ = (call);                                                         \
		if (e_ != hipSuccess)                                                           \
			throw ::pdc::HipError(std::string(#call) + ": " + hipGetErrorString(e_),    \
			                      PDC_ERR_DEVICE);                                      \
	} while (0)

#define PDC_CHECK_LAUNCH() PDC_HIP(hipGetLastError())

inline int grid_for(int64_t n, int block = kBlock) {
	int64_t g = (n + block - 1) / block;
	return (int)(g < 1 ? 1 : g);
}

// Device memory handed out by a bump allocator over a fixed region (the
// library's static scratch array, used by the entry point that allocates
// nothing).  While an arena is active on the calling thread, every buffer
// that grows takes its memory from the arena instead of hipMalloc, and
// `restore` puts each such buffer back as it was before the call.
struct Arena {
	char* base = nullptr;
	size_t cap = 0, used = 0;
	struct Undo {
		void** pp;
		size_t* pcap;
		void* old_p;
		size_t old_cap;
	};
	std::vector<Undo> undo;
	bool owns(const void* p) const {
		return p != nullptr && (const char*)p >= base && (const char*)p < base + cap;
	}
	void note(void** pp, size_t* pcap) {
		for (const Undo& u : undo)
			if (u.pp == pp) return;
		undo.push_back(Undo{pp, pcap, *pp, *pcap});
	}
	void* take(size_t bytes, void** pp, size_t* pcap) {
		note(pp, pcap);
		const size_t b = (bytes + 255) & ~(size_t)255;
		if (used + b > cap)
			throw HipError("the call needs more device scratch than the " + std::to_string(cap >> 20) +
			                   " MiB of the no-allocation entry point",
			               PDC_ERR_CAPACITY);
		void* p = base + used;
		used += b;
		return p;
	}
	void restore() {
		for (auto it = undo.rbegin(); it != undo.rend(); ++it) {
			*it->pp = it->old_p;
			*it->pcap = it->old_cap;
		}
		undo.clear();
		used = 0;
	}
};
extern thread_local Arena* t_arena;  // the calling thread's active arena, or null
// Device bytes held by the buffers of the context the calling thread is
// working on (set by every entry point), or null.
extern thread_local int64_t* t_dev_bytes;
inline void count_bytes(int64_t d) {
	if (t_dev_bytes) *t_dev_bytes += d;
}

// Growable device buffer (reallocates only when the capacity grows).
template <class T>
struct DBuf {
	T* p = nullptr;
	size_t cap = 0;
	void ensure(size_t n) {
		if (t_arena) {
			if (t_arena->owns(p) && n <= cap) return;
			void* q = t_arena->take(std::max<size_t>(n, 1) * sizeof(T), reinterpret_cast<void**>(&p), &cap);
			p = static_cast<T*>(q);
			cap = n;
			return;
		}
		if (n <= cap) return;
		size_t nc = cap ? cap : 256;
		while (nc < n) nc = nc + nc / 2 + 1;
		if (p) {
			PDC_HIP(hipFree(p));
			count_bytes(-(int64_t)(cap * sizeof(T)));
		}
		p = nullptr;
		cap = 0;
		PDC_HIP(hipMalloc(&p, nc * sizeof(T)));
		count_bytes((int64_t)(nc * sizeof(T)));
		cap = nc;
	}
	// Grows to at least n elements (exactly n with `exact`) keeping the first
	// `keep` elements (copied on `s`, which is synchronised).
	void grow_keep(size_t n, size_t keep, hipStream_t s, bool exact = false) {
		if (n <= cap) return;
		if (t_arena) throw HipError("a body table cannot grow inside the no-allocation call", PDC_ERR_STATE);
		size_t nc = exact ? n : (cap ? cap : 256);
		while (nc < n) nc = nc + nc / 2 + 1;
		T* q = nullptr;
		PDC_HIP(hipMalloc(&q, nc * sizeof(T)));
		count_bytes((int64_t)(nc * sizeof(T)));
		if (p && keep) {
			PDC_HIP(hipMemcpyAsync(q, p, std::min(keep, cap) * sizeof(T), hipMemcpyDeviceToDevice, s));
			PDC_HIP(hipStreamSynchronize(s));
		}
		if (p) {
			PDC_HIP(hipFree(p));
			count_bytes(-(int64_t)(cap * sizeof(T)));
		}
		p = q;
		cap = nc;
	}
	void release() {
		if (t_arena) {  // the buffer is put back by the arena's restore
			t_arena->note(reinterpret_cast<void**>(&p), &cap);
			p = nullptr;
			cap = 0;
			return;
		}
		if (p) {
			(void)hipFree(p);
			count_bytes(-(int64_t)(cap * sizeof(T)));
		}
		p = nullptr;
		cap = 0;
	}
	operator T*() const { return p; }
};

constexpr int kPcgRowsPerWG = 8;  // PCG block rows (wavefronts) per workgroup

// Work tile: a run of consecutive elements (contacts, bodies, pairs or PCG
// block rows) of one scene, processed by one workgroup.  Tiles never cross a
// scene boundary, so per-workgroup partial sums are per-scene sums.
struct Tile {
	int scene;   // -1 for padding tiles
	int first;   // first element
	int count;   // elements in the tile
};

// PCG workgroup record, 64 bytes, read at the start of every PCG launch: the
// tile's block rows, its scene's tile range, the row pointers of its rows and
// the tile's own copy of the scene's PCG state.  Every workgroup of a scene
// takes the same decisions, so each one keeps and updates its own copy and no
// workgroup reads another's record.
struct PcgTile {
	int scene;    // -1 for padding tiles
	int first;    // first block row (local body numbering)
	int count;    // block rows (<= kPcgRowsPerWG)
	int t0, W;    // the scene's tiles [t0, t0 + W)
	int done;     // the scene's PCG has converged, or the scene is not solving
	int it;       // iterations completed
	int pad[kPcgRowsPerWG + 1];
};
static_assert(sizeof(PcgTile) == 64, "one 64-byte record per PCG workgroup");

// Per-scene state of the solve, advanced on the device (the Newton loop's
// decisions are taken per scene by the kernels) and read by the host once per
// Newton iteration and at the end of the solve.
struct SceneSc {
	double char_mass, char_speed;
	double gap, muc;
	double ru2, wrf2, rl2, ac2;
	double err, sErr, sacErr, eps, exitErr;
	double t_inErr, t_sacIn, t_gErr;
	double Hu;           // eps = Hu err^0.333335
	double ls_step;      // accepted step of this iteration's line search (0: none)
	int run;             // the scene's Newton loop continues
	int it;              // completed Newton iterations
	int ls_state;        // LS_* below
	int gd_count;        // line searches that fell back to gradient descent
	int trials;          // line-search trial points evaluated
	int pcg_total;       // PCG iterations, summed over Newton iterations
	int pcg_hits;        // PCG solves that reached the iteration cap
	int nonfinite;       // Newton iterations with a non-finite error or PCG result
};
enum { LS_PENDING = 0, LS_ACCEPTED = 1, LS_GD_PENDING = 2, LS_GD_DECIDED = 3 };

// Per-scene control written by the host once per solve.
struct SceneCtl {
	int act;     // the scene has contacts and takes part in the solve
	int pad;
	double mu;   // the scene's friction coefficient
};

// Per-scene PCG outcome (written by the scene's first tile; read on the
// device by k_newton_end, and by the host after each group of PCG chunks
// when it waits for the stream) and `run`, the scene's Newton state.
struct PcgScene {
	int iters;   // iterations of the last solve
	int done;    // converged flag (also set when the solve broke down)
	int run;     // the scene's Newton loop continues (copied by k_err_final)
	int bad;     // the solve met a non-finite residual: its increment is taken as zero
};

// Progress of a scene's Newton loop and PCG, written by the kernels into
// pinned host memory while the stream runs and read by the host without
// waiting for the stream (the host sizes the PCG launches it queues from
// it).  Each 64-bit field is written with one store, so the host never sees
// half of one.
struct PcgProgress {
	long long runtag;   // (Newton iteration << 32) | run, written by k_err_final
	long long outcome;  // (1 << 32) | iterations once the PCG converged; 0 before (reset by the host)
	double rel2;        // |r|^2 / |b|^2 at the start of the latest active iteration
	int it;             // PCG iterations started in the current solve
	int pad;
};

// Detection counters (one device struct).
struct Scalars {
	int n_pairs;         // broadphase candidate pairs (may exceed the pair capacity)
	int n_stage;         // staging slots the pairs' culled samples need (may exceed its capacity)
	int n_contacts;      // contacts after merge
	int n_raw;           // raw sample hits
	int ties;
	int n_big;           // pairs with more than kSortCap hits (merged in global memory)
	int pad[2];
	unsigned long long all_samples;   // samples of the emitting sides of every pair
	unsigned long long read_samples;  // samples in the cells the cull keeps
	unsigned long long brick_nodes;   // grid nodes under the partner's box, per pair and side
	unsigned long long ingrid_samples;  // tested samples inside the partner's grid box
};

// Counts of the last contact detection (pdc_get_detection_stats).
struct DetectionStats {
	int64_t pairs = 0, raw_hits = 0, ties = 0;
	int64_t samples_all = 0, samples_read = 0, brick_nodes = 0, large_pairs = 0, samples_in_grid = 0;
};

// Body-frame sample cell: the box of its FP32 samples and their range
// [beg, end) relative to the body's first sample.
struct SampleCell {
	float lo[3], hi[3];
	int beg, end;
};

// A triangle mesh registered with pdc_register_mesh.
struct MeshShape {
	std::vector<double> verts;  // 3 per vertex, body frame (the surface samples)
	std::vector<double> recs;   // kTriRec per triangle of nonzero area (the distance)
	double ext[3] = {0, 0, 0};  // half extents of the grid box before padding
	DBuf<double> d_recs;
	DBuf<float> grid;           // its SDF nodes at grid_n / grid_pad (built on first use)
	int grid_n = 0;
	double grid_pad = -1.0;
};

// Size class of a scene of n bodies (index into pdc_shapes' per-class arrays).
inline int size_class(int n) { return n < 64 ? 0 : n < 256 ? 1 : n < 1024 ? 2 : 3; }

struct PcgGraph {
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	int nblk = 0;
	int chunk = 0;
	unsigned char args[512] = {};  // the kernel arguments the chunk was recorded with
};

struct KernelTimer {
	int which = PDC_KT_NONE;
	double total_ms = 0;
	int64_t launches = 0;
	double bytes = 0;              // algorithmic bytes of the timed launches
	int64_t pending_contact_slots = 0;
	std::vector<hipEvent_t> pool;  // start/stop pairs awaiting read-back
	std::vector<int> pair_launches;  // launches each pair spans (one per used pair)
	size_t used = 0;
};

}  // namespace pdc

struct pdc_context {
	int device = 0;
	hipStream_t stream = nullptr;
	pdc_params p{};
	std::string err;
	int cus = 0;

	// ---- bodies (global numbering) ----
	int N = 0;
	int64_t n_samples = 0;
	pdc::DBuf<int> b_shape, b_static, b_emit;
	pdc::DBuf<double> b_dims, b_mass, b_I0;
	pdc::DBuf<double> b_pos, b_quat, b_vel, b_ang, b_acc;  // acc: 6 per body
	pdc::DBuf<double> b_R, b_aabb;                         // 9 and 6 per body
	pdc::DBuf<pdc::GridDesc> b_grid;
	pdc::DBuf<float> grid_pool;
	pdc::DBuf<int> b_soff;                                 // N+1 sample offsets
	pdc::DBuf<float> samples;                              // 3 per sample (FP32, body frame, by cell)
	pdc::DBuf<pdc::SampleCell> cells;                      // body b's at floor(soff[b] / kCellStride)
	pdc::DBuf<int> b_ncell;                                // cells per body
	std::vector<std::unique_ptr<pdc::MeshShape>> meshes;   // registered meshes, by id
	std::vector<int> h_soff;                               // host copy of sample offsets
	std::vector<int> h_static, h_emit;                     // host copies of the flags
	int64_t dev_bytes = 0;                                 // device bytes held by the context's buffers

	// ---- broadphase / pairs ----
	pdc::DBuf<int> bp_cnt, bp_off;                          // per body a: candidates b > a
	pdc::DBuf<int> pr_a, pr_b;                              // candidate pairs
	pdc::DBuf<int> pr_soff;                                 // staging offsets (P+1)
	pdc::DBuf<int> pr_nh;                                   // raw hits per pair
	pdc::DBuf<int> pr_nc, pr_coff;                          // merged contacts per pair, offsets (P+1)
	pdc::DBuf<int> pr_tmp;
	pdc::DBuf<double> stage_raw;                            // 7 per slot: pos, n, depth
	pdc::DBuf<double> stage_m;                              // 7 per slot: merged
	pdc::DBuf<int> stage_idx;                               // 2 per slot: sort indices of large pairs
	pdc::DBuf<int> pr_big;                                  // pairs with more than kSortCap hits
	pdc::DBuf<int> pr_st;                                   // per pair detection counters
	int pr_cap = 0;                                         // pair capacity
	int64_t stage_cap = 0;                                  // staging slots

	// ---- contacts (SoA) ----
	int C = 0;
	pdc::DBuf<int> c_id0, c_id1;
	pdc::DBuf<double> c_pos;    // [3][C]
	pdc::DBuf<double> c_depth;  // [C]
	pdc::DBuf<double> c_T;      // [C][36]: (side*18 + col*6 + row), one contact's 36 values together
	pdc::DBuf<double> c_lf;     // [3][C]
	int c_stride = 0;           // SoA row stride (the contact capacity)
	int n_pairs = 0, n_raw = 0, n_ties = 0;
	pdc::DetectionStats det;
	double tie_band = 0;

	// ---- topology (local numbering) ----
	int NL = 0, NP = 0, NNZB = 0;
	pdc::DBuf<int> t_key_flag, t_pid;                       // per contact
	pdc::DBuf<int> ap_a, ap_b, ap_start;                    // active pairs (global ids), start (NP+1)
	pdc::DBuf<int> g_flag, g2l, l2g;                        // per global body
	pdc::DBuf<int> li0, li1;                                // per contact local ids
	pdc::DBuf<int> adj_cnt, adj_ptr, adj_cur, adj_nb, adj_pair, adj_pos, adj_blk;
	pdc::DBuf<int> adj_nb_u, adj_pair_u, adj_row_u, pair_adj;
	pdc::DBuf<int> inc_cnt, inc_ptr, inc, c_pid;
	pdc::DBuf<int> a_cnt, a_ptr, a_diag;
	int pcg_last_iters = 0;
	pdc::PcgGraph pcg_graph[4];                               // recorded chunks (FP32, FP64; x2 forms)
	pdc::PcgGraph phase_graph[2];                             // launches before the first PCG; after a PCG and before the next
	bool use_graphs = true;
	std::vector<hipEvent_t> kt_direct;                        // events for direct-launch timing
	hipEvent_t* kt_direct_ev(size_t n) {
		while (kt_direct.size() < n) {
			hipEvent_t e;
			if (hipEventCreate(&e) != hipSuccess) break;
			kt_direct.push_back(e);
		}
		return kt_direct.data();
	}

	// ---- solver (local numbering) ----
	pdc::DBuf<int> l_static;
	pdc::DBuf<double> l_v, l_vold, l_Ms, l_Minv, l_vphys;   // 6,6,7,7,6 per body
	pdc::DBuf<double> l_speed;                               // objSpeed per body
	pdc::DBuf<double> l_ru, l_rhs, l_D, l_du, l_gdu;        // 6 per body
	pdc::DBuf<double> s_K, s_f, s_lam, s_c, s_g, s_H;        // per contact
	pdc::DBuf<double> s_rf, s_rl, s_sf, s_invC, s_W, s_kf, s_df, s_dl;
	pdc::DBuf<double> partials;                              // [k][kMaxPartialBlocks]
	pdc::DBuf<float> A32, P32, b32;                           // FP32 PCG operator, preconditioner, rhs
	pdc::DBuf<double> A64, P64, b64, pcg64;                  // FP64 PCG
	pdc::DBuf<double> x64, y64;                              // PCG solution (from its start x0) and M^-1 x0
	pdc::DBuf<pdc::Scalars> d_sc;
	pdc::Scalars* h_sc = nullptr;                            // pinned mirror

	// ---- scenes ----
	int S = 1;                                                // independent scenes
	std::vector<int> h_sbeg{0, 0};                           // per scene: first body (S + 1)
	std::vector<int> h_sclass;                                // per scene: bodies that are dynamic or emit samples
	                                                          // (picks the work shapes: inert padding never counts)
	std::vector<int> h_step;                                  // per scene: stepped by the current call
	std::vector<double> h_mu;                                 // per scene friction coefficient
	pdc::DBuf<int> b_scene;                                   // per body: scene
	pdc::DBuf<int> d_sbeg;                                    // per scene: first body (S + 1)
	pdc::DBuf<int> d_step;                                    // per scene: stepped (broadphase, integration)
	pdc::DBuf<int> d_act;                                     // the scenes of the current solve (ascending)
	pdc::DBuf<int> ls_gate;                                   // per line-search tile: its scene's search state
	std::vector<int> h_act;                                   // its host copy (source of the upload)
	pdc::DBuf<int> r_cbeg, r_lbeg, r_pbeg;                    // per scene: first contact, local body, pair
	pdc::DBuf<int> r_nnzb;                                    // per scene: Newton-matrix blocks (+ total)
	std::vector<int> h_cbeg, h_lbeg, h_pbeg, h_nnzb;
	pdc::DBuf<pdc::Tile> t_c, t_b, t_w, t_p;                  // tiles: contacts, bodies, body waves, pair waves
	pdc::DBuf<int> tb_c, tb_b, tb_w, tb_p;                    // per scene: first tile (S + 1)
	int n_tc = 0, n_tb = 0, n_tw = 0, n_tp = 0, n_tr = 0;
	int n_tr_act = 0;  // PCG tiles of the current solve's rows (the first n_tr_act of the table)
	int one_t0 = 0, one_W = 0, one_scn = 0;  // a solve whose rows are one scene's: its PCG tiles
	                                         // [one_t0, one_t0 + one_W) and its index
	int cap_tr = 0;                                           // PCG tiles launched (fixed per scene set)
	pdc_shapes shapes{{256, 256, 256, 256}, {4, 4, 4, 4}, 0};   // work shapes per size class
	// PCG launches per group the host queues (and per recorded chunk where
	// the PCG is replayed from a graph): the table's, or when it says 0, 8 for
	// one scene and 16 for several.
	int pcg_chunk() const { return shapes.pcg_chunk > 0 ? shapes.pcg_chunk : S == 1 ? 8 : 16; }
	pdc::DBuf<pdc::PcgTile> t_pcg;                            // PCG workgroup records (cap_tr)
	// The PCG works in slot space: slot kPcgRowsPerWG * tile + k is block row
	// k of PCG tile `tile` (slots past a tile's count are padding).  The
	// operator is stored ELL by slot: ell_R blocks per slot, column slots in
	// e_col (-1: empty).
	pdc::DBuf<int> tb_r;                                      // per scene: first PCG tile (S + 1)
	pdc::DBuf<int> l_slot;                                    // per local body: its slot
	// Groups of the standard PCG form's preconditioner (pcg_tol below the
	// single-reduction form's range): per local body its group's key, per key
	// the size, contact count and members; per scene the groups of each size
	// class (8, 4, 2, 1 bodies); per slot log2 of its group's size; the
	// groups' inverse blocks by slot (6 x 48 per slot: the slot's rows against
	// the 8 slots of its tile, entries outside the group unused).
	pdc::DBuf<int> g_rep, g_size, g_mem, g_best, g_cnt, l_gsz;
	bool groups_valid = false;                                // g_rep holds the last topology's groups
	pdc::DBuf<long long> g_w;
	std::vector<int> h_gcnt;
	pdc::DBuf<float> G32;
	pdc::DBuf<double> G64;
	pdc::DBuf<int> e_col;                                     // [slots][ell_R]
	pdc::DBuf<int> r_amax;                                    // most blocks in a block row
	int ell_R = 8;
	pdc::DBuf<pdc::SceneSc> d_ss;                             // per-scene scalars
	pdc::SceneSc* h_ss = nullptr;                             // pinned mirror
	size_t h_ss_cap = 0;
	pdc::DBuf<pdc::SceneCtl> d_sctl;                          // per-scene control
	std::vector<pdc::SceneCtl> h_sctl;
	pdc::DBuf<pdc::PcgScene> d_ps;                            // per-scene PCG control
	std::vector<pdc::PcgScene> h_ps;
	pdc::PcgProgress* h_prog = nullptr;                       // pinned, written by the kernels (S)
	pdc::PcgProgress* d_prog = nullptr;                       // its device address
	size_t h_prog_cap = 0;
	std::vector<pdc_solve_stats> scene_stats;                 // per scene, last solve

	// ---- statistics ----
	pdc_timings timings{};
	pdc_solve_stats last{};
	pdc::KernelTimer kt;
	hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace pdc {

// pdc_util.hip
void scan_exclusive(pdc_context* ctx, const int* in, int* out, int n);  // out has n+1 entries
// The same over min(cap, *n_dev) entries, the count read on the device.
void scan_exclusive_dev(pdc_context* ctx, const int* in, int* out, int cap, const int* n_dev);
int read_int(pdc_context* ctx, const int* dptr);
void sync_scalars(pdc_context* ctx);  // device Scalars -> ctx->h_sc (blocking)
void kt_begin(pdc_context* ctx, int which, hipEvent_t* ev);
void kt_end(pdc_context* ctx, int which, hipEvent_t ev0);
void kt_collect(pdc_context* ctx);
void kt_end_nocollect(pdc_context* ctx, hipEvent_t ev0);
// kt_end for a pair of events around `launches` launches of the timed kernel
void kt_end_span(pdc_context* ctx, hipEvent_t ev0, int launches);

// pdc_contacts.hip
// Registers a triangle mesh (vertices in the body frame); returns its id.
int register_mesh(pdc_context* ctx, int nv, const double* v, int nt, const int* tri);
void upload_bodies(pdc_context* ctx, int n, const int* shape, const double* dims, const double* mass,
                   const double* inertia, const int* is_static, const int* emit);
// Bodies appended at the end of the table as a new last scene (grids and
// samples built for them only); returns the index of the first.
int append_bodies(pdc_context* ctx, int n, const int* shape, const double* dims, const double* mass,
                  const double* inertia, const int* is_static, const int* emit);
// Bodies [first, first + n) replaced in place; their samples must fit the
// sample range those bodies held, the last body (if it emits no samples)
// keeping what the others leave.
void replace_bodies(pdc_context* ctx, int first, int n, const int* shape, const double* dims,
                    const double* mass, const double* inertia, const int* is_static, const int* emit);
// Capacity of the body table and its samples grown ahead (contents kept).
void reserve_bodies(pdc_context* ctx, int64_t bodies, int64_t samples);
// Per-scene count of bodies that are dynamic or emit samples.
void count_scene_classes(pdc_context* ctx);
void detect_contacts(pdc_context* ctx);  // fills c_* and ctx->C

// pdc_solver.hip
void build_topology(pdc_context* ctx);   // from c_id0/c_id1 (sorted by pair key)
// Newton solve of every scene with contacts; per-scene statistics in
// ctx->scene_stats, their aggregate from aggregate_stats().
void solve(pdc_context* ctx, double dt, const double* g);
pdc_solve_stats aggregate_stats(const pdc_context* ctx);
// Scene table: S scenes over the uploaded bodies (sbeg: S + 1 body offsets);
// allocates the per-scene buffers and resets friction to the configured one.
void set_scenes(pdc_context* ctx, int S, const int* sbeg);
// Marks the scenes stepped by the next detection / integration (all when
// ids == nullptr).
void set_stepped(pdc_context* ctx, int n, const int* ids);
void write_acc_default(pdc_context* ctx, double dt, const double* g);
void integrate(pdc_context* ctx, double dt);
void ensure_contact_capacity(pdc_context* ctx, size_t C);

}  // namespace pdc
