// pdc_solver.hip — (b)-(e): the primal-dual friction Newton solve on the device.
//
// Restates PrimalDualForceSolver::fillForcesIntl (reference
// src/Solver/PrimalDual/PrimalDualForceSolver.cpp:167-561) with the
// NonSmoothContactForce / OnlyNormalForce plugins
// (src/Solver/PrimalDual/NonSmoothForces/*.cpp).  Everything is FP64 except the
// inner PCG's stored operator, which is FP32 or FP64 (pdc_params.pcg_precision).
//
// Layout: bodies in the local solve set are numbered by ascending global id;
// per-body vectors are AoS (6 doubles per body), per-contact arrays are SoA
// with row stride ctx->c_stride.  Contact-to-body sums are gathers over a
// per-body pair adjacency (each pair's contacts are contiguous), so no
// floating-point atomics are used and every sum has a fixed order.

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>

#include "pdc_internal.h"
#include "pdc_reduce.h"
#include "pdc_stall.h"


namespace pdc {

constexpr int kMaxWG = 1024;  // workgroups of the reducing 1-D kernels
constexpr int kTrials = 20;   // line-search trials of the gradient-descent search, one pass (the most per pass)
// Line-search trials per pass in the Newton direction (10 in all): a search
// usually decides within its first two trials (about 2 per Newton iteration
// on the settled pile), so the first pass is short and the second, which
// returns at once for a decided search, takes the rest.
constexpr int kNewtonPasses = 2;
constexpr int kNewtonPass[kNewtonPasses] = {2, 8};
enum {
	P_GAP = 0, P_WRF, P_RL, P_AC, P_RU, P_ND_TRACE, P_ND_CNT,
	P_LS,                                  // 7 slots per trial: wrf2 rl2 ac2 gf gl ru2 gu
	P_COUNT = P_LS + 7 * kTrials
};

inline int wg_for(int64_t n) { return std::min(grid_for(n), kMaxWG); }

// Below this relative tolerance the standard form (k_pcgs_*) is used: the
// single-reduction recurrences stall near 1e-10 on ill-conditioned systems.
// The standard form is preconditioned by body groups (k_grp_*), the
// single-reduction form by the diagonal blocks.
constexpr double kSingleReductionMinTol = 1e-9;
inline bool standard_form(const pdc_context* ctx) { return ctx->p.pcg_tol < kSingleReductionMinTol; }

// ===========================================================================
// topology
// ===========================================================================
__global__ __launch_bounds__(256) void k_pair_flags(int C, const int* __restrict__ id0, const int* __restrict__ id1,
                             int* __restrict__ flag) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= C) return;
	const int a = min(id0[i], id1[i]), b = max(id0[i], id1[i]);
	int f = 1;
	if (i > 0) {
		const int a0 = min(id0[i - 1], id1[i - 1]), b0 = max(id0[i - 1], id1[i - 1]);
		f = (a != a0 || b != b0) ? 1 : 0;
	}
	flag[i] = f;
}

__global__ __launch_bounds__(256) void k_pair_fill(int C, const int* __restrict__ id0, const int* __restrict__ id1,
                            const int* __restrict__ flag, const int* __restrict__ pex,
                            int* __restrict__ pa, int* __restrict__ pb, int* __restrict__ pstart,
                            int* __restrict__ gflag) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= C) return;
	if (flag[i]) {
		const int p = pex[i];
		const int a = min(id0[i], id1[i]), b = max(id0[i], id1[i]);
		pa[p] = a;
		pb[p] = b;
		pstart[p] = i;
		gflag[a] = 1;
		gflag[b] = 1;
	}
	if (i == C - 1) pstart[pex[C]] = C;
}

__global__ __launch_bounds__(256) void k_l2g(int N, const int* __restrict__ gflag, const int* __restrict__ g2l,
                      int* __restrict__ l2g) {
	const int g = blockIdx.x * blockDim.x + threadIdx.x;
	if (g < N && gflag[g]) l2g[g2l[g]] = g;
}

__global__ __launch_bounds__(256) void k_contact_local(int C, const int* __restrict__ id0, const int* __restrict__ id1,
                                const int* __restrict__ g2l, int* __restrict__ li0,
                                int* __restrict__ li1) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= C) return;
	li0[i] = g2l[id0[i]];
	li1[i] = g2l[id1[i]];
}

__global__ __launch_bounds__(256) void k_adj_count(int NP, const int* __restrict__ pa, const int* __restrict__ pb,
                            const int* __restrict__ g2l, int* __restrict__ cnt) {
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= NP) return;
	atomicAdd(&cnt[g2l[pa[p]]], 1);
	atomicAdd(&cnt[g2l[pb[p]]], 1);
}

// Unsorted adjacency (the slot order depends on the atomics; k_adj_rank
// reorders it deterministically).
__global__ __launch_bounds__(256) void k_adj_fill(int NP, const int* __restrict__ pa, const int* __restrict__ pb,
                           const int* __restrict__ g2l, int* __restrict__ cur, int* __restrict__ nb,
                           int* __restrict__ pair, int* __restrict__ row) {
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= NP) return;
	const int la = g2l[pa[p]], lb = g2l[pb[p]];
	int k = atomicAdd(&cur[la], 1);
	nb[k] = lb;
	pair[k] = p;
	row[k] = la;
	k = atomicAdd(&cur[lb], 1);
	nb[k] = la;
	pair[k] = p;
	row[k] = lb;
}

// Sorted position of each adjacency entry = its rank by partner within the
// row (partners are unique per row).  Also records, per pair, the sorted
// adjacency entry of each of its two ends.
__global__ __launch_bounds__(256) void k_adj_rank(int E, const int* __restrict__ ptr, const int* __restrict__ nb_u,
                           const int* __restrict__ pair_u, const int* __restrict__ row_u,
                           const int* __restrict__ pa, const int* __restrict__ g2l,
                           int* __restrict__ nb, int* __restrict__ pair, int* __restrict__ pair_adj) {
	const int k = blockIdx.x * blockDim.x + threadIdx.x;
	if (k >= E) return;
	const int r = row_u[k], me = nb_u[k];
	int rank = 0;
	for (int q = ptr[r]; q < ptr[r + 1]; ++q) rank += nb_u[q] < me ? 1 : 0;
	const int dst = ptr[r] + rank;
	const int p = pair_u[k];
	nb[dst] = me;
	pair[dst] = p;
	pair_adj[2 * p + (g2l[pa[p]] == r ? 0 : 1)] = dst;
}

// One wavefront per local body: exclusive prefix of the pair sizes along the
// sorted adjacency row (the body's incidence offsets), the row's contact
// total, the static flag and the block count of its Newton-matrix row.
__global__ __launch_bounds__(256) void k_adj_pos(int NL, const int* __restrict__ ptr, const int* __restrict__ nb,
                          const int* __restrict__ pair, const int* __restrict__ pstart,
                          const int* __restrict__ l2g, const int* __restrict__ bstat,
                          int* __restrict__ adj_pos, int* __restrict__ inc_cnt, int* __restrict__ a_cnt,
                          int* __restrict__ lstat) {
	const int lane = threadIdx.x & 63;
	const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
	const int nw = (gridDim.x * blockDim.x) >> 6;
	for (int r = wave; r < NL; r += nw) {
		const int st = bstat[l2g[r]];
		int run = 0, act = 0;
		for (int base = ptr[r]; base < ptr[r + 1]; base += 64) {
			const int k = base + lane;
			int sz = 0, ac = 0;
			if (k < ptr[r + 1]) {
				const int p = pair[k];
				sz = pstart[p + 1] - pstart[p];
				ac = bstat[l2g[nb[k]]] ? 0 : 1;
			}
			int x = sz;
#pragma unroll
			for (int off = 1; off < 64; off <<= 1) {
				const int y = __shfl_up(x, off, 64);
				if (lane >= off) x += y;
			}
			if (k < ptr[r + 1]) adj_pos[k] = run + x - sz;
			run += __shfl(x, 63, 64);
			act += wave_sum(ac);
		}
		if (lane == 0) {
			inc_cnt[r] = run;
			a_cnt[r] = 1 + (st ? 0 : act);
			lstat[r] = st;
		}
	}
}

// Incidence list: entry (contact i, side j) of body r at
// inc_ptr[r] + adj_pos[entry of (r, pair)] + (i - pair start).
__global__ __launch_bounds__(256) void k_inc_fill(int C, const int* __restrict__ flag, const int* __restrict__ pex,
                           const int* __restrict__ li0, const int* __restrict__ li1,
                           const int* __restrict__ pa, const int* __restrict__ g2l,
                           const int* __restrict__ pstart, const int* __restrict__ pair_adj,
                           const int* __restrict__ adj_pos, const int* __restrict__ inc_ptr,
                           int* __restrict__ pid, int* __restrict__ inc) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= C) return;
	const int p = pex[i] + flag[i] - 1;
	pid[i] = p;
	const int ra = g2l[pa[p]];
	const int off = i - pstart[p];
	for (int j = 0; j < 2; ++j) {
		const int r = j ? li1[i] : li0[i];
		const int k = pair_adj[2 * p + (r == ra ? 0 : 1)];
		inc[inc_ptr[r] + adj_pos[k] + off] = 2 * i + j;
	}
}

// Block pattern of the Newton matrix (linsysReserve, NonSmoothForceBase.h:151-187):
// the diagonal block of every body plus one block per active-active pair,
// columns ascending, stored ELL by PCG slot: block q of the row of local body
// r is entry slot(r) * R + q, whose column is the slot of the partner
// (e_col).  adiag and adj_blk give the entries of the diagonal block and of
// each adjacency entry's block (-1: none).
__global__ __launch_bounds__(256) void k_a_fill(int NL, const int* __restrict__ lstat, const int* __restrict__ ptr,
                         const int* __restrict__ nb, const int* __restrict__ slot, int R,
                         int* __restrict__ ecol, int* __restrict__ adiag, int* __restrict__ adj_blk) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= NL) return;
	const int base = slot[r] * R;
	int o = base;
	bool diag_done = false;
	for (int k = ptr[r]; k < ptr[r + 1]; ++k) {
		const int c = nb[k];
		if (lstat[r] || lstat[c]) {
			adj_blk[k] = -1;
			continue;
		}
		if (!diag_done && c > r) {
			adiag[r] = o;
			ecol[o++] = slot[r];
			diag_done = true;
		}
		adj_blk[k] = o;
		ecol[o++] = slot[c];
	}
	if (!diag_done) {
		adiag[r] = o;
		ecol[o] = slot[r];
	}
}

// Most blocks in one block row (the ELL width before rounding).
__global__ __launch_bounds__(256) void k_amax(int NL, const int* __restrict__ cnt, int* __restrict__ out) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	int m = r < NL ? cnt[r] : 0;
	m = wave_max(m);
	if ((threadIdx.x & 63) == 0 && m > 0) atomicMax(out, m);
}

// PCG slot of every local body: row k of the scene's tile t is slot
// kPcgRowsPerWG * (first PCG tile of the scene + t) + k.
__global__ __launch_bounds__(256) void k_row_slot(int NL, const int* __restrict__ l2g, const int* __restrict__ bscene,
                                                const int* __restrict__ lbeg, const int* __restrict__ tbr,
                                                int* __restrict__ slot) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= NL) return;
	const int sc = bscene[l2g[r]];
	const int off = r - lbeg[sc];
	slot[r] = kPcgRowsPerWG * (tbr[sc] + off / kPcgRowsPerWG) + off % kPcgRowsPerWG;
}

// ---- groups of strongly coupled bodies --------------------------------------
// The standard PCG form is preconditioned by the inverse blocks of groups of up
// to kPcgRowsPerWG bodies, each group inside one PCG tile, so that a tile
// applies its groups' blocks from its own rows within the iteration's launch.
// (The reference's PCG branch uses AMGCL's smoothed-aggregation AMG,
// PrimalDualForceSolver.cpp:395-415; the grouping resembles its aggregation,
// with the strength read off the contact list, so groups are fixed for the
// timestep.)  Strength of the coupling of groups G and H of dynamic bodies:
// n_GH / sqrt(w_G w_H), n_GH the contacts between them, w the contacts of a
// group's bodies (static partners included).  Three levels merge groups of
// size 1, 2, 4 into groups of 2, 4, 8; a level runs three rounds of handshake
// matching: every unmerged group of the level's size picks the unmerged
// neighbour group of that size with the strongest coupling (ties: the smaller
// key), and two groups that pick each other merge.  A group's key is its
// smallest local index (local bodies are numbered by ascending global id, so
// this is the smallest global id).  Static bodies stay alone.  Per body:
// g_rep, the group's key; per key: g_size, g_w and the members g_mem (at most
// kPcgRowsPerWG each).  A round is a propose launch and an accept launch; a
// round without merges leaves the state as it was, so the rounds need no
// early exit.
constexpr int kGroupLevels = 3, kGroupRounds = 3;

__global__ __launch_bounds__(256) void k_grp_init(int NL, const int* __restrict__ inc_cnt, int* __restrict__ rep,
                                                int* __restrict__ gsize, long long* __restrict__ gw,
                                                int* __restrict__ mem) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= NL) return;
	rep[r] = r;
	gsize[r] = 1;
	gw[r] = inc_cnt[r];
	mem[kPcgRowsPerWG * r] = r;
}

// Group r's pick among the neighbour groups of size s: the largest n^2 / w
// (n the contacts between the groups, w the neighbour's contacts), compared
// exactly; ties go to the smaller key.  One wavefront per body (the group's
// key holder works, the others leave): the lanes take the adjacency entries
// of the group's members, 64 at a time; each lane sums the contacts of every
// entry that leads to its entry's neighbour group (the entries broadcast by
// shuffles), keeps the better of its candidates, and the wavefront reduces
// the lanes' picks under the same order (a total order, so every lane ends
// with the same pick).
__device__ __forceinline__ bool grp_better(int h1, long long n1, long long w1, int h2, long long n2, long long w2) {
	if (h1 < 0) return false;
	if (h2 < 0) return true;
	const unsigned __int128 l = (unsigned __int128)(unsigned long long)(n1 * n1) * (unsigned long long)w2;
	const unsigned __int128 r = (unsigned __int128)(unsigned long long)(n2 * n2) * (unsigned long long)w1;
	return l > r || (l == r && h1 < h2);
}

__global__ __launch_bounds__(256) void k_grp_propose(int NL, int s, const int* __restrict__ lstat,
                                                   const int* __restrict__ aptr, const int* __restrict__ anb,
                                                   const int* __restrict__ apair, const int* __restrict__ pstart,
                                                   const int* __restrict__ rep, const int* __restrict__ gsize,
                                                   const long long* __restrict__ gw, const int* __restrict__ mem,
                                                   int* __restrict__ best) {
	const int r = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
	const int lane = threadIdx.x & 63;
	if (r >= NL) return;
	int bh = -1;
	long long bn = 0, bw = 0;
	if (rep[r] == r && !lstat[r] && gsize[r] == s) {
		// the members' adjacency ranges, concatenated
		int beg[kPcgRowsPerWG / 2], len[kPcgRowsPerWG / 2];
		int E = 0;
#pragma unroll
		for (int k = 0; k < kPcgRowsPerWG / 2; ++k) {
			beg[k] = 0;
			len[k] = 0;
			if (k < s) {
				const int m = mem[kPcgRowsPerWG * r + k];
				beg[k] = aptr[m];
				len[k] = aptr[m + 1] - beg[k];
				E += len[k];
			}
		}
		// entry e of the concatenation: its neighbour group of size s (-1: none) and contacts
		auto entry = [&](int e, int& h, long long& n) {
			h = -1;
			n = 0;
			if (e >= E) return;
			int q = -1;
#pragma unroll
			for (int k = 0; k < kPcgRowsPerWG / 2; ++k) {
				if (q < 0 && e < len[k]) q = beg[k] + e;
				else if (q < 0) e -= len[k];
			}
			const int c = anb[q];
			if (lstat[c]) return;
			const int g = rep[c];
			if (g == r || gsize[g] != s) return;
			const int p = apair[q];
			h = g;
			n = pstart[p + 1] - pstart[p];
		};
		for (int a0 = 0; a0 < E; a0 += 64) {
			int ha;
			long long na;
			entry(a0 + lane, ha, na);
			long long tot = 0;
			for (int b0 = 0; b0 < E; b0 += 64) {
				int hb = ha;
				long long nb = na;
				if (b0 != a0) entry(b0 + lane, hb, nb);
				for (int l = 0; l < 64; ++l) {
					const int hl = __shfl(hb, l, 64);
					const long long nl = __shfl(nb, l, 64);
					if (ha >= 0 && hl == ha) tot += nl;
				}
			}
			if (ha >= 0) {
				const long long wa = gw[ha];
				if (grp_better(ha, tot, wa, bh, bn, bw)) {
					bh = ha;
					bn = tot;
					bw = wa;
				}
			}
		}
#pragma unroll
		for (int off = 32; off > 0; off >>= 1) {
			const int oh = __shfl_xor(bh, off, 64);
			const long long on = __shfl_xor(bn, off, 64), ow = __shfl_xor(bw, off, 64);
			if (grp_better(oh, on, ow, bh, bn, bw)) {
				bh = oh;
				bn = on;
				bw = ow;
			}
		}
	}
	if (lane == 0) best[r] = bh;
}

// Mutual picks merge into the group with the smaller key.
__global__ __launch_bounds__(256) void k_grp_accept(int NL, int s, const int* __restrict__ best, int* __restrict__ rep,
                                                  int* __restrict__ gsize, long long* __restrict__ gw,
                                                  int* __restrict__ mem) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= NL) return;
	const int h = best[r];
	if (h <= r || best[h] != r) return;
	for (int k = 0; k < s; ++k) {
		const int m = mem[kPcgRowsPerWG * h + k];
		mem[kPcgRowsPerWG * r + s + k] = m;
		rep[m] = r;
	}
	gsize[r] = 2 * s;
	gsize[h] = 0;
	gw[r] += gw[h];
}

// Size class of a group: 0 for 8 bodies, 1 for 4, 2 for 2, 3 for 1.
__device__ __forceinline__ int group_class(int size) {
	return size >= 8 ? 0 : size >= 4 ? 1 : size >= 2 ? 2 : 3;
}

// Groups per scene and size class (one workgroup per scene).
__global__ __launch_bounds__(256) void k_grp_count(int S, const int* __restrict__ lbeg, const int* __restrict__ rep,
                                                 const int* __restrict__ gsize, int* __restrict__ cnt) {
	__shared__ int c[4];
	const int sc = blockIdx.x;
	if (sc >= S) return;
	if (threadIdx.x < 4) c[threadIdx.x] = 0;
	__syncthreads();
	for (int r = lbeg[sc] + threadIdx.x; r < lbeg[sc + 1]; r += blockDim.x)
		if (rep[r] == r) atomicAdd(&c[group_class(gsize[r])], 1);
	__syncthreads();
	if (threadIdx.x < 4) cnt[4 * sc + threadIdx.x] = c[threadIdx.x];
}

// PCG slots by group: a scene's groups in order of size (8, 4, 2, 1), groups
// of one size by key, members in merge order.  Every group then starts at a
// multiple of its size, so none straddles a tile, and the scene's slots are
// exactly its ceil(NL / kPcgRowsPerWG) tiles.  gsz[slot] = log2 of the
// slot's group size.  One workgroup of 256 per scene; ranks by a workgroup
// scan of four 16-bit counters packed in 64 bits.
__global__ __launch_bounds__(256) void k_grp_slots(int S, const int* __restrict__ lbeg, const int* __restrict__ tbr,
                                                 const int* __restrict__ cnt, const int* __restrict__ rep,
                                                 const int* __restrict__ gsize, const int* __restrict__ mem,
                                                 int* __restrict__ slot, int* __restrict__ gsz) {
	__shared__ unsigned long long wtot[4];
	const int sc = blockIdx.x;
	if (sc >= S) return;
	const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
	int off[4];
	off[0] = 0;
	off[1] = 8 * cnt[4 * sc];
	off[2] = off[1] + 4 * cnt[4 * sc + 1];
	off[3] = off[2] + 2 * cnt[4 * sc + 2];
	int run[4] = {0, 0, 0, 0};
	const int base = kPcgRowsPerWG * tbr[sc];
	for (int r0 = lbeg[sc]; r0 < lbeg[sc + 1]; r0 += 256) {
		const int r = r0 + (int)threadIdx.x;
		int cls = -1, size = 0;
		if (r < lbeg[sc + 1] && rep[r] == r) {
			size = gsize[r];
			cls = group_class(size);
		}
		const unsigned long long mine = cls >= 0 ? 1ull << (16 * cls) : 0ull;
		unsigned long long x = mine;
#pragma unroll
		for (int o = 1; o < 64; o <<= 1) {
			const unsigned long long y = __shfl_up(x, o, 64);
			if (lane >= o) x += y;
		}
		if (lane == 63) wtot[wv] = x;
		__syncthreads();
		unsigned long long before = 0, total = 0;
		for (int k = 0; k < 4; ++k) {
			if (k < wv) before += wtot[k];
			total += wtot[k];
		}
		__syncthreads();
		const unsigned long long excl = before + x - mine;
		if (cls >= 0) {
			const int rank = run[cls] + (int)((excl >> (16 * cls)) & 0xffff);
			const int first = off[cls] + size * rank;
			int lg = 0;
			while ((1 << lg) < size) ++lg;
			for (int j = 0; j < size; ++j) {
				const int m = mem[kPcgRowsPerWG * r + j];
				slot[m] = base + first + j;
				gsz[base + first + j] = lg;
			}
		}
		for (int k = 0; k < 4; ++k) run[k] += (int)((total >> (16 * k)) & 0xffff);
	}
}

void ensure_contact_capacity(pdc_context* ctx, size_t C) {
	if ((int64_t)C <= ctx->c_stride) return;
	size_t cs = std::max<size_t>(C, (size_t)ctx->c_stride + ctx->c_stride / 2);
	cs = (cs + 63) & ~(size_t)63;
	auto re = [&](auto& buf, size_t k) {
		buf.release();
		buf.ensure(k * cs);
	};
	re(ctx->c_id0, 1); re(ctx->c_id1, 1); re(ctx->c_pos, 3); re(ctx->c_depth, 1);
	re(ctx->c_T, 36); re(ctx->c_lf, 3);
	re(ctx->t_key_flag, 1); ctx->t_pid.release(); ctx->t_pid.ensure(cs + 1);
	re(ctx->li0, 1); re(ctx->li1, 1);
	re(ctx->s_K, 1); re(ctx->s_f, 3); re(ctx->s_lam, 2); re(ctx->s_c, 2); re(ctx->s_g, 2);
	re(ctx->s_H, 3); re(ctx->s_rf, 3); re(ctx->s_rl, 2); re(ctx->s_sf, 3); re(ctx->s_invC, 2);
	re(ctx->s_W, 4); re(ctx->s_kf, 3); re(ctx->s_df, 3); re(ctx->s_dl, 2);
	ctx->c_stride = (int)cs;
}

// First element of every scene in a list grouped by scene: element k belongs
// to the scene of global body ids[k]; beg[s] = first k of scene >= s, beg[S] = n.
__global__ void k_scene_bounds(int n, int S, const int* __restrict__ ids, const int* __restrict__ bscene,
                               int* __restrict__ beg) {
	const int i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i > n) return;
	const int hi = i < n ? bscene[ids[i]] : S;
	const int lo = i > 0 ? bscene[ids[i - 1]] : -1;
	for (int s = lo + 1; s <= hi; ++s) beg[s] = i;
}

// Newton-matrix blocks per scene (block rows of a scene are contiguous).
__global__ void k_scene_nnzb(int S, const int* __restrict__ lbeg, const int* __restrict__ aptr, int* __restrict__ out) {
	const int s = blockIdx.x * blockDim.x + threadIdx.x;
	if (s < S) out[s] = aptr[lbeg[s + 1]] - aptr[lbeg[s]];
	if (s == S) out[S] = aptr[lbeg[S]];
}

// Tiles of per(scene) consecutive elements inside each scene; padded with
// empty tiles (scene -1) up to `pad_to`.
template <class Per>
static void make_tiles(const std::vector<int>& beg, int S, Per per_of, std::vector<Tile>& t, std::vector<int>& tb,
                       int pad_to = 0) {
	t.clear();
	tb.assign(S + 1, 0);
	for (int sc = 0; sc < S; ++sc) {
		tb[sc] = (int)t.size();
		const int per = per_of(sc);
		for (int e = beg[sc]; e < beg[sc + 1]; e += per) t.push_back(Tile{sc, e, std::min(per, beg[sc + 1] - e)});
	}
	tb[S] = (int)t.size();
	while ((int)t.size() < pad_to) t.push_back(Tile{-1, 0, 0});
}

static int upload_tiles(pdc_context* ctx, const std::vector<Tile>& t, const std::vector<int>& tb, DBuf<Tile>& dt,
                        DBuf<int>& dtb) {
	hipStream_t s = ctx->stream;
	dt.ensure(std::max<size_t>(t.size(), 1));
	dtb.ensure(tb.size());
	if (!t.empty()) PDC_HIP(hipMemcpyAsync(dt.p, t.data(), t.size() * sizeof(Tile), hipMemcpyHostToDevice, s));
	PDC_HIP(hipMemcpyAsync(dtb.p, tb.data(), tb.size() * sizeof(int), hipMemcpyHostToDevice, s));
	return (int)t.size();
}

// Per-scene contact, local-body and pair ranges, the Newton-matrix size, and
// the work tiles of every kernel kind (one read-back).
static void scene_ranges(pdc_context* ctx) {
	hipStream_t s = ctx->stream;
	const int S = ctx->S, C = ctx->C, NL = ctx->NL, NP = ctx->NP;
	ctx->r_cbeg.ensure(S + 1); ctx->r_lbeg.ensure(S + 1); ctx->r_pbeg.ensure(S + 1);
	ctx->r_nnzb.ensure(S + 1);
	hipLaunchKernelGGL(k_scene_bounds, dim3(grid_for(C + 1)), dim3(kBlock), 0, s, C, S, ctx->c_id0.p,
	                   ctx->b_scene.p, ctx->r_cbeg.p);
	hipLaunchKernelGGL(k_scene_bounds, dim3(grid_for(NL + 1)), dim3(kBlock), 0, s, NL, S, ctx->l2g.p,
	                   ctx->b_scene.p, ctx->r_lbeg.p);
	hipLaunchKernelGGL(k_scene_bounds, dim3(grid_for(NP + 1)), dim3(kBlock), 0, s, NP, S, ctx->ap_a.p,
	                   ctx->b_scene.p, ctx->r_pbeg.p);
	hipLaunchKernelGGL(k_scene_nnzb, dim3(grid_for(S + 1)), dim3(kBlock), 0, s, S, ctx->r_lbeg.p, ctx->a_ptr.p,
	                   ctx->r_nnzb.p);
	const bool grouped = standard_form(ctx);
	if (grouped) {
		ctx->g_cnt.ensure(4 * (size_t)S);
		hipLaunchKernelGGL(k_grp_count, dim3(S), dim3(256), 0, s, S, ctx->r_lbeg.p, ctx->g_rep.p, ctx->g_size.p,
		                   ctx->g_cnt.p);
		ctx->h_gcnt.resize(4 * (size_t)S);
		PDC_HIP(hipMemcpyAsync(ctx->h_gcnt.data(), ctx->g_cnt.p, 4 * S * sizeof(int), hipMemcpyDeviceToHost, s));
	}
	ctx->r_amax.ensure(1);
	PDC_HIP(hipMemsetAsync(ctx->r_amax.p, 0, sizeof(int), s));
	if (NL > 0) hipLaunchKernelGGL(k_amax, dim3(grid_for(NL)), dim3(kBlock), 0, s, NL, ctx->a_cnt.p, ctx->r_amax.p);
	PDC_CHECK_LAUNCH();
	int amax = 0;
	PDC_HIP(hipMemcpyAsync(&amax, ctx->r_amax.p, sizeof(int), hipMemcpyDeviceToHost, s));
	ctx->h_cbeg.resize(S + 1); ctx->h_lbeg.resize(S + 1); ctx->h_pbeg.resize(S + 1); ctx->h_nnzb.resize(S + 1);
	PDC_HIP(hipMemcpyAsync(ctx->h_cbeg.data(), ctx->r_cbeg.p, (S + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
	PDC_HIP(hipMemcpyAsync(ctx->h_lbeg.data(), ctx->r_lbeg.p, (S + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
	PDC_HIP(hipMemcpyAsync(ctx->h_pbeg.data(), ctx->r_pbeg.p, (S + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
	PDC_HIP(hipMemcpyAsync(ctx->h_nnzb.data(), ctx->r_nnzb.p, (S + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
	PDC_HIP(hipStreamSynchronize(s));
	ctx->NNZB = ctx->h_nnzb[S];
	ctx->ell_R = std::max(8, (amax + 7) / 8 * 8);
	std::vector<Tile> t;
	std::vector<int> tb;
	// work shapes of each scene's size class
	auto cls = [&](int sc) { return size_class(ctx->h_sclass[sc]); };
	auto per_contact = [&](int sc) { return ctx->shapes.contact_tile[cls(sc)]; };
	auto per_wave = [&](int sc) { return ctx->shapes.body_waves[cls(sc)]; };
	auto fixed = [](int per) { return [per](int) { return per; }; };
	make_tiles(ctx->h_cbeg, S, per_contact, t, tb);
	ctx->n_tc = upload_tiles(ctx, t, tb, ctx->t_c, ctx->tb_c);
	make_tiles(ctx->h_lbeg, S, fixed(kBlock), t, tb);
	ctx->n_tb = upload_tiles(ctx, t, tb, ctx->t_b, ctx->tb_b);
	make_tiles(ctx->h_lbeg, S, per_wave, t, tb);
	ctx->n_tw = upload_tiles(ctx, t, tb, ctx->t_w, ctx->tb_w);
	make_tiles(ctx->h_pbeg, S, per_wave, t, tb);
	ctx->n_tp = upload_tiles(ctx, t, tb, ctx->t_p, ctx->tb_p);
	make_tiles(ctx->h_lbeg, S, fixed(kPcgRowsPerWG), t, tb, ctx->cap_tr);
	if ((int)t.size() > ctx->cap_tr) throw HipError("PCG tile table exceeds its capacity", PDC_ERR_STATE);
	ctx->tb_r.ensure(S + 1);
	PDC_HIP(hipMemcpyAsync(ctx->tb_r.p, tb.data(), (S + 1) * sizeof(int), hipMemcpyHostToDevice, s));
	ctx->l_slot.ensure(NL + 1);
	if (grouped) {
		ctx->l_gsz.ensure((size_t)kPcgRowsPerWG * std::max(1, ctx->cap_tr));
		hipLaunchKernelGGL(k_grp_slots, dim3(S), dim3(256), 0, s, S, ctx->r_lbeg.p, ctx->tb_r.p, ctx->g_cnt.p,
		                   ctx->g_rep.p, ctx->g_size.p, ctx->g_mem.p, ctx->l_slot.p, ctx->l_gsz.p);
	} else if (NL > 0) {
		hipLaunchKernelGGL(k_row_slot, dim3(grid_for(NL)), dim3(kBlock), 0, s, NL, ctx->l2g.p, ctx->b_scene.p,
		                   ctx->r_lbeg.p, ctx->tb_r.p, ctx->l_slot.p);
	}
	PDC_CHECK_LAUNCH();
	std::vector<PcgTile> pt(t.size());
	for (size_t k = 0; k < t.size(); ++k) {
		PcgTile r{};
		r.scene = t[k].scene;
		r.first = t[k].first;
		r.count = t[k].count;
		r.t0 = r.scene >= 0 ? tb[r.scene] : 0;
		r.W = r.scene >= 0 ? tb[r.scene + 1] - tb[r.scene] : 0;
		r.done = 1;
		pt[k] = r;
	}
	ctx->n_tr = (int)t.size();
	ctx->n_tr_act = tb[S];
	int solving = 0, only = 0;
	for (int sc = 0; sc < S; ++sc)
		if (tb[sc + 1] > tb[sc]) {
			++solving;
			only = sc;
		}
	ctx->one_t0 = solving == 1 ? tb[only] : 0;
	ctx->one_W = solving == 1 ? tb[only + 1] - tb[only] : 0;
	ctx->one_scn = only;
	ctx->t_pcg.ensure(std::max<size_t>(pt.size(), 1));
	PDC_HIP(hipMemcpyAsync(ctx->t_pcg.p, pt.data(), pt.size() * sizeof(PcgTile), hipMemcpyHostToDevice, ctx->stream));
}

// The groups of the local bodies (k_grp_init, then the levels' rounds).
static void group_bodies(pdc_context* ctx) {
	hipStream_t s = ctx->stream;
	const int NL = ctx->NL;
	const size_t nb = (size_t)NL + 1;
	ctx->g_rep.ensure(nb); ctx->g_size.ensure(nb); ctx->g_best.ensure(nb); ctx->g_w.ensure(nb);
	ctx->g_mem.ensure(kPcgRowsPerWG * nb);
	if (NL == 0) return;
	const int grid = grid_for(NL);
	hipLaunchKernelGGL(k_grp_init, dim3(grid), dim3(kBlock), 0, s, NL, ctx->inc_cnt.p, ctx->g_rep.p, ctx->g_size.p,
	                   ctx->g_w.p, ctx->g_mem.p);
	for (int level = 0, size = 1; level < kGroupLevels; ++level, size *= 2)
		for (int round = 0; round < kGroupRounds; ++round) {
			hipLaunchKernelGGL(k_grp_propose, dim3(grid_for(64 * (int64_t)NL)), dim3(kBlock), 0, s, NL, size,
			                   ctx->l_static.p,
			                   ctx->adj_ptr.p, ctx->adj_nb.p, ctx->adj_pair.p, ctx->ap_start.p, ctx->g_rep.p,
			                   ctx->g_size.p, ctx->g_w.p, ctx->g_mem.p, ctx->g_best.p);
			hipLaunchKernelGGL(k_grp_accept, dim3(grid), dim3(kBlock), 0, s, NL, size, ctx->g_best.p, ctx->g_rep.p,
			                   ctx->g_size.p, ctx->g_w.p, ctx->g_mem.p);
		}
	PDC_CHECK_LAUNCH();
}

void build_topology(pdc_context* ctx) {
	hipStream_t s = ctx->stream;
	const int C = ctx->C, N = ctx->N;
	hipLaunchKernelGGL(k_pair_flags, dim3(grid_for(C)), dim3(kBlock), 0, s, C, ctx->c_id0.p,
	                   ctx->c_id1.p, ctx->t_key_flag.p);
	PDC_CHECK_LAUNCH();
	scan_exclusive(ctx, ctx->t_key_flag.p, ctx->t_pid.p, C);
	const int NP = read_int(ctx, ctx->t_pid.p + C);
	ctx->NP = NP;
	ctx->ap_a.ensure(NP); ctx->ap_b.ensure(NP); ctx->ap_start.ensure(NP + 1);
	ctx->g_flag.ensure(N + 1); ctx->g2l.ensure(N + 1); ctx->l2g.ensure(N + 1);
	PDC_HIP(hipMemsetAsync(ctx->g_flag.p, 0, (N + 1) * sizeof(int), s));
	hipLaunchKernelGGL(k_pair_fill, dim3(grid_for(C)), dim3(kBlock), 0, s, C, ctx->c_id0.p,
	                   ctx->c_id1.p, ctx->t_key_flag.p, ctx->t_pid.p, ctx->ap_a.p, ctx->ap_b.p,
	                   ctx->ap_start.p, ctx->g_flag.p);
	PDC_CHECK_LAUNCH();
	scan_exclusive(ctx, ctx->g_flag.p, ctx->g2l.p, N);
	const int NL = read_int(ctx, ctx->g2l.p + N);
	ctx->NL = NL;
	hipLaunchKernelGGL(k_l2g, dim3(grid_for(N)), dim3(kBlock), 0, s, N, ctx->g_flag.p, ctx->g2l.p,
	                   ctx->l2g.p);
	PDC_CHECK_LAUNCH();
	hipLaunchKernelGGL(k_contact_local, dim3(grid_for(C)), dim3(kBlock), 0, s, C, ctx->c_id0.p,
	                   ctx->c_id1.p, ctx->g2l.p, ctx->li0.p, ctx->li1.p);
	PDC_CHECK_LAUNCH();
	const size_t E = 2 * (size_t)NP;
	ctx->adj_cnt.ensure(NL + 1); ctx->adj_ptr.ensure(NL + 1); ctx->adj_cur.ensure(NL + 1);
	ctx->adj_nb.ensure(E); ctx->adj_pair.ensure(E); ctx->adj_pos.ensure(E); ctx->adj_blk.ensure(E);
	ctx->adj_nb_u.ensure(E); ctx->adj_pair_u.ensure(E); ctx->adj_row_u.ensure(E);
	ctx->pair_adj.ensure(E);
	PDC_HIP(hipMemsetAsync(ctx->adj_cnt.p, 0, (NL + 1) * sizeof(int), s));
	hipLaunchKernelGGL(k_adj_count, dim3(grid_for(NP)), dim3(kBlock), 0, s, NP, ctx->ap_a.p,
	                   ctx->ap_b.p, ctx->g2l.p, ctx->adj_cnt.p);
	PDC_CHECK_LAUNCH();
	scan_exclusive(ctx, ctx->adj_cnt.p, ctx->adj_ptr.p, NL);
	PDC_HIP(hipMemcpyAsync(ctx->adj_cur.p, ctx->adj_ptr.p, (NL + 1) * sizeof(int),
	                       hipMemcpyDeviceToDevice, s));
	hipLaunchKernelGGL(k_adj_fill, dim3(grid_for(NP)), dim3(kBlock), 0, s, NP, ctx->ap_a.p,
	                   ctx->ap_b.p, ctx->g2l.p, ctx->adj_cur.p, ctx->adj_nb_u.p, ctx->adj_pair_u.p,
	                   ctx->adj_row_u.p);
	PDC_CHECK_LAUNCH();
	hipLaunchKernelGGL(k_adj_rank, dim3(grid_for((int64_t)E)), dim3(kBlock), 0, s, (int)E, ctx->adj_ptr.p,
	                   ctx->adj_nb_u.p, ctx->adj_pair_u.p, ctx->adj_row_u.p, ctx->ap_a.p, ctx->g2l.p,
	                   ctx->adj_nb.p, ctx->adj_pair.p, ctx->pair_adj.p);
	PDC_CHECK_LAUNCH();
	ctx->l_static.ensure(NL + 1); ctx->a_cnt.ensure(NL + 1); ctx->a_ptr.ensure((size_t)ctx->N + 1);
	ctx->a_diag.ensure(NL + 1); ctx->inc_cnt.ensure(NL + 1); ctx->inc_ptr.ensure(NL + 1);
	hipLaunchKernelGGL(k_adj_pos, dim3(std::max(1, std::min((NL + 3) / 4, 1024))), dim3(kBlock), 0, s, NL,
	                   ctx->adj_ptr.p, ctx->adj_nb.p, ctx->adj_pair.p, ctx->ap_start.p, ctx->l2g.p,
	                   ctx->b_static.p, ctx->adj_pos.p, ctx->inc_cnt.p, ctx->a_cnt.p, ctx->l_static.p);
	PDC_CHECK_LAUNCH();
	ctx->groups_valid = false;
	if (standard_form(ctx)) {
		group_bodies(ctx);
		ctx->groups_valid = true;
	}
	scan_exclusive(ctx, ctx->inc_cnt.p, ctx->inc_ptr.p, NL);
	scan_exclusive(ctx, ctx->a_cnt.p, ctx->a_ptr.p, NL);
	ctx->inc.ensure(2 * (size_t)C);
	ctx->c_pid.ensure((size_t)C);
	hipLaunchKernelGGL(k_inc_fill, dim3(grid_for(C)), dim3(kBlock), 0, s, C, ctx->t_key_flag.p, ctx->t_pid.p,
	                   ctx->li0.p, ctx->li1.p, ctx->ap_a.p, ctx->g2l.p, ctx->ap_start.p, ctx->pair_adj.p,
	                   ctx->adj_pos.p, ctx->inc_ptr.p, ctx->c_pid.p, ctx->inc.p);
	PDC_CHECK_LAUNCH();
	scene_ranges(ctx);
	// ELL pattern over every slot of the PCG tiles (empty entries -1); the
	// storage holds the tile capacity, the solve uses its own tiles' slots
	const size_t ell = (size_t)kPcgRowsPerWG * ctx->cap_tr * ctx->ell_R;
	ctx->e_col.ensure(ell);
	const size_t ell_act = (size_t)kPcgRowsPerWG * std::max(1, ctx->n_tr_act) * ctx->ell_R;
	PDC_HIP(hipMemsetAsync(ctx->e_col.p, 0xff, ell_act * sizeof(int), s));
	hipLaunchKernelGGL(k_a_fill, dim3(grid_for(NL)), dim3(kBlock), 0, s, NL, ctx->l_static.p,
	                   ctx->adj_ptr.p, ctx->adj_nb.p, ctx->l_slot.p, ctx->ell_R, ctx->e_col.p, ctx->a_diag.p,
	                   ctx->adj_blk.p);
	PDC_CHECK_LAUNCH();
}

// ===========================================================================
// solver kernels
// ===========================================================================
struct SArgs {
	int C, cs, NL, nsc, S;
	const int* act;       // the scenes in this solve (ascending): workgroup k of a per-scene kernel takes act[k]
	int* lsgate;          // per line-search tile (contact tiles, then body tiles): its scene's ls_state while
	                      // the scene runs, else -1 (written by the per-scene kernels that change the state)
	int ntc;              // contact tiles (the body tiles' gates follow them)
	int pstride;          // partial-sum slot stride (>= every tile count)
	double restitution, compliance, reduction, dt, gdt[3];
	int spring;
	// scenes and work tiles (a workgroup's tile lies inside one scene)
	const Tile* tc;       // contacts, kBlock per tile
	const Tile* tb;       // local bodies, kBlock per tile
	const Tile* tw;       // local bodies, one wavefront each, 4 per tile
	const Tile* tp;       // active pairs, one wavefront each, 4 per tile
	const int* tbc;       // per scene: first tile of each kind (S + 1)
	const int* tbb;
	const int* tbw;
	const int* cbeg;      // per scene: first contact, first local body (S + 1)
	const int* lbeg;
	SceneSc* ss;
	const SceneCtl* ctl;
	PcgScene* ps;
	PcgProgress* prog;    // pinned host memory (see PcgProgress)
	// contacts
	const int* li0;
	const int* li1;
	const double* T;        // [C][36]
	const double* depth;
	// topology
	const int* pstart;    // per pair: first contact (NP+1)
	const int* pa;        // per pair: global id of the first body
	const int* g2l;
	const int* lstat;
	const int* l2g;
	const int* inc_ptr;   // per body: incidence offsets (NL+1)
	const int* inc;       // incidence entries 2*contact + side
	const int* pair_adj;  // per pair: sorted adjacency entry of each end
	const int* adj_blk;   // per adjacency entry: Newton-matrix block (ELL entry) or -1
	const int* slot;      // per local body: PCG slot (index of its PCG vectors and preconditioner block)
	// bodies
	double* v;
	double* vold;
	double* vphys;
	double* Ms;
	double* Minv;
	double* speed;
	double* ru;
	double* rhs;
	double* D;
	double* du;           // Newton increment; after the commit, the next PCG's start
	double* xs;           // PCG start x0 (then its solution) and S x0 (S: diagonal block)
	double* ys;
	// contact state
	double* K;
	double* f;
	double* lam;
	double* c;
	double* g;
	double* H;
	double* rf;
	double* rl;
	double* sf;
	double* invC;
	double* W;
	double* kf;
	double* df;
	double* dl;
	double* part;
};

__device__ __forceinline__ double Tc(const SArgs& a, int i, int j, int k) {
	return a.T[36 * (size_t)i + j * 18 + k];
}

// u = toLocal(v) = -T0^T v_a - T1^T v_b (ContactInfo.h:46-51); v given per side.
__device__ __forceinline__ void to_local(const SArgs& a, int i, const double* va, const double* vb,
                                         double* u) {
	for (int col = 0; col < 3; ++col) {
		double s0 = 0, s1 = 0;
		for (int r = 0; r < 6; ++r) {
			s0 += Tc(a, i, 0, col * 6 + r) * va[r];
			s1 += Tc(a, i, 1, col * 6 + r) * vb[r];
		}
		u[col] = -s0 - s1;
	}
}

template <int K>
__device__ __forceinline__ void write_partials(double (&v)[K], double* part, int stride, const int* slots,
                                               int tix = -1) {
	__shared__ double lds[K * 16];
	double out[K];
	block_sum<K, double>(v, lds, out);
	if (tix < 0) tix = blockIdx.x;
	if (threadIdx.x == 0)
		for (int k = 0; k < K; ++k) part[slots[k] * (size_t)stride + tix] = out[k];
}

// Sum of partial slot `slot` over tiles [t0, t1) by the first wavefront, in a
// fixed order; every thread of the workgroup gets it (`tmp`: an LDS scalar).
// The wavefront's loads are all in flight together (partial_lane_sum).
__device__ __forceinline__ double tile_sum(const double* part, int stride, int slot, int t0, int t1,
                                           double* tmp) {
	if (threadIdx.x < 64) {
		const double v = wave_sum_dpp(partial_lane_sum(part + slot * (size_t)stride + t0, t1 - t0));
		if (threadIdx.x == 0) *tmp = v;
	}
	__syncthreads();
	const double r = *tmp;
	__syncthreads();
	return r;
}

// The workgroup's tile and its scene; the workgroup returns at once when the
// scene is not in the solve (PDC_TILE) or has left the Newton loop (PDC_TILE_RUN).
#define PDC_TILE_BEGIN(TILES)                \
	const Tile tile = (TILES)[blockIdx.x];   \
	if (tile.scene < 0) return;              \
	const int scn = tile.scene;
#define PDC_TILE(TILES)      \
	PDC_TILE_BEGIN(TILES)    \
	if (!a.ctl[scn].act) return;
#define PDC_TILE_RUN(TILES)  \
	PDC_TILE_BEGIN(TILES)    \
	if (!a.ss[scn].run) return;

// ---- setup ----------------------------------------------------------------
__global__ __launch_bounds__(256) void k_setup_bodies(SArgs a, const double* __restrict__ vel, const double* __restrict__ ang,
                               const double* __restrict__ quat, const double* __restrict__ mass,
                               const double* __restrict__ I0) {
	PDC_TILE(a.tb);
	(void)scn;
	double acc[2] = {0, 0};
	if ((int)threadIdx.x < tile.count) {
		const int r = tile.first + threadIdx.x;
		const int g = a.l2g[r];
		const int st = a.lstat[r];
		double v[6] = {vel[3 * g], vel[3 * g + 1], vel[3 * g + 2], ang[3 * g], ang[3 * g + 1], ang[3 * g + 2]};
		for (int k = 0; k < 6; ++k) {
			a.vphys[6 * r + k] = v[k];
			a.v[6 * r + k] = v[k];
			a.vold[6 * r + k] = v[k];
		}
		if (!st)
			for (int k = 0; k < 3; ++k) a.vold[6 * r + k] = v[k] + a.gdt[k];
		double q[4] = {quat[4 * g], quat[4 * g + 1], quat[4 * g + 2], quat[4 * g + 3]};
		double R[9], Iw[6], Ii[6];
		quat_rot(q, R);
		const double d[3] = {I0[3 * g], I0[3 * g + 1], I0[3 * g + 2]};
		const double di[3] = {1.0 / d[0], 1.0 / d[1], 1.0 / d[2]};
		rot_diag_rot_sym(R, d, Iw);
		rot_diag_rot_sym(R, di, Ii);
		const double m = mass[g];
		a.Ms[7 * r] = m;
		a.Minv[7 * r] = 1.0 / m;
		for (int k = 0; k < 6; ++k) {
			a.Ms[7 * r + 1 + k] = Iw[k];
			a.Minv[7 * r + 1 + k] = Ii[k];
		}
		if (!st) {
			acc[0] += 3 * m + (Iw[0] + Iw[3] + Iw[5]);
			acc[1] += 1.0;
		}
	}
	const int slots[2] = {P_ND_TRACE, P_ND_CNT};
	write_partials<2>(acc, a.part, a.pstride, slots);
}

// objSpeed per body (PrimalDualForceSolver.cpp:38-63): one wavefront per
// body takes the maximum over its incident contacts of the norm of its own
// half relative velocity and, when the contact involves a static body, of
// the full relative velocity.
__global__ __launch_bounds__(256) void k_nondim_speed(SArgs a) {
	PDC_TILE(a.tw);
	(void)scn;
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	if (wid < tile.count) {
		const int r = tile.first + wid;
		double m = 0;
		if (a.lstat[r]) {
			if (lane == 0) a.speed[r] = 0;
			return;  // charSpeed averages over active bodies only
		}
		for (int k = a.inc_ptr[r] + lane; k < a.inc_ptr[r + 1]; k += 64) {
			const int e = a.inc[k];
			const int i = e >> 1, js = e & 1;
			const int l[2] = {a.li0[i], a.li1[i]};
			bool hasStatic = false;
			double rel[3] = {0, 0, 0}, own = 0;
			for (int j = 0; j < 2; ++j) {
				double ve[6];
				for (int q = 0; q < 6; ++q) ve[q] = a.vphys[6 * l[j] + q];
				if (!a.lstat[l[j]]) {
					for (int q = 0; q < 3; ++q) ve[q] += a.gdt[q];
				} else {
					hasStatic = true;
				}
				double half[3];
				for (int col = 0; col < 3; ++col) {
					double t = 0;
					for (int q = 0; q < 6; ++q) t += Tc(a, i, j, col * 6 + q) * ve[q];
					half[col] = -t;
				}
				for (int q = 0; q < 3; ++q) rel[q] += half[q];
				if (j == js) own = sqrt(dot3(half, half));
			}
			m = own > m ? own : m;
			if (hasStatic) {
				const double rn = sqrt(dot3(rel, rel));
				m = rn > m ? rn : m;
			}
		}
		m = wave_max(m);
		if (lane == 0) a.speed[r] = m;
	}
}

// charMass, charSpeed (PrimalDualForceSolver.cpp:23-78); one workgroup per scene.
__global__ __launch_bounds__(1024) void k_nondim_final(SArgs a, const double* __restrict__ mass, double gnorm_dt) {
	__shared__ double slot;
	__shared__ double lds[2 * 16];
	const int sc = a.act[blockIdx.x];
	if (!a.ctl[sc].act) return;
	const double tr = tile_sum(a.part, a.pstride, P_ND_TRACE, a.tbb[sc], a.tbb[sc + 1], &slot);
	const double cnt = tile_sum(a.part, a.pstride, P_ND_CNT, a.tbb[sc], a.tbb[sc + 1], &slot);
	double v[2] = {0, 0};
	for (int r = a.lbeg[sc] + threadIdx.x; r < a.lbeg[sc + 1]; r += blockDim.x) {
		if (a.lstat[r]) continue;
		const double m = mass[a.l2g[r]];
		v[0] += m;
		v[1] += a.speed[r] * m;
	}
	double out[2];
	block_sum<2, double>(v, lds, out);
	if (threadIdx.x == 0) {
		const double avg = cnt > 0 ? tr / cnt / 6 : 1.0;
		double s = out[1] / out[0];
		s = s > gnorm_dt * 0.25 ? s : gnorm_dt * 0.25;
		s = s > 1e-4 ? s : 1e-4;
		a.ss[sc].char_mass = avg;
		a.ss[sc].char_speed = s;
	}
}

// Scale velocities by charSpeed and masses by charMass; K, f, lambda init
// (NonSmoothContactForce.cpp:8-52, PrimalDualForceSolver.cpp:216-252).
__global__ __launch_bounds__(256) void k_scale_bodies(SArgs a) {
	PDC_TILE(a.tb);
	const double s = a.ss[scn].char_speed, m = a.ss[scn].char_mass;
	if ((int)threadIdx.x < tile.count) {
		const int r = tile.first + threadIdx.x;
		for (int k = 0; k < 6; ++k) {
			a.v[6 * r + k] /= s;
			a.vold[6 * r + k] /= s;
		}
		if (!a.lstat[r])
			for (int k = 0; k < 7; ++k) a.Ms[7 * r + k] /= m;
	}
}

__global__ __launch_bounds__(256) void k_preprocess(SArgs a) {
	PDC_TILE(a.tc);
	const double s = a.ss[scn].char_speed;
	if ((int)threadIdx.x < tile.count) {
		const int i = tile.first + threadIdx.x;
		double Ki = 0;
		if (a.spring) {
			Ki += -a.reduction * a.depth[i] / a.dt / s;
		} else {
			double u[3];
			to_local(a, i, &a.v[6 * a.li0[i]], &a.v[6 * a.li1[i]], u);
			Ki += u[0] * a.restitution;
		}
		a.K[i] = Ki;
		a.f[i] = 1.0;
		a.f[a.cs + i] = 0.0;
		a.f[2 * a.cs + i] = 0.0;
		a.lam[i] = 1e-2;
		a.lam[a.cs + i] = 1e-2;
	}
}

// ---- (b) per-contact terms -------------------------------------------------
// Constraint values, gradients, Hessians; partial sums of lambda.c for the
// surrogate duality gap (NonSmoothContactForce.cpp:105-118, NonSmoothForceBase.h:82-86).
__global__ __launch_bounds__(256) void k_constraints(SArgs a) {
	PDC_TILE_RUN(a.tc);
	const double mu = a.ctl[scn].mu;
	double acc[1] = {0};
	if ((int)threadIdx.x < tile.count) {
		const int i = tile.first + threadIdx.x;
		const double f0 = a.f[i];
		if (a.nsc) {
			const double f1 = a.f[a.cs + i], f2 = a.f[2 * a.cs + i];
			const double c0 = -f0, c1 = Lorentz::constraint(mu, f0, f1, f2);
			double g[2], H[3];
			Lorentz::gradient(f1, f2, g);
			Lorentz::hessian(f1, f2, H);
			a.c[i] = c0; a.c[a.cs + i] = c1;
			a.g[i] = g[0]; a.g[a.cs + i] = g[1];
			a.H[i] = H[0]; a.H[a.cs + i] = H[1]; a.H[2 * a.cs + i] = H[2];
			acc[0] += a.lam[i] * c0 + a.lam[a.cs + i] * c1;
		} else {
			a.c[i] = -f0;
			acc[0] += a.lam[i] * -f0;
		}
	}
	const int slots[1] = {P_GAP};
	write_partials<1>(acc, a.part, a.pstride, slots);
}

// Residual terms of one contact at (v', f', lambda') (NonSmoothContactForce.cpp:53-172).
// Returns rf (3), rl (2), wrf (3), ac (3) with fresh constraint values.
__device__ __forceinline__ void contact_terms(const SArgs& a, int i, const double* va, const double* vb,
                                              const double* f, const double* lam, double muc, double mu,
                                              double cm, double* rf, double* rl, double* wrf, double* ac) {
	double u[3];
	to_local(a, i, va, vb, u);
	const double Ki = a.K[i];
	const double spring_term = a.spring ? a.compliance * cm * f[0] : 0.0;
	if (a.nsc) {
		double r0 = -u[0], r1 = -u[1], r2 = -u[2];
		if (a.spring) r0 -= spring_term;
		r0 += lam[0];
		double g[2];
		Lorentz::gradient(f[1], f[2], g);
		r1 -= lam[1] * g[0];
		r2 -= lam[1] * g[1];
		r0 -= Ki;
		rf[0] = r0; rf[1] = r1; rf[2] = r2;
		rl[0] = lam[0] * -f[0] + muc;
		rl[1] = lam[1] * Lorentz::constraint(mu, f[0], f[1], f[2]) + muc;
		// tangential weight (NonSmoothContactForce.cpp:74-104)
		const double tsq = u[1] * u[1] + u[2] * u[2];
		double tn[2] = {u[1], u[2]};
		if (tsq > 0) {
			const double l = sqrt(tsq);
			tn[0] = u[1] / l;
			tn[1] = u[2] / l;
		}
		const int lb[2] = {a.li0[i], a.li1[i]};
		double fw = 0;
		for (int j = 0; j < 2; ++j) {
			if (a.lstat[lb[j]]) continue;
			double hh[6];
			for (int r = 0; r < 6; ++r) hh[r] = Tc(a, i, j, 6 + r) * tn[0] + Tc(a, i, j, 12 + r) * tn[1];
			const double* Mi = &a.Minv[7 * lb[j]];
			double t[3];
			sym3_mv(Mi + 1, hh + 3, t);
			const double q = Mi[0] * (hh[0] * hh[0] + hh[1] * hh[1] + hh[2] * hh[2]) +
			                 (hh[3] * t[0] + hh[4] * t[1] + hh[5] * t[2]);
			fw += q * cm;
		}
		const double rad = Lorentz::radius(mu, f[0]);
		double w = rad / (sqrt(tsq) + 1e-8) * fw;
		w = w < 1.0 ? w : 1.0;
		wrf[0] = r0; wrf[1] = r1 * w; wrf[2] = r2 * w;
		// Alart-Curnier (NonSmoothContactForce.cpp:53-73)
		const double tpn = f[0] - (u[0] + Ki + spring_term);
		ac[0] = (tpn > 0.0 ? tpn : 0.0) - f[0];
		const double td0 = f[1] - u[1], td1 = f[2] - u[2];
		const double tdn = sqrt(td0 * td0 + td1 * td1);
		double p0 = td0, p1 = td1;
		if (!(tdn < rad)) { p0 = td0 * rad / tdn; p1 = td1 * rad / tdn; }
		ac[1] = p0 - f[1];
		ac[2] = p1 - f[2];
	} else {
		double r0 = -u[0];
		if (a.spring) r0 -= spring_term;
		r0 += lam[0];
		r0 -= Ki;
		rf[0] = r0; rf[1] = 0; rf[2] = 0;
		wrf[0] = r0; wrf[1] = 0; wrf[2] = 0;
		rl[0] = lam[0] * -f[0] + muc;
		rl[1] = 0;
		const double tpn = f[0] - (u[0] + Ki + spring_term);
		ac[0] = (tpn > 0.0 ? tpn : 0.0) - f[0];
		ac[1] = 0; ac[2] = 0;
	}
}

// Main-iteration residuals of the contacts: stores rf, rl; partial |wrf|^2, |rl|^2, |ac|^2.
__device__ __forceinline__ void contact_res_tile(const SArgs& a, int tix) {
	__shared__ double slot;
	const Tile tile = a.tc[tix];
	if (tile.scene < 0) return;
	const int scn = tile.scene;
	if (!a.ss[scn].run) return;
	const double gapsum = tile_sum(a.part, a.pstride, P_GAP, a.tbc[scn], a.tbc[scn + 1], &slot);
	const int Cs = a.cbeg[scn + 1] - a.cbeg[scn];
	const double gap = -gapsum / ((double)(a.nsc ? 2 : 1) * Cs);
	const double muc = gap * 0.1;
	if (tix == a.tbc[scn] && threadIdx.x == 0) { a.ss[scn].gap = gap; a.ss[scn].muc = muc; }
	const double mu = a.ctl[scn].mu, cm = a.ss[scn].char_mass;
	double acc[3] = {0, 0, 0};
	if ((int)threadIdx.x < tile.count) {
		const int i = tile.first + threadIdx.x;
		const double f[3] = {a.f[i], a.f[a.cs + i], a.f[2 * a.cs + i]};
		const double lam[2] = {a.lam[i], a.lam[a.cs + i]};
		double rf[3], rl[2], wrf[3], ac[3];
		contact_terms(a, i, &a.v[6 * a.li0[i]], &a.v[6 * a.li1[i]], f, lam, muc, mu, cm, rf, rl, wrf, ac);
		for (int k = 0; k < 3; ++k) a.rf[k * a.cs + i] = rf[k];
		a.rl[i] = rl[0];
		a.rl[a.cs + i] = rl[1];
		acc[0] += wrf[0] * wrf[0] + wrf[1] * wrf[1] + wrf[2] * wrf[2];
		acc[1] += rl[0] * rl[0] + rl[1] * rl[1];
		acc[2] += ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
	}
	const int slots[3] = {P_WRF, P_RL, P_AC};
	write_partials<3>(acc, a.part, a.pstride, slots, tix);
}
__global__ __launch_bounds__(256) void k_contact_res(SArgs a) { contact_res_tile(a, blockIdx.x); }

// ru = M (v - vold) + sum_j T_j f (static rows 0)  (PrimalDualForceSolver.cpp:151-163).
// One wavefront per body: lanes take the body's incident contacts with stride
// 64, the six sums are reduced in a fixed butterfly order.
__device__ __forceinline__ void body_res_tile(const SArgs& a, int tix) {
	const Tile tile = a.tw[tix];
	if (tile.scene < 0 || !a.ss[tile.scene].run) return;
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	double acc[1] = {0};
	const double* F = a.f;
	const int fd = a.nsc ? 3 : 1;
	if (wid < tile.count) {
		const int r = tile.first + wid;
		double ru[6] = {0, 0, 0, 0, 0, 0};
		if (!a.lstat[r]) {
			// the body's rows first (every lane, same address), then the gathers
			double d[6], M[7];
			for (int q = 0; q < 6; ++q) d[q] = a.v[6 * r + q] - a.vold[6 * r + q];
			for (int q = 0; q < 7; ++q) M[q] = a.Ms[7 * r + q];
			double add[6] = {0, 0, 0, 0, 0, 0};
			for (int k = a.inc_ptr[r] + lane; k < a.inc_ptr[r + 1]; k += 64) {
				const int e = a.inc[k];
				const int i = e >> 1, j = e & 1;
				// every load of the contact issued before the first use
				const double f0 = F[i];
				const double f1 = fd == 3 ? F[a.cs + i] : 0.0, f2 = fd == 3 ? F[2 * a.cs + i] : 0.0;
#pragma unroll
				for (int row = 0; row < 6; ++row)
					add[row] += Tc(a, i, j, row) * f0 + Tc(a, i, j, 6 + row) * f1 + Tc(a, i, j, 12 + row) * f2;
			}
			for (int q = 0; q < 6; ++q) add[q] = wave_sum_dpp(add[q]);
			for (int q = 0; q < 3; ++q) ru[q] = M[0] * d[q];
			double t[3];
			sym3_mv(M + 1, d + 3, t);
			ru[3] = t[0]; ru[4] = t[1]; ru[5] = t[2];
			for (int q = 0; q < 6; ++q) ru[q] += add[q];
		}
		if (lane == 0) {
			for (int q = 0; q < 6; ++q) {
				a.ru[6 * r + q] = ru[q];
				acc[0] += ru[q] * ru[q];
			}
		}
	}
	const int slots[1] = {P_RU};
	write_partials<1>(acc, a.part, a.pstride, slots, tix);
}
__global__ __launch_bounds__(256) void k_body_res(SArgs a) { body_res_tile(a, blockIdx.x); }

// Both residual kernels in one launch (the contact part writes rf, rl and the
// scene's gap, the body part ru; neither reads what the other writes):
// blocks [0, ntc) take the contact tiles, blocks [ntc, ntc + ntw) the
// body-wavefront tiles.
__global__ __launch_bounds__(256) void k_residuals(SArgs a, int ntc) {
	if ((int)blockIdx.x < ntc) contact_res_tile(a, blockIdx.x);
	else body_res_tile(a, blockIdx.x - ntc);
}

// std::max(a, b) and std::min(a, b): (a < b) ? b : a and (b < a) ? b : a.
__device__ __forceinline__ double ref_max(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double ref_min(double a, double b) { return (b < a) ? b : a; }

// err, eps and the exit test (PrimalDualForceSolver.cpp:295-320); one workgroup per scene.
// The line-search gates of scene sc's tiles: its state while it runs, else
// -1 (every thread of the calling workgroup takes a share).
__device__ __forceinline__ void write_ls_gates(const SArgs& a, int sc, int g) {
	for (int t = a.tbc[sc] + (int)threadIdx.x; t < a.tbc[sc + 1]; t += blockDim.x) a.lsgate[t] = g;
	for (int t = a.tbw[sc] + (int)threadIdx.x; t < a.tbw[sc + 1]; t += blockDim.x) a.lsgate[a.ntc + t] = g;
}

__global__ __launch_bounds__(256) void k_err_final(SArgs a, double tol) {
	__shared__ double sums[4];
	const int sc = a.act[blockIdx.x];
	if (!a.ctl[sc].act || !a.ss[sc].run) return;
	const int c0 = a.tbc[sc], c1 = a.tbc[sc + 1], w0 = a.tbw[sc], w1 = a.tbw[sc + 1];
	// |wrf|^2, |rl|^2, |ac|^2 over the contact tiles and |ru|^2 over the body
	// tiles, one wavefront each, all loads in flight together
	{
		const int wid = threadIdx.x >> 6;
		const int slot = wid == 0 ? P_WRF : wid == 1 ? P_RL : wid == 2 ? P_AC : P_RU;
		const int t0 = wid < 3 ? c0 : w0, t1 = wid < 3 ? c1 : w1;
		const double v = wave_sum_dpp(partial_lane_sum(a.part + slot * (size_t)a.pstride + t0, t1 - t0));
		if ((threadIdx.x & 63) == 0) sums[wid] = v;
	}
	__syncthreads();
	const double wrf2 = sums[0], rl2 = sums[1], ac2 = sums[2], ru2 = sums[3];
	if (threadIdx.x == 0) {
		const double nC = a.cbeg[sc + 1] - a.cbeg[sc];
		const double nR = 6.0 * (a.lbeg[sc + 1] - a.lbeg[sc]);
		const double nF = (a.nsc ? 3.0 : 1.0) * nC, nLc = (a.nsc ? 2.0 : 1.0) * nC;
		SceneSc* S = &a.ss[sc];
		const double Hu = S->Hu;
		S->ru2 = ru2; S->wrf2 = wrf2; S->rl2 = rl2; S->ac2 = ac2;
		const double err = ru2 + wrf2 + rl2;
		S->err = err;
		S->sErr = err;
		S->sacErr = ru2 + ac2 + rl2;
		S->eps = Hu * pow(err, 0.66667 * 0.5);
		const double momen = sqrt(ru2) / nR;
		// std::max / std::min as the reference calls them: a NaN first
		// argument is kept, so a non-finite error never passes the exit test
		// (fmax / fmin would return the other, finite argument)
		double e = ref_max(momen, sqrt(wrf2) / nF);
		e = ref_max(e, sqrt(rl2) / nLc);
		e = ref_max(e, S->gap);
		e = ref_min(e, ref_max(momen, sqrt(ac2) / nC));
		S->exitErr = e;
		if (!isfinite(err) || !isfinite(e)) S->nonfinite += 1;
		// the exit test; a scene that stays starts this iteration's line search afresh
		if (e < tol) S->run = 0;
		S->ls_state = LS_PENDING;
		S->ls_step = 0;
		a.ps[sc].run = S->run;
		a.prog[sc].runtag = ((long long)S->it << 32) | (long long)(unsigned)S->run;
	}
	__syncthreads();
	write_ls_gates(a, sc, a.ss[sc].run ? (int)LS_PENDING : -1);
}

// ---- (c) assembly ----------------------------------------------------------
// Per contact: invC, sf, Kn^-1, Kt^-1, Kt^-1 sf (NonSmoothContactForce.cpp:184-219).
__global__ __launch_bounds__(256) void k_asm_prep(SArgs a) {
	PDC_TILE_RUN(a.tc);
	const double eps = a.ss[scn].eps, cm = a.ss[scn].char_mass;
	if ((int)threadIdx.x < tile.count) {
		const int i = tile.first + threadIdx.x;
		const double l0 = a.lam[i];
		if (a.nsc) {
			const double l1 = a.lam[a.cs + i];
			const double ic0 = 1. / (a.c[i] - eps * l0);
			const double ic1 = 1. / (a.c[a.cs + i] - eps * l1);
			const double g0 = a.g[i], g1 = a.g[a.cs + i];
			double lN = -l0 * ic0;
			if (a.spring) lN += a.compliance * cm;
			const double lm = -l1 * ic1;
			const double L00 = lm * g0 * g0, L01 = lm * g0 * g1, L11 = lm * g1 * g1;
			const double rl0 = a.rl[i], rl1 = a.rl[a.cs + i];
			const double sf0 = a.rf[i] + (-rl0 * ic0);
			const double sf1 = a.rf[a.cs + i] + g0 * rl1 * ic1;
			const double sf2 = a.rf[2 * a.cs + i] + g1 * rl1 * ic1;
			const double kn = 1. / (eps + lN);
			const double k00 = eps + L00 + l1 * a.H[i];
			const double k01 = 0.0 + L01 + l1 * a.H[a.cs + i];
			const double k11 = eps + L11 + l1 * a.H[2 * a.cs + i];
			const double det = k00 * k11 - k01 * k01;
			const double i00 = k11 / det, i01 = -k01 / det, i11 = k00 / det;
			a.invC[i] = ic0; a.invC[a.cs + i] = ic1;
			a.sf[i] = sf0; a.sf[a.cs + i] = sf1; a.sf[2 * a.cs + i] = sf2;
			a.W[i] = kn; a.W[a.cs + i] = i00; a.W[2 * a.cs + i] = i01; a.W[3 * a.cs + i] = i11;
			a.kf[i] = sf0 * kn;
			a.kf[a.cs + i] = i00 * sf1 + i01 * sf2;
			a.kf[2 * a.cs + i] = i01 * sf1 + i11 * sf2;
		} else {
			const double ic = 1. / (a.c[i] - eps);
			double reg = -l0 * ic;
			if (a.spring) reg += a.compliance * cm;
			const double sf0 = a.rf[i] + (-a.rl[i] * ic);
			const double kn = 1. / (eps + reg);
			a.invC[i] = ic;
			a.sf[i] = sf0;
			a.W[i] = kn;
			a.W[a.cs + i] = 0; a.W[2 * a.cs + i] = 0; a.W[3 * a.cs + i] = 0;
			a.kf[i] = sf0 * kn;
			a.kf[a.cs + i] = 0; a.kf[2 * a.cs + i] = 0;
		}
	}
}

// Value q (lane dependent) of six, as a select chain: a runtime index into a
// private array would place the array in scratch or LDS.
__device__ __forceinline__ double sel6(int q, double v0, double v1, double v2, double v3, double v4, double v5) {
	return q == 0 ? v0 : q == 1 ? v1 : q == 2 ? v2 : q == 3 ? v3 : q == 4 ? v4 : v5;
}
#define PDC_PICK6(v, q) sel6((q), (v)[0], (v)[1], (v)[2], (v)[3], (v)[4], (v)[5])
// element (row, k) of a row-major 6x6 array, row lane dependent, k constant
#define PDC_ROW6(m, row, k) sel6((row), (m)[k], (m)[6 + (k)], (m)[12 + (k)], (m)[18 + (k)], (m)[24 + (k)], (m)[30 + (k)])

// Diagonal blocks, rhs, Jacobi scaling D, block-Jacobi preconditioner P, scaled
// b; the stored matrix is the preconditioned operator A P (block (a, a) =
// A_aa P_a, the identity up to rounding).  One wavefront per body: lanes
// accumulate T_j W T_j^T (upper triangle) and y = -sum T_j Kinv sf over the
// body's incident contacts; after the wavefront sums every lane holds the
// block.  Every lane factors it (Cholesky), lane c < 6 solves for column c of
// the inverse, and lane 6 i + j < 36 forms entry (i, j) of the stored blocks.
template <class T, bool PLAIN>
__global__ __launch_bounds__(256) void k_asm_diag(SArgs a, const int* __restrict__ adiag, T* __restrict__ Aval,
                           T* __restrict__ Pval, T* __restrict__ bvec) {
	PDC_TILE_RUN(a.tw);
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	const double eps = a.ss[scn].eps;
	const bool warm = a.ss[scn].it > 0;
	const int fd = a.nsc ? 3 : 1;
	if (wid >= tile.count) return;
	const int r = tile.first + wid;
	const bool st = a.lstat[r] != 0;
	// body rows, requested before the contact gathers (every lane, same address)
	double M[7], ru[6], x0[6];
	for (int q = 0; q < 7; ++q) M[q] = a.Ms[7 * r + q];
	for (int q = 0; q < 6; ++q) {
		ru[q] = a.ru[6 * r + q];
		// PCG start: the previous iteration's increment, zero in the first
		// (PrimalDualForceSolver.cpp:407-414 hands AMGCL X = du)
		x0[q] = warm ? a.du[6 * r + q] : 0.0;
	}
	double U[21], y[6];
	for (int q = 0; q < 21; ++q) U[q] = 0;
	for (int q = 0; q < 6; ++q) y[q] = 0;
	if (!st) {
		for (int k = a.inc_ptr[r] + lane; k < a.inc_ptr[r + 1]; k += 64) {
			const int e = a.inc[k];
			const int i = e >> 1, j = e & 1;
			double tj0[6], tj1[6], tj2[6];
			for (int q = 0; q < 6; ++q) {
				tj0[q] = Tc(a, i, j, q);
				tj1[q] = Tc(a, i, j, 6 + q);
				tj2[q] = Tc(a, i, j, 12 + q);
			}
			const double kf0 = a.kf[i];
			const double kf1 = fd == 3 ? a.kf[a.cs + i] : 0.0, kf2 = fd == 3 ? a.kf[2 * a.cs + i] : 0.0;
			for (int q = 0; q < 6; ++q) y[q] -= tj0[q] * kf0 + tj1[q] * kf1 + tj2[q] * kf2;
			const double kn = a.W[i];
			const double w00 = a.W[a.cs + i], w01 = a.W[2 * a.cs + i], w11 = a.W[3 * a.cs + i];
			int u = 0;
			for (int row = 0; row < 6; ++row) {
				const double a0 = tj0[row] * kn;
				const double at0 = tj1[row] * w00 + tj2[row] * w01;
				const double at1 = tj1[row] * w01 + tj2[row] * w11;
				for (int col = row; col < 6; ++col) U[u++] += a0 * tj0[col] + at0 * tj1[col] + at1 * tj2[col];
			}
		}
		for (int q = 0; q < 21; ++q) U[q] = wave_sum_dpp(U[q]);
		for (int q = 0; q < 6; ++q) y[q] = wave_sum_dpp(y[q]);
	}
	// the block B (every lane)
	double S[36];
#pragma unroll
	for (int q = 0; q < 36; ++q) S[q] = 0;
	if (st) {
#pragma unroll
		for (int q = 0; q < 6; ++q) S[q * 6 + q] = 1.0 + eps;
	} else {
		int u = 0;
#pragma unroll
		for (int row = 0; row < 6; ++row)
#pragma unroll
			for (int col = row; col < 6; ++col) {
				S[row * 6 + col] = U[u];
				S[col * 6 + row] = U[u];
				++u;
			}
#pragma unroll
		for (int q = 0; q < 3; ++q) S[q * 6 + q] += M[0];
		S[3 * 6 + 3] += M[1]; S[3 * 6 + 4] += M[2]; S[3 * 6 + 5] += M[3];
		S[4 * 6 + 3] += M[2]; S[4 * 6 + 4] += M[4]; S[4 * 6 + 5] += M[5];
		S[5 * 6 + 3] += M[3]; S[5 * 6 + 4] += M[5]; S[5 * 6 + 5] += M[6];
#pragma unroll
		for (int q = 0; q < 6; ++q) S[q * 6 + q] += eps;
	}
	double Dr[6];
#pragma unroll
	for (int q = 0; q < 6; ++q) Dr[q] = 1 / sqrt(S[q * 6 + q]);
	const int sr = a.slot[r];
	if (lane < 6) {
		const double rhs = st ? 0.0 : PDC_PICK6(y, lane) - PDC_PICK6(ru, lane);
		const double dq = PDC_PICK6(Dr, lane);
		a.D[6 * r + lane] = dq;
		a.rhs[6 * r + lane] = rhs;
		bvec[6 * sr + lane] = (T)(rhs * dq);
	}
	// scaled block S = D B D
#pragma unroll
	for (int ii = 0; ii < 6; ++ii)
#pragma unroll
		for (int jj = 0; jj < 6; ++jj) S[ii * 6 + jj] = S[ii * 6 + jj] * (Dr[ii] * Dr[jj]);
	// Cholesky S = L L^T with reciprocal pivots (every lane)
	double L[36], Ld[6];
	bool spd = true;
#pragma unroll
	for (int q = 0; q < 36; ++q) L[q] = 0;
#pragma unroll
	for (int j = 0; j < 6; ++j) {
		double d = S[j * 6 + j];
#pragma unroll
		for (int k = 0; k < j; ++k) d -= L[j * 6 + k] * L[j * 6 + k];
		spd = spd && d > 0;
		Ld[j] = 1 / sqrt(d > 0 ? d : 1.0);
		L[j * 6 + j] = 1 / Ld[j];
#pragma unroll
		for (int i = j + 1; i < 6; ++i) {
			double t = S[i * 6 + j];
#pragma unroll
			for (int k = 0; k < j; ++k) t -= L[i * 6 + k] * L[j * 6 + k];
			L[i * 6 + j] = t * Ld[j];
		}
	}
	// column c = lane of S^-1 (lanes 0..5): L z = e_c, L^T x = z; not SPD:
	// the inverse of the diagonal
	const int c = lane < 6 ? lane : 0;
	double e[6];
	if (spd) {
#pragma unroll
		for (int i = 0; i < 6; ++i) {
			double t = i == c ? 1.0 : 0.0;
#pragma unroll
			for (int k = 0; k < i; ++k) t -= L[i * 6 + k] * e[k];
			e[i] = t * Ld[i];
		}
#pragma unroll
		for (int i = 5; i >= 0; --i) {
			double t = e[i];
#pragma unroll
			for (int k = i + 1; k < 6; ++k) t -= L[k * 6 + i] * e[k];
			e[i] = t * Ld[i];
		}
	} else {
#pragma unroll
		for (int i = 0; i < 6; ++i) e[i] = i == c ? 1.0 / S[i * 6 + i] : 0.0;
	}
	// the preconditioner as stored
#pragma unroll
	for (int i = 0; i < 6; ++i) e[i] = (double)(T)e[i];
	// lane 6 ii + jj: column jj of the inverse from lane jj
	const int ii = lane < 36 ? lane / 6 : 0, jj = lane < 36 ? lane - 6 * (lane / 6) : 0;
	double col[6];
#pragma unroll
	for (int k = 0; k < 6; ++k) col[k] = __shfl(e[k], jj, 64);
	if (lane < 36) {
		double acc = 0;
		// PLAIN: the block itself (column jj of the identity: exact); else S P
#pragma unroll
		for (int k = 0; k < 6; ++k) acc += PDC_ROW6(S, ii, k) * (PLAIN ? (k == jj ? 1.0 : 0.0) : col[k]);
		Pval[36 * (size_t)sr + lane] = (T)PDC_PICK6(col, ii);
		Aval[36 * (size_t)adiag[r] + lane] = (T)acc;
	}
	// PCG start x0, and the vector the stored operator maps to A x0: x0 itself
	// (PLAIN) or S x0 (the operator A P)
	if (lane < 6) {
		double acc = 0;
#pragma unroll
		for (int k = 0; k < 6; ++k) acc += PDC_ROW6(S, lane, k) * x0[k];
		a.xs[6 * sr + lane] = PDC_PICK6(x0, lane);
		a.ys[6 * sr + lane] = PLAIN ? PDC_PICK6(x0, lane) : acc;
	}
}

// Off-diagonal blocks, one wavefront per body pair: S = sum over the pair's
// contacts of T_side(a) W T_side(b)^T; A_ab = D_a S D_b and A_ba = A_ab^T (the
// Newton matrix is symmetric).  Stored are the blocks of the preconditioned
// operator, A_ab P_b and A_ba P_a; P comes from k_asm_diag, which runs first.
// Pairs of at most kOffDirect contacts (the common case, LARGE = false): lane
// 6 ii + jj < 36 sums entry (ii, jj) over the contacts itself, with no
// cross-lane reduction.  Larger pairs (a launch of their own, LARGE = true, so
// that the small pairs' kernel keeps few registers): the lanes take the
// contacts with stride 64 and the 36 entries are reduced across the
// wavefront.  The rows and columns for the products with P come by shuffles.
constexpr int kOffDirect = 16;
template <class T, bool LARGE, bool PLAIN>
__global__ __launch_bounds__(256) void k_asm_off(SArgs a, const T* __restrict__ Pval, T* __restrict__ Aval) {
	PDC_TILE_RUN(a.tp);
	(void)scn;
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	const int ii = lane < 36 ? lane / 6 : 0, jj = lane < 36 ? lane - 6 * (lane / 6) : 0;
	if (wid >= tile.count) return;
	const int p = tile.first + wid;
	const int ka = a.pair_adj[2 * p], kb = a.pair_adj[2 * p + 1];
	const int blk_ab = a.adj_blk[ka], blk_ba = a.adj_blk[kb];
	if (blk_ab < 0) return;  // a static end: no coupling block
	const int c0 = a.pstart[p], c1 = a.pstart[p + 1];
	if ((c1 - c0 > kOffDirect) != LARGE) return;  // the other launch's pair
	const int ra = a.g2l[a.pa[p]];
	const int rb = a.li0[c0] == ra ? a.li1[c0] : a.li0[c0];
	// scaling and preconditioner entries, requested before the contact loop
	const double da = a.D[6 * ra + ii], db = a.D[6 * rb + jj];
	const int sa = a.slot[ra], sb = a.slot[rb];
	double pb[6], pa[6];
#pragma unroll
	for (int k = 0; k < 6; ++k) {
		pb[k] = (double)Pval[36 * (size_t)sb + 6 * k + jj];
		pa[k] = (double)Pval[36 * (size_t)sa + 6 * k + jj];
	}
	double sum = 0;
	if (!LARGE) {
#pragma unroll 4
		for (int i = c0; i < c1; ++i) {
			const int j = a.li0[i] == ra ? 0 : 1;
			const double kn = a.W[i];
			const double w00 = a.W[a.cs + i], w01 = a.W[2 * a.cs + i], w11 = a.W[3 * a.cs + i];
			const double t0 = Tc(a, i, j, ii), t1 = Tc(a, i, j, 6 + ii), t2 = Tc(a, i, j, 12 + ii);
			const double u0 = Tc(a, i, 1 - j, jj), u1 = Tc(a, i, 1 - j, 6 + jj), u2 = Tc(a, i, 1 - j, 12 + jj);
			sum += (t0 * kn) * u0 + (t1 * w00 + t2 * w01) * u1 + (t1 * w01 + t2 * w11) * u2;
		}
	} else {
		double B[36];
#pragma unroll
		for (int q = 0; q < 36; ++q) B[q] = 0;
		for (int i = c0 + lane; i < c1; i += 64) {
			const int j = a.li0[i] == ra ? 0 : 1, k = 1 - j;
			const double kn = a.W[i];
			const double w00 = a.W[a.cs + i], w01 = a.W[2 * a.cs + i], w11 = a.W[3 * a.cs + i];
			double u0[6], u1[6], u2[6];
#pragma unroll
			for (int c = 0; c < 6; ++c) {
				u0[c] = Tc(a, i, k, c);
				u1[c] = Tc(a, i, k, 6 + c);
				u2[c] = Tc(a, i, k, 12 + c);
			}
#pragma unroll
			for (int row = 0; row < 6; ++row) {
				const double t0 = Tc(a, i, j, row) * kn;
				const double t1 = Tc(a, i, j, 6 + row), t2 = Tc(a, i, j, 12 + row);
				const double s1 = t1 * w00 + t2 * w01, s2 = t1 * w01 + t2 * w11;
#pragma unroll
				for (int c = 0; c < 6; ++c) B[row * 6 + c] += t0 * u0[c] + s1 * u1[c] + s2 * u2[c];
			}
		}
#pragma unroll
		for (int q = 0; q < 36; ++q) {
			const double t = wave_sum_dpp(B[q]);
			sum = q == lane ? t : sum;
		}
	}
	// lane (ii, jj) < 36 holds A_ab[ii][jj]
	const double v = sum * (da * db);
	double bab = 0, bba = 0;
	if (PLAIN) {
		bab = v;
		bba = __shfl(v, 6 * jj + ii, 64);  // A_ba[ii][jj] = A_ab[jj][ii]
	} else {
#pragma unroll
		for (int k = 0; k < 6; ++k) {
			const double a_ik = __shfl(v, 6 * ii + k, 64);  // A_ab[ii][k]
			const double a_ki = __shfl(v, 6 * k + ii, 64);  // A_ab[k][ii] = A_ba[ii][k]
			bab += a_ik * pb[k];
			bba += a_ki * pa[k];
		}
	}
	if (lane < 36) {
		Aval[36 * (size_t)blk_ab + ii * 6 + jj] = (T)bab;
		Aval[36 * (size_t)blk_ba + ii * 6 + jj] = (T)bba;
	}
}

// ---- (d) the groups' inverse blocks -------------------------------------------
// Inverse of a symmetric positive definite 6x6 block (row-major) through its
// Cholesky factor; false when a pivot is not positive and finite.
__device__ __forceinline__ bool spd6_inverse(const double* S, double (&P)[36]) {
	double L[21], rd[6];
	bool ok = true;
#pragma unroll
	for (int j = 0; j < 6; ++j) {
#pragma unroll
		for (int i = j; i < 6; ++i) {
			double t = S[i * 6 + j];
#pragma unroll
			for (int k = 0; k < j; ++k) t -= L[i * (i + 1) / 2 + k] * L[j * (j + 1) / 2 + k];
			if (i == j) {
				ok = ok && t > 0.0 && isfinite(t);
				rd[j] = 1.0 / sqrt(t > 0.0 ? t : 1.0);
				L[j * (j + 1) / 2 + j] = 1.0 / rd[j];
			} else {
				L[i * (i + 1) / 2 + j] = t * rd[j];
			}
		}
	}
#pragma unroll
	for (int c = 0; c < 6; ++c) {
		double e[6];
#pragma unroll
		for (int i = 0; i < 6; ++i) {
			double t = i == c ? 1.0 : 0.0;
#pragma unroll
			for (int k = 0; k < i; ++k) t -= L[i * (i + 1) / 2 + k] * e[k];
			e[i] = t * rd[i];
		}
#pragma unroll
		for (int i = 5; i >= 0; --i) {
			double t = e[i];
#pragma unroll
			for (int k = i + 1; k < 6; ++k) t -= L[k * (k + 1) / 2 + i] * e[k];
			e[i] = t * rd[i];
		}
#pragma unroll
		for (int i = 0; i < 6; ++i) P[i * 6 + c] = e[i];
	}
	return ok;
}

// Inverse blocks of the body groups (standard PCG form) from the assembled,
// scaled operator: one workgroup per PCG tile, one wavefront per slot (block
// row i of its group's g x g matrix of 6 x 6 blocks: the diagonal blocks and
// the blocks between the group's bodies, found in the ELL rows), lane l
// holding elements l, l + 64, ... of the row's 36 g entries.  The matrix, a
// principal submatrix of the symmetric positive definite Newton matrix, is
// inverted in place by block Gauss-Jordan elimination without pivoting: step
// j publishes block row j through LDS, every wavefront of the group inverts
// the pivot block D = G_jj (Cholesky) and updates its row: G_jj <- D^-1,
// G_jk <- D^-1 G_jk, G_ij <- -G_ij D^-1, G_ik <- G_ik - (G_ij D^-1) G_jk.  A
// matrix with a pivot block that is not positive definite gets the inverse
// of its diagonal instead.  Stored per slot: its six rows against the tile's
// eight slots, entry (i, 6 m + c) at G[288 slot + 48 i + 6 m + c], written
// and read for the group's slots only.
constexpr int kGroupRow = 6 * kPcgRowsPerWG;             // 48 columns of a tile row
constexpr int kGroupBlock = 6 * kGroupRow;               // 288 entries per slot
static_assert((1 << kGroupLevels) == kPcgRowsPerWG, "the largest group fills one PCG tile");
template <class T>
__global__ __launch_bounds__(64 * kPcgRowsPerWG) void k_grp_inv(const PcgTile* __restrict__ tiles,
                                                        const SceneSc* __restrict__ ss, const int* __restrict__ ecol,
                                                        int R, const T* __restrict__ Aval,
                                                        const int* __restrict__ gsz, T* __restrict__ G) {
	constexpr int NW = kPcgRowsPerWG;
	constexpr int kQ = (36 * NW + 63) / 64;    // elements of a block row per lane
	__shared__ double prow[2][NW][36];         // the pivot rows by step parity, block m at column slot m
	__shared__ double wcol[NW][36];            // per wavefront: its block in the pivot column, G_ij
	__shared__ double wx[NW][36];              // per wavefront: X = G_ij D^-1
	__shared__ double wd[NW][36];              // per wavefront: D^-1 (read by element index)
	const PcgTile* Tp = tiles + blockIdx.x;
	const int scn = Tp->scene, cnt = Tp->count;
	if (scn < 0 || !ss[scn].run) return;
	const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
	const int base = NW * (int)blockIdx.x;
	const bool row = w < cnt;
	// this row's group and the tile's largest group (every wavefront runs
	// that many steps: groups of different sizes can share a tile)
	const int lg = row ? gsz[base + w] : 0;
	int lgmax = 0;
	for (int k = 0; k < cnt; ++k) lgmax = max(lgmax, gsz[base + k]);
	const int gs = 1 << lg, g0 = w & ~(gs - 1), bi = w - g0;
	const int sr = base + w;
	// the ELL entry of block (bi, lane) of the row, lane < gs (-1: no block):
	// the lanes read 64 entries of the row at once and find each of the
	// group's column slots by a ballot
	int qb = -1;
	if (row)
		for (int e0 = 0; e0 < R; e0 += 64) {
			const int v = e0 + lane < R ? ecol[(size_t)sr * R + e0 + lane] : -2;
			for (int bk = 0; bk < gs; ++bk) {
				const unsigned long long hit = __ballot(v == base + g0 + bk);
				if (hit && lane == bk && qb < 0) qb = sr * R + e0 + (__ffsll((long long)hit) - 1);
			}
		}
	double M[kQ], d0[kQ];
#pragma unroll
	for (int q = 0; q < kQ; ++q) {
		const int el = lane + 64 * q, bk = min(el / 36, NW - 1), rc = el % 36;
		const int qk = __shfl(qb, bk, 64);
		M[q] = row && el < 36 * gs && qk >= 0 ? (double)Aval[36 * (size_t)qk + rc] : 0.0;
		d0[q] = M[q];
	}
	bool ok = true;
	for (int j = 0; j < (1 << lgmax); ++j) {
		const bool step = row && j < gs;
		double (*pr)[36] = prow[j & 1];
		if (step) {
#pragma unroll
			for (int q = 0; q < kQ; ++q) {
				const int el = lane + 64 * q, bk = el / 36, rc = el % 36;
				if (el < 36 * gs) {
					if (bi == j) pr[g0 + bk][rc] = M[q];
					if (bk == j) wcol[w][rc] = M[q];
				}
			}
		}
		__syncthreads();
		if (step) {
			{
				double Dinv[36];
				ok = spd6_inverse(pr[g0 + j], Dinv) && ok;
				if (lane == 0)
#pragma unroll
					for (int t = 0; t < 36; ++t) wd[w][t] = Dinv[t];
			}
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			const double* Dinv = wd[w];
			if (bi != j) {
				if (lane < 36) {
					const int r = lane / 6, c = lane % 6;
					double t = 0;
#pragma unroll
					for (int k = 0; k < 6; ++k) t += wcol[w][r * 6 + k] * Dinv[k * 6 + c];
					wx[w][lane] = t;
				}
				__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
				__builtin_amdgcn_wave_barrier();
				__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			}
#pragma unroll
			for (int q = 0; q < kQ; ++q) {
				const int el = lane + 64 * q, bk = el / 36, rc = el % 36, r = rc / 6, c = rc % 6;
				if (el < 36 * gs) {
					const double* gjk = pr[g0 + bk];
					if (bi == j) {
						if (bk == j) {
							M[q] = Dinv[rc];
						} else {
							double t = 0;
#pragma unroll
							for (int k = 0; k < 6; ++k) t += Dinv[r * 6 + k] * gjk[k * 6 + c];
							M[q] = t;
						}
					} else if (bk == j) {
						M[q] = -wx[w][rc];
					} else {
						double t = 0;
#pragma unroll
						for (int k = 0; k < 6; ++k) t += wx[w][r * 6 + k] * gjk[k * 6 + c];
						M[q] -= t;
					}
				}
			}
		}
	}
	if (!row) return;
#pragma unroll
	for (int q = 0; q < kQ; ++q) {
		const int el = lane + 64 * q, bk = el / 36, rc = el % 36, r = rc / 6, c = rc % 6;
		if (el < 36 * gs)
			G[kGroupBlock * (size_t)sr + kGroupRow * r + 6 * (g0 + bk) + c] =
			    (T)(ok ? M[q] : (bk == bi && r == c ? 1.0 / d0[q] : 0.0));
	}
}

// ---- (d) PCG ----------------------------------------------------------------
// Block-Jacobi preconditioned CG on the scaled block-CSR system from the
// initial guess x0 (the previous Newton increment, zero in the first
// iteration), in the single-reduction form of Chronopoulos and Gear
// ("s-step iterative methods for symmetric linear systems", J. Comput. Appl.
// Math. 25 (1989); the preconditioned recurrences as listed by Ghysels and
// Vanroose, Parallel Computing 40 (2014), Alg. 3).  With M = P the
// block-Jacobi inverse and the assembled operator B = A M (B_ij = A_ij P_j):
//   start: x = x0, r = b - A x0 = b - B (M^-1 x0)
//   init: u = M r, w = A u = B r, z = q = s = p = 0
//   iteration i:  gamma = (r, u), delta = (w, u), |r|^2   (from the partials)
//     stop when |r|^2 <= tol^2 |b|^2
//     beta = gamma / gamma_prev, alpha = gamma / (delta - beta gamma / alpha_prev)
//     (i = 0: beta = 0, alpha = gamma / delta)
//     m = M w, n = A m = B w
//     z = n + beta z, q = m + beta q, s = w + beta s, p = u + beta p
//     x += alpha p, r -= alpha s, u -= alpha q, w -= alpha z
//   u = M r and q = M s hold throughout; neither is stored: a block row
//   forms M r and M s from its own r and s and the recurrences above from them
//     partials of (r, u), (w, u), (r, r) for iteration i + 1
// One kernel per iteration, so the kernel boundary is the only grid-wide
// synchronisation.  Every workgroup re-reduces the partial sums in the same
// fixed order and takes the same decisions.  w (read by neighbouring block
// rows) and the partials are double-buffered by iteration parity; gamma and
// alpha of the previous iteration pass through a two-slot history written by
// workgroup 0.  The iteration counter alternates between two slots of the
// control block, so a chunk of iterations is a fixed sequence of launches
// (recorded once as a hipGraph and replayed); after convergence every launch
// returns at once.
//
// The recursively updated u and w carry absolute rounding errors of order
// eps |u_0|, so their relative accuracy falls as the residual falls; the
// vectors, the reductions and the scalars are therefore FP64 on both paths,
// while B and P are stored as T (FP32 on the FP32 path, which then solves the
// FP32-rounded operator).
//
// Layout: slot space (one block row per slot, kPcgRowsPerWG slots per
// workgroup: slot = kPcgRowsPerWG * workgroup + wavefront), so that a
// wavefront's row, its vectors, its preconditioner block and its ELL operator
// row are addressed from its indices alone.  One wavefront per block row;
// lane = 8 * row + s for the six scalar rows of the block row, the eight s
// lanes of a row take the row's ELL entries s, s + 8, ... and are summed with
// three xor shuffles.
template <class T>
struct PcgArgs {
	const int* ecol;  // [slots][R]: column slot of each ELL entry, -1 empty
	const T* B;       // A M, ELL by slot, row-major 6x6 blocks
	const T* P;       // block-Jacobi inverse, one 6x6 block per slot
	const T* G;       // standard form: the groups' inverse blocks (k_grp_inv)
	const int* gsz;   // standard form: per slot, log2 of its group's size
	const T* b;
	double* x;
	double* r;
	double* w[2];
	double* z;
	double* q;
	double* s;
	double* p;
	double* zv[2];   // standard form: z = P r, w = P q and p by iteration parity
	double* wv[2];
	double* pv[2];
	int ntiles;         // the solve's tiles (workgroups past them return)
	int one_t0, one_W, one_scn;  // a solve of one scene: its tiles [one_t0, one_t0 + one_W) and index; else one_W = 0
	int pstride;        // partial array k (gam 0, 1, del 2, 3, rr 4, 5, bb 6, e1 7, 8, e2 9, 10, e1b 11, 12)
	                    // starts at gam[0] + k pstride
	double* gam[2];  // partials, one per tile, by iteration parity
	double* del[2];
	double* rr[2];
	double* e1[2];   // standard form: (q, z), (r, w), (q, w)
	double* e1b[2];
	double* e2[2];
	double* bb;
	double* hist;    // per scene: gamma, alpha of iteration i at [4 s + 2 (i & 1)], [+ 1]
	PcgTile* tiles;         // one record per workgroup (kPcgRowsPerWG block rows)
	const SceneSc* ss;      // per scene: `run`, the scene is still in its Newton loop
	PcgScene* ps;           // per scene: done flag, iteration counters
	PcgProgress* prog;      // per scene, pinned host memory: iterations started, |r|^2/|b|^2, outcome
	double tol2;            // squared relative tolerance
	int R;                  // ELL entries per slot (a multiple of 8)
	int maxit;              // iteration cap of a solve
};

// Element k (0 or 1) of a two-slot kernel argument, selected without a runtime
// index into the argument struct (which would give it a per-thread copy).
template <class P>
__device__ __forceinline__ P pick(const P (&slots)[2], int k) {
	return k ? slots[1] : slots[0];
}

// The PCG tile of this workgroup.  Workgroups are dispatched to the eight
// XCDs in turn (workgroup b to XCD b mod 8; the grid is a multiple of 8), so
// tile (b mod 8) (G / 8) + b / 8 gives each XCD a contiguous eighth of the
// tiles: tiles near each other in slot order read their neighbours' rows
// through the same L2.  Workgroups past the solve's tiles return.
__device__ __forceinline__ int pcg_tile_id() {
	const int b = (int)blockIdx.x, per = (int)(gridDim.x >> 3);
	return (b & 7) * per + (b >> 3);
}

// The wavefront's slot in tile `tile`.
__device__ __forceinline__ int pcg_slot(int tile) {
	return kPcgRowsPerWG * tile + (int)(threadIdx.x >> 6);
}

// Row `i` (0..5) of the ELL row of slot `sr` of B times v; every lane of the
// row's eight-lane group returns the sum.  Lanes of rows 6, 7 return 0.
// Empty entries (column -1) are skipped.
template <class T, class V>
__device__ __forceinline__ double spmv_ell(const PcgArgs<T>& a, const V* __restrict__ v, int sr, int i, int sl) {
	double acc = 0;
	if (i < 6)
		for (int e = sl; e < a.R; e += 8) {
			const size_t k = (size_t)sr * a.R + e;
			const int c = a.ecol[k];
			if (c < 0) continue;
			const T* blk = a.B + 36 * k + 6 * i;
			const V* vc = v + 6 * (size_t)c;
			double d = 0;
#pragma unroll
			for (int cc = 0; cc < 6; ++cc) d += (double)blk[cc] * (double)vc[cc];
			acc += d;
		}
	acc += __shfl_xor(acc, 4, 64);
	acc += __shfl_xor(acc, 2, 64);
	acc += __shfl_xor(acc, 1, 64);
	return acc;
}

constexpr int kPcgThreads = 64 * kPcgRowsPerWG;

// The standard form's preconditioner, row i of the tile row `wid` (slot br,
// group size 2^lg): lane (i, m) holds the six entries of the row against slot
// m of the tile when m lies in the row's group (gin), and takes their product
// with slot m's entries of a vector the tile has placed in LDS (6 per tile
// row); the eight lanes of the row are summed.
template <class T>
__device__ __forceinline__ bool group_row(const PcgArgs<T>& a, int br, int wid, int i, int m, int cnt, int lg,
                                          double (&pg)[6]) {
	const bool gin = i < 6 && m < cnt && (m >> lg) == (wid >> lg);
	if (gin) {
		const T* src = a.G + kGroupBlock * (size_t)br + kGroupRow * i + 6 * m;
#pragma unroll
		for (int cc = 0; cc < 6; ++cc) pg[cc] = (double)src[cc];
	} else {
#pragma unroll
		for (int cc = 0; cc < 6; ++cc) pg[cc] = 0.0;
	}
	return gin;
}
__device__ __forceinline__ double group_apply(const double (&pg)[6], bool gin, const double* vl, int m) {
	double t = 0;
	if (gin)
#pragma unroll
		for (int cc = 0; cc < 6; ++cc) t += pg[cc] * vl[6 * m + cc];
	return group8_sum(t);
}
// Totals of N <= 4 partial arrays p0..p{N-1} over the scene's tiles
// [t0, t0 + W): wavefront k sums array k; every thread gets all totals.
// (Separate pointer arguments, not an array of pointers: a runtime-indexed
// private array would be placed in LDS.)
template <int N>
__device__ __forceinline__ void reduce_partials(const double* p0, const double* p1, const double* p2,
                                                const double* p3, int t0, int W, double (&out)[N], double* lds) {
	static_assert(N <= 4 && N <= kPcgRowsPerWG, "one partial array per wavefront");
	const int wid = threadIdx.x >> 6;
	if (wid < N) {
		const double* p = wid == 0 ? p0 : wid == 1 ? p1 : wid == 2 ? p2 : p3;
		const double t = wave_sum_dpp(partial_lane_sum(p + t0, W));
		if ((threadIdx.x & 63) == 0) lds[wid] = t;
	}
	__syncthreads();
#pragma unroll
	for (int k = 0; k < N; ++k) out[k] = lds[k];
}

// Start of a solve for one tile: its state; the scene's first tile also
// resets the scene's outcome and history.
template <class T>
__device__ __forceinline__ void pcg_tile_reset(const PcgArgs<T>& a, PcgTile* Tp, int scn, int tile) {
	if (threadIdx.x == 0) {
		Tp->done = 0;
		Tp->it = 0;
		if (tile == Tp->t0) {
			a.ps[scn].done = 0;
			a.ps[scn].iters = 0;
			a.ps[scn].bad = 0;
			// history read by iteration 0: beta = gamma / inf = 0, alpha = gamma / delta
			a.hist[4 * scn + 2] = __builtin_huge_val();
			a.hist[4 * scn + 3] = 1.0;
		}
	}
}

// End of a scene's solve after `it` iterations, seen by every tile of the
// scene (thread 0): converged, or broken down on a non-finite sum (`bad`:
// the reference takes du = 0 when its factorisation fails,
// PrimalDualForceSolver.cpp:390-394; the scene's first tile publishes it).
template <class T>
__device__ __forceinline__ void pcg_finish(const PcgArgs<T>& a, PcgTile* Tp, int scn, int it, bool bad,
                                           bool first_tile) {
	Tp->done = 1;
	if (first_tile) {
		a.ps[scn].done = 1;
		a.ps[scn].iters = it;
		a.ps[scn].bad = bad ? 1 : 0;
		a.prog[scn].outcome = (1LL << 32) | (long long)(unsigned)it;
	}
}

// A tile whose scene is not solving marks itself done.
template <class T>
__device__ __forceinline__ bool pcg_tile_skip(const PcgArgs<T>& a, PcgTile* Tp) {
	const int scn = Tp->scene;
	if (scn < 0) return true;
	if (!a.ss[scn].run) {
		if (threadIdx.x == 0) Tp->done = 1;
		return true;
	}
	return false;
}

// Initial residual of the single-reduction form, r = b - A x0 = b - B y with
// y = M^-1 x0 (x0 itself is already in x), and the tile's partial sums of
// |r|^2 (rr[1]) and |b|^2.  A zero start has y = 0 and gives r = b exactly.
template <class T>
__global__ __launch_bounds__(kPcgThreads) void k_pcg_start(PcgArgs<T> a, const double* __restrict__ y) {
	__shared__ double lds[2 * 16];
	const int tile = pcg_tile_id();
	if (tile >= a.ntiles) return;
	PcgTile* Tp = a.tiles + tile;
	if (pcg_tile_skip(a, Tp)) return;
	const int cnt = Tp->count;
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	const int i = lane >> 3, sl = lane & 7;
	double v[2] = {0, 0};
	if (wid < cnt) {
		const int br = pcg_slot(tile);
		const double ay = spmv_ell(a, y, br, i, sl);
		if (i < 6 && sl == 0) {
			const double bt = (double)a.b[6 * br + i], rt = bt - ay;
			a.r[6 * br + i] = rt;
			v[0] = rt * rt;
			v[1] = bt * bt;
		}
	}
	double out[2];
	block_sum<2, double>(v, lds, out);
	if (threadIdx.x == 0) {
		a.rr[1][tile] = out[0];
		a.bb[tile] = out[1];
	}
}

// The start a solve keeps: the warm start x0 (the previous Newton increment,
// as the reference hands AMGCL X = du) when its residual is smaller than
// that of a zero start, |b - A x0| < |b|; otherwise x0 = 0, r = b.  A warm
// start far from the new solution (the increment of a gradient-descent
// fallback, or of an earlier iteration when the Newton loop has converged
// further) would put the relative tolerance below the rounding level of
// |b - A x0| and the iteration could not reach it.  Every tile of a scene
// sums the same partials in the same order and decides alike.
template <class T>
__device__ __forceinline__ bool pcg_zero_start(const PcgArgs<T>& a, const PcgTile* Tp, double* lds) {
	double tot[2];
	reduce_partials<2>(a.rr[1], a.bb, nullptr, nullptr, Tp->t0, Tp->W, tot, lds);
	return !(tot[0] < tot[1]);
}

template <class T>
__global__ __launch_bounds__(kPcgThreads) void k_pcg_init(PcgArgs<T> a) {
	__shared__ double lds[4 * 16];
	const int tile = pcg_tile_id();
	if (tile >= a.ntiles) return;
	PcgTile* Tp = a.tiles + tile;
	if (pcg_tile_skip(a, Tp)) return;
	const int scn = Tp->scene, cnt = Tp->count;
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	const int i = lane >> 3, sl = lane & 7;
	const bool zero = pcg_zero_start(a, Tp, lds);
	__syncthreads();
	double v[3] = {0, 0, 0};
	if (wid < cnt) {
		const int br = pcg_slot(tile);
		const double w0 = zero ? spmv_ell(a, a.b, br, i, sl) : spmv_ell(a, a.r, br, i, sl);
		if (i < 6 && sl == 0) {
			const int t = 6 * br + i;
			double rrow[6];
#pragma unroll
			for (int c = 0; c < 6; ++c) rrow[c] = zero ? (double)a.b[6 * br + c] : a.r[6 * br + c];
			const double rt = PDC_PICK6(rrow, i);
			double ut = 0;
#pragma unroll
			for (int c = 0; c < 6; ++c) ut += (double)a.P[36 * (size_t)br + 6 * i + c] * rrow[c];
			if (zero) {
				a.r[t] = rt;
				a.x[t] = 0;
			}
			a.w[0][t] = w0;
			a.z[t] = 0;
			a.s[t] = 0;
			a.p[t] = 0;
			v[0] += rt * ut;
			v[1] += w0 * ut;
			v[2] += rt * rt;
		}
	}
	double out[3];
	block_sum<3, double>(v, lds, out);
	if (threadIdx.x == 0) {
		a.gam[0][tile] = out[0];
		a.del[0][tile] = out[1];
		a.rr[0][tile] = out[2];
	}
	pcg_tile_reset(a, Tp, scn, tile);
}

// Iteration parity equals the launch slot: every chunk has an even number of
// launches and each active launch advances the scene's counter by one.
// Memory round trips: (1) the workgroup's record (scalar loads) together
// with everything addressed by the slot alone (ELL columns, the
// preconditioner block, the row vectors); (2) the partial sums of the
// previous launch (their range comes with the record), and the operator
// blocks and gathered neighbour values of the occupied ELL entries (known
// from the columns; empty entries are not read).  EM is the ELL width in
// groups of 8 entries (the loads of all of them are issued at once); EM = 0
// reads the width from the arguments.  A tile whose scene has converged
// returns after its record arrives.  Lane sums are DPP moves (every lane is
// active where they are taken), and the closing workgroup sum has its own
// LDS array, so that no wavefront writes what another may still read.
template <class T, int EM, bool ONE>
__global__ __launch_bounds__(kPcgThreads) void k_pcg_iter(PcgArgs<T> a, int slot) {
	__shared__ double lds[4 * 16];
	static_assert(kPcgRowsPerWG >= 4, "four partial arrays, one per wavefront");
	const int tile = pcg_tile_id();
	if (tile >= a.ntiles) return;
	PcgTile* __restrict__ Tp = a.tiles + tile;
	// the record by vector loads (see k_pcgs_iter)
	int vz = 0;
	asm volatile("" : "+v"(vz));
	const int4* __restrict__ rv = reinterpret_cast<const int4*>(Tp) + vz;
	const int4 ra = rv[0], rb = rv[1];
	const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
	const int lane = threadIdx.x & 63;
	const int i = lane >> 3, sl = lane & 7;
	const int ir = i < 6 ? i : 5;  // rows 6, 7 of the lane layout read row 5's addresses
	const int br = kPcgRowsPerWG * tile + wid;
	const int cur = slot, nxt = slot ^ 1;
	const double* __restrict__ wc = pick(a.w, cur);
	// round 1: the preconditioner row, the row vectors and the ELL columns of
	// the row (entries sl, sl + 8, ...)
	constexpr int kE = EM > 0 ? EM : 1;
	const int R = EM > 0 ? 8 * EM : a.R;
	double pr[6], wr[6], rrow[6], srow[6];
#pragma unroll
	for (int cc = 0; cc < 6; ++cc) {
		pr[cc] = (double)a.P[36 * (size_t)br + 6 * ir + cc];
		wr[cc] = wc[6 * br + cc];
		rrow[cc] = a.r[6 * br + cc];
		srow[cc] = a.s[6 * br + cc];
	}
	const int t = 6 * br + ir;
	const double zo = a.z[t], po = a.p[t], xo = a.x[t];
	int ec[kE];
#pragma unroll
	for (int m = 0; m < kE; ++m) ec[m] = a.ecol[(size_t)br * R + sl + 8 * m];
	// the partial sums of the previous launch, the scene's first 256 tiles
	// here (wavefronts 0..3: gam, del, rr of parity cur, bb; arrays 0, 2, 4
	// (+ cur), 6 of the partial block), and the scene's gamma and alpha of the
	// previous iteration; with the first round when the solve is one scene's,
	// else after the record
	constexpr unsigned kPartArray = 0x6420u;
	constexpr int kFirst = 4;
	const int pk = (int)((kPartArray >> (4 * min(wid, 3))) & 15u) + (wid < 3 ? cur : 0);
	double pv[kFirst];
	const double* psrc = nullptr;
	int t0 = 0, W = 0;
	double gprev = 0, aprev = 0;
	auto load_partials = [&](int first, int count, int sc) {
		t0 = first;
		W = count;
		psrc = a.gam[0] + (size_t)pk * a.pstride + t0;
#pragma unroll
		for (int k = 0; k < kFirst; ++k) {
			pv[k] = 0.0;
			if (wid < 4 && lane + 64 * k < W) pv[k] = psrc[lane + 64 * k];
		}
		gprev = a.hist[4 * sc + 2 * nxt];
		aprev = a.hist[4 * sc + 2 * nxt + 1];
	};
	if (ONE) load_partials(a.one_t0, a.one_W, a.one_scn);
	// every lane holds the same record
	const int scn = __builtin_amdgcn_readfirstlane(ra.x), cnt = __builtin_amdgcn_readfirstlane(ra.z);
	const int done = __builtin_amdgcn_readfirstlane(rb.y), it = __builtin_amdgcn_readfirstlane(rb.z);
	if (scn < 0 || done || it >= a.maxit) return;
	if (!ONE) load_partials(__builtin_amdgcn_readfirstlane(ra.w), __builtin_amdgcn_readfirstlane(rb.x), scn);
	// round 2: the operator blocks and the neighbour values of the row's
	// occupied entries; entries sl without a branch (a lane whose entry is
	// empty, or of rows 6, 7, reads the row's first block and its own w, lines
	// its row reads anyway, and drops the product)
	// (all loads issued before the first product: a product between them
	// would wait for the loads before it)
	double bv[kE][6], vv[kE][6];
	{
		const bool use = i < 6 && ec[0] >= 0;
		const int c = use ? ec[0] : br;
		const T* blk = a.B + 36 * ((size_t)br * R + (use ? sl : 0)) + 6 * ir;
		const double* vc = wc + 6 * (size_t)c;
#pragma unroll
		for (int cc = 0; cc < 6; ++cc) {
			const double b0 = (double)blk[cc], v0 = vc[cc];
			bv[0][cc] = use ? b0 : 0.0;
			vv[0][cc] = use ? v0 : 0.0;
		}
	}
#pragma unroll
	for (int m = 1; m < kE; ++m) {
		const int c = i < 6 ? ec[m] : -1;
		if (c >= 0) {
			const T* blk = a.B + 36 * ((size_t)br * R + sl + 8 * m) + 6 * i;
			const double* vc = wc + 6 * (size_t)c;
#pragma unroll
			for (int cc = 0; cc < 6; ++cc) {
				bv[m][cc] = (double)blk[cc];
				vv[m][cc] = vc[cc];
			}
		} else {
#pragma unroll
			for (int cc = 0; cc < 6; ++cc) bv[m][cc] = vv[m][cc] = 0.0;
		}
	}
	double n = 0;
#pragma unroll
	for (int m = 0; m < kE; ++m) {
		double d = 0;
#pragma unroll
		for (int cc = 0; cc < 6; ++cc) d += bv[m][cc] * vv[m][cc];
		n += d;
	}
	if (i < 6) {
		if (EM == 0)
			for (int e = sl + 8; e < R; e += 8) {
				const size_t k = (size_t)br * R + e;
				const int c = a.ecol[k];
				if (c < 0) continue;
				const T* blk = a.B + 36 * k + 6 * i;
				const double* vc = wc + 6 * (size_t)c;
				double d = 0;
#pragma unroll
				for (int cc = 0; cc < 6; ++cc) d += (double)blk[cc] * vc[cc];
				n += d;
			}
	}
	n = group8_sum(n);
	// the previous iteration's scalars as reciprocals, formed while the
	// partial sums are still on their way (one division after the barrier)
	const double inv_gprev = 1.0 / gprev, inv_aprev = 1.0 / aprev;
	// m = M w, u = M r (not stored), M s of the old s
	double m = 0, u = 0, ms = 0;
#pragma unroll
	for (int cc = 0; cc < 6; ++cc) {
		m += pr[cc] * wr[cc];
		u += pr[cc] * rrow[cc];
		ms += pr[cc] * srow[cc];
	}
	if (wid < 4) {
		// (not 0 + pv[0]: see k_pcgs_iter)
		double part = pv[0];
#pragma unroll
		for (int k = 1; k < kFirst; ++k) part += pv[k];
		if (W > 64 * kFirst) part += partial_lane_sum(psrc + 64 * kFirst, W - 64 * kFirst);
		part = wave_sum_dpp(part);
		if (lane == 0) lds[wid] = part;
	}
	__syncthreads();
	const double gamma = lds[0], delta = lds[1], rr = lds[2], bb = lds[3];
	const bool first_tile = tile == t0;
	// a non-finite sum ends the solve as a breakdown (a NaN would otherwise
	// read as converged: every comparison with it is false)
	const bool bad = !(isfinite(rr) && isfinite(bb) && isfinite(gamma) && isfinite(delta));
	if (bad || !(rr > a.tol2 * bb)) {
		if (threadIdx.x == 0) pcg_finish(a, Tp, scn, it, bad, first_tile);
		return;
	}
	const double beta = gamma * inv_gprev;
	const double alpha = gamma / (delta - beta * gamma * inv_aprev);
	double* __restrict__ wn = pick(a.w, nxt);
	double v[3] = {0, 0, 0};
	if (wid < cnt && i < 6 && sl == 0) {
		const double wt = PDC_PICK6(wr, i), rt0 = PDC_PICK6(rrow, i), st0 = PDC_PICK6(srow, i);
		const double zt = n + beta * zo;
		const double qt = m + beta * ms;   // M s of the updated s
		const double st = wt + beta * st0;
		const double pt = u + beta * po;
		a.z[t] = zt;
		a.s[t] = st;
		a.p[t] = pt;
		a.x[t] = xo + alpha * pt;
		const double rt = rt0 - alpha * st;
		const double ut = u - alpha * qt;    // M r of the updated r
		const double wnt = wt - alpha * zt;
		a.r[t] = rt;
		wn[t] = wnt;
		v[0] += rt * ut;
		v[1] += wnt * ut;
		v[2] += rt * rt;
	}
	__shared__ double lds_b[3 * kPcgRowsPerWG];
	double out[3];
	block_sum_last<3, kPcgRowsPerWG>(v, lds_b, out);
	if (threadIdx.x == 0) {
		pick(a.gam, nxt)[tile] = out[0];
		pick(a.del, nxt)[tile] = out[1];
		pick(a.rr, nxt)[tile] = out[2];
		Tp->it = it + 1;
		if (first_tile) {
			a.hist[4 * scn + 2 * cur] = gamma;
			a.hist[4 * scn + 2 * cur + 1] = alpha;
			a.prog[scn].rel2 = rr / bb;
			a.prog[scn].it = it + 1;
		}
	}
}

// Standard PCG in one launch per iteration, for tolerances below the
// attainable accuracy of the single-reduction recurrences.  The assembly
// stores A itself (not A P) for this form, and the iteration is the textbook
// one, q = A p, x += alpha p, r -= alpha q, with p = z + beta p and z = P r;
// z and w = P q are carried per row (z_i = z_{i-1} - alpha w_{i-1}), so that a
// row forms its neighbours' p_i from their stored z, w, p with the same
// expression, and the same bits, as they form it themselves: the q = A p of a
// launch is the product with exactly the p that updates x.  (Folding P into
// the stored operator, A P rounded once, separates the two by about
// cond(diagonal block) x eps and stalls tight solves on stiff blocks.)
//   start (k_pcgs_start): r0 = b - A x0, z0 = p0 = P r0, and P b
//   F(0) (k_pcgs_init): the start kept (pcg_zero_start), q0 = A p0,
//         w0 = P q0, partials of delta = (p, q),
//         gamma = (r, z), e1 = (q, z), e1' = (r, w), e2 = (q, w), |r|^2, |b|^2
//   F(i), i >= 1 (k_pcgs_iter), from the totals of F(i-1):
//         stop when |r_{i-1}|^2 <= tol^2 |b|^2 (x_{i-1} is the solution);
//         alpha = gamma / delta, x_i = x_{i-1} + alpha p_{i-1},
//         r_i = r_{i-1} - alpha q_{i-1}, z_i = z_{i-1} - alpha w_{i-1},
//         gamma_i = (r_i, z_i) = gamma - alpha e1 - alpha e1' + alpha^2 e2,
//         beta = gamma_i / gamma, p_i = z_i + beta p_{i-1}, q_i = A p_i,
//         w_i = P q_i, and the partials of F(i)
// alpha always takes the directly summed gamma, so the identity's rounding
// enters only one beta.  z, w and p (read by neighbouring rows) and the
// partials are double-buffered by iteration parity; r, q and x are read and
// written by their own row only.  A launch whose sums are no longer positive
// (the recurrence has reached its rounding floor) ends the solve at x_{i-1};
// a non-finite sum ends it as a breakdown.  The iteration count is the count
// of x updates, as in the textbook loop.
template <class T>
__global__ __launch_bounds__(kPcgThreads) void k_pcgs_start(PcgArgs<T> a, const double* __restrict__ x0) {
	__shared__ double lds[2 * 16];
	__shared__ double rl[6 * kPcgRowsPerWG], bl[6 * kPcgRowsPerWG];
	const int tile = pcg_tile_id();
	if (tile >= a.ntiles) return;
	PcgTile* Tp = a.tiles + tile;
	if (pcg_tile_skip(a, Tp)) return;
	const int cnt = Tp->count;
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	const int i = lane >> 3, sl = lane & 7;
	const int ir = i < 6 ? i : 5;
	const int br = pcg_slot(tile);
	const int lg = wid < cnt ? a.gsz[br] : 0;
	double pg[6];
	const bool gin = group_row(a, br, wid, i, sl, cnt, lg, pg);
	double ri = 0, bi = 0;
	if (wid < cnt) {
		const double ax = spmv_ell(a, x0, br, i, sl);
		bi = (double)a.b[6 * br + ir];
		ri = bi - ax;
		if (i < 6 && sl == 0) {
			rl[6 * wid + i] = ri;
			bl[6 * wid + i] = bi;
		}
	}
	__syncthreads();
	const double zi = group_apply(pg, gin, rl, sl);
	const double zb = group_apply(pg, gin, bl, sl);
	double v[2] = {0, 0};
	if (wid < cnt && i < 6 && sl == 0) {
		const int t = 6 * br + i;
		a.r[t] = ri;
		a.zv[0][t] = zi;
		a.pv[0][t] = zi;
		a.zv[1][t] = zb;  // z of the zero start, P b
		v[0] = ri * ri;
		v[1] = bi * bi;
	}
	double out[2];
	block_sum<2, double>(v, lds, out);
	if (threadIdx.x == 0) {
		a.rr[1][tile] = out[0];
		a.bb[tile] = out[1];
	}
}

template <class T>
__global__ __launch_bounds__(kPcgThreads) void k_pcgs_init(PcgArgs<T> a) {
	__shared__ double lds[7 * 16];
	const int tile = pcg_tile_id();
	if (tile >= a.ntiles) return;
	PcgTile* Tp = a.tiles + tile;
	if (pcg_tile_skip(a, Tp)) return;
	const int scn = Tp->scene, cnt = Tp->count;
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	const int i = lane >> 3, sl = lane & 7;
	const bool zero = pcg_zero_start(a, Tp, lds);
	__syncthreads();
	__shared__ double ql[6 * kPcgRowsPerWG];
	const int br = pcg_slot(tile);
	const int lg = wid < cnt ? a.gsz[br] : 0;
	double pg[6];
	const bool gin = group_row(a, br, wid, i, sl, cnt, lg, pg);
	// q0 = A p0 with p0 = z0 of the start kept (P b for a zero start)
	const double n = wid < cnt ? spmv_ell(a, zero ? a.zv[1] : a.pv[0], br, i, sl) : 0.0;
	if (i < 6 && sl == 0) ql[6 * wid + i] = n;
	__syncthreads();
	const double wi = group_apply(pg, gin, ql, sl);
	double v[6] = {0, 0, 0, 0, 0, 0};
	if (wid < cnt) {
		if (i < 6 && sl == 0) {
			const int t = 6 * br + i;
			const double rt = zero ? (double)a.b[t] : a.r[t];
			const double zt = zero ? a.zv[1][t] : a.zv[0][t];
			if (zero) {
				a.r[t] = rt;
				a.x[t] = 0;
				a.zv[0][t] = zt;
				a.pv[0][t] = zt;
			}
			a.q[t] = n;
			a.wv[0][t] = wi;
			v[0] = zt * n;
			v[1] = rt * zt;
			v[2] = n * zt;
			v[3] = rt * wi;
			v[4] = n * wi;
			v[5] = rt * rt;
		}
	}
	double out[6];
	block_sum<6, double>(v, lds, out);
	if (threadIdx.x == 0) {
		a.del[0][tile] = out[0];
		a.gam[0][tile] = out[1];
		a.e1[0][tile] = out[2];
		a.e1b[0][tile] = out[3];
		a.e2[0][tile] = out[4];
		a.rr[0][tile] = out[5];
	}
	pcg_tile_reset(a, Tp, scn, tile);
}

// One iteration F(i), i >= 1, launch slot = parity of i - 1.  Memory round
// trips: (1) the record, the ELL columns, the preconditioner row and the
// row's own r, q, x, z, w, p; (2) the operator blocks and the neighbours'
// z, w, p of the occupied entries (their addresses come with the columns),
// combined into p_i after alpha and beta are known.  The partials of F(i-1)
// (seven arrays, one per wavefront) load with (1) when the solve's rows are
// one scene's (ONE: its tiles come with the arguments), else with (2) (the
// record gives the scene's tiles).  EM as in k_pcg_iter.
template <class T, int EM, bool ONE>
__global__ __launch_bounds__(kPcgThreads) void k_pcgs_iter(PcgArgs<T> a, int slot) {
	__shared__ double lds[8];
	static_assert(kPcgRowsPerWG >= 7, "seven partial arrays, one per wavefront");
	const int tile = pcg_tile_id();
	if (tile >= a.ntiles) return;
	PcgTile* __restrict__ Tp = a.tiles + tile;
	// the record by vector loads (an offset the compiler holds in a vector
	// register): a scalar load of it would be waited for together with the
	// kernel arguments that address the first round
	int vz = 0;
	asm volatile("" : "+v"(vz));
	const int4* __restrict__ rv = reinterpret_cast<const int4*>(Tp) + vz;
	const int4 ra = rv[0], rb = rv[1];
	const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
	const int lane = threadIdx.x & 63;
	const int i = lane >> 3, sl = lane & 7;
	const int ir = i < 6 ? i : 5;
	const int br = kPcgRowsPerWG * tile + wid;
	const int cur = slot, nxt = slot ^ 1;
	const double* __restrict__ zc = pick(a.zv, cur);
	const double* __restrict__ wc = pick(a.wv, cur);
	const double* __restrict__ pc = pick(a.pv, cur);
	// round 1
	constexpr int kE = EM > 0 ? EM : 1;
	const int R = EM > 0 ? 8 * EM : a.R;
	const size_t t = 6 * (size_t)br + ir;
	const double xo = a.x[t], ro = a.r[t], qo = a.q[t], zo = zc[t], wo = wc[t], po = pc[t];
	int ec[kE];
#pragma unroll
	for (int m = 0; m < kE; ++m) ec[m] = a.ecol[(size_t)br * R + sl + 8 * m];
	const int lgr = a.gsz[br];
	// the partial sums of F(i-1), the scene's first 256 tiles here: wavefront
	// wid < 7 sums del, gam, e1, e1b, e2, rr (parity cur) or bb, arrays 2, 0,
	// 7, 11, 9, 4, 6 (+ cur) of the partial block.  Addressed by arithmetic
	// and loaded under lane conditions: selecting among the argument pointers,
	// or a uniform branch per load, compiles to code that waits for the
	// record, or for every load, before the next load issues.
	constexpr unsigned kPartArray = 0x649B702u;
	constexpr int kFirst = 4;
	const int pk = (int)((kPartArray >> (4 * min(wid, 6))) & 15u) + (wid < 6 ? cur : 0);
	double pvl[kFirst];
	const double* psrc = nullptr;
	int t0 = 0, W = 0;
	auto load_partials = [&](int first, int count) {
		t0 = first;
		W = count;
		psrc = a.gam[0] + (size_t)pk * a.pstride + t0;
#pragma unroll
		for (int k = 0; k < kFirst; ++k) {
			pvl[k] = 0.0;
			if (wid < 7 && lane + 64 * k < W) pvl[k] = psrc[lane + 64 * k];
		}
	};
	if (ONE) load_partials(a.one_t0, a.one_W);
	// every lane holds the same record
	const int scn = __builtin_amdgcn_readfirstlane(ra.x), cnt = __builtin_amdgcn_readfirstlane(ra.z);
	const int done = __builtin_amdgcn_readfirstlane(rb.y), it = __builtin_amdgcn_readfirstlane(rb.z);
	if (scn < 0 || done || it >= a.maxit) return;
	if (!ONE) load_partials(__builtin_amdgcn_readfirstlane(ra.w), __builtin_amdgcn_readfirstlane(rb.x));
	// round 2: operator blocks, neighbour z, w, p and the row's
	// preconditioner entries
	double pg[6];
	const bool gin = group_row(a, br, wid, i, sl, cnt, wid < cnt ? lgr : 0, pg);
	double blk[kE][6], gz[kE][6], gw[kE][6], gp[kE][6];
	{
		// entries sl (about half of them occupied) without a branch: a lane
		// whose entry is empty (or of rows 6, 7) reads the row's first block
		// and its own z, w, p, lines its row reads anyway, and drops them, so
		// that no wait for these loads is placed before the later entries issue
		const bool use = i < 6 && ec[0] >= 0;
		const int c = use ? ec[0] : br;
		const T* bp = a.B + 36 * ((size_t)br * R + (use ? sl : 0)) + 6 * ir;
#pragma unroll
		for (int cc = 0; cc < 6; ++cc) {
			const double b0 = (double)bp[cc], z0 = zc[6 * (size_t)c + cc], w0 = wc[6 * (size_t)c + cc],
			             p0 = pc[6 * (size_t)c + cc];
			blk[0][cc] = use ? b0 : 0.0;
			gz[0][cc] = use ? z0 : 0.0;
			gw[0][cc] = use ? w0 : 0.0;
			gp[0][cc] = use ? p0 : 0.0;
		}
	}
#pragma unroll
	for (int m = 1; m < kE; ++m) {
		const int c = i < 6 ? ec[m] : -1;
		if (c >= 0) {
			const T* bp = a.B + 36 * ((size_t)br * R + sl + 8 * m) + 6 * i;
#pragma unroll
			for (int cc = 0; cc < 6; ++cc) {
				blk[m][cc] = (double)bp[cc];
				gz[m][cc] = zc[6 * (size_t)c + cc];
				gw[m][cc] = wc[6 * (size_t)c + cc];
				gp[m][cc] = pc[6 * (size_t)c + cc];
			}
		} else {
#pragma unroll
			for (int cc = 0; cc < 6; ++cc) blk[m][cc] = gz[m][cc] = gw[m][cc] = gp[m][cc] = 0.0;
		}
	}
	if (wid < 7) {
		// (not 0 + pvl[0]: the compiler would form that sum in the load's
		// branch and wait for every load issued so far there)
		double part = pvl[0];
#pragma unroll
		for (int k = 1; k < kFirst; ++k) part += pvl[k];
		if (W > 64 * kFirst) part += partial_lane_sum(psrc + 64 * kFirst, W - 64 * kFirst);
		part = wave_sum_dpp(part);
		if (lane == 0) lds[wid] = part;
	}
	__syncthreads();
	const double delta = lds[0], gamma = lds[1], e1 = lds[2], e1b = lds[3], e2 = lds[4], rr = lds[5], bb = lds[6];
	const bool first_tile = tile == t0;
	const bool bad = !(isfinite(rr) && isfinite(bb) && isfinite(gamma) && isfinite(delta) && isfinite(e1) &&
	                   isfinite(e1b) && isfinite(e2));
	// converged, at the rounding floor (sums no longer positive), or broken down
	if (bad || !(rr > a.tol2 * bb) || !(delta > 0.0) || !(gamma > 0.0)) {
		if (threadIdx.x == 0) pcg_finish(a, Tp, scn, it, bad, first_tile);
		return;
	}
	const double alpha = gamma / delta;
	const double beta = (gamma - alpha * e1 - alpha * e1b + alpha * alpha * e2) / gamma;
	// q_i = A p_i, each neighbour's p_i formed as it forms its own
	double n = 0;
#pragma unroll
	for (int m = 0; m < kE; ++m) {
		double d = 0;
#pragma unroll
		for (int cc = 0; cc < 6; ++cc) d += blk[m][cc] * ((gz[m][cc] - alpha * gw[m][cc]) + beta * gp[m][cc]);
		n += d;
	}
	if (EM == 0 && i < 6)
		for (int e = sl + 8; e < R; e += 8) {
			const size_t k = (size_t)br * R + e;
			const int c = a.ecol[k];
			if (c < 0) continue;
			const T* bp = a.B + 36 * k + 6 * i;
			double d = 0;
#pragma unroll
			for (int cc = 0; cc < 6; ++cc) {
				const size_t v = 6 * (size_t)c + cc;
				d += (double)bp[cc] * ((zc[v] - alpha * wc[v]) + beta * pc[v]);
			}
			n += d;
		}
	n = group8_sum(n);
	// w_i = P q_i over the row's group, from the tile's q_i in LDS
	__shared__ double ql[6 * kPcgRowsPerWG];
	if (i < 6 && sl == 0) ql[6 * wid + i] = n;
	__syncthreads();
	const double wn = group_apply(pg, gin, ql, sl);
	double v[6] = {0, 0, 0, 0, 0, 0};
	if (wid < cnt && i < 6 && sl == 0) {
		const double rn = ro - alpha * qo;
		const double zn = zo - alpha * wo;
		const double pn = zn + beta * po;
		a.x[t] = xo + alpha * po;
		a.r[t] = rn;
		a.q[t] = n;
		pick(a.zv, nxt)[t] = zn;
		pick(a.wv, nxt)[t] = wn;
		pick(a.pv, nxt)[t] = pn;
		v[0] = pn * n;
		v[1] = rn * zn;
		v[2] = n * zn;
		v[3] = rn * wn;
		v[4] = n * wn;
		v[5] = rn * rn;
	}
	__shared__ double lds_b[6 * kPcgRowsPerWG];
	double out[6];
	block_sum_last<6, kPcgRowsPerWG>(v, lds_b, out);
	if (threadIdx.x == 0) {
		pick(a.del, nxt)[tile] = out[0];
		pick(a.gam, nxt)[tile] = out[1];
		pick(a.e1, nxt)[tile] = out[2];
		pick(a.e1b, nxt)[tile] = out[3];
		pick(a.e2, nxt)[tile] = out[4];
		pick(a.rr, nxt)[tile] = out[5];
		Tp->it = it + 1;
		if (first_tile) {
			a.prog[scn].rel2 = rr / bb;
			a.prog[scn].it = it + 1;
		}
	}
}

// du from the PCG solution (undoing the Jacobi scaling), and the
// gradient-descent direction -ru of the fallback search
// (PrimalDualForceSolver.cpp:451), used only if that search runs.
// du of local body r: the PCG solution by slot, unscaled.
__device__ __forceinline__ double body_du(const SArgs& a, const double* __restrict__ x, int r, int k) {
	return x[6 * a.slot[r] + k] * a.D[6 * r + k];
}

// A solve that broke down on a non-finite sum gives du = 0, as the
// reference's failed factorisation does (PrimalDualForceSolver.cpp:390-394).
__device__ __forceinline__ void du_tile(const SArgs& a, const double* __restrict__ x, double* __restrict__ gdu,
                                        int tix) {
	const Tile tile = a.tb[tix];
	if (tile.scene < 0 || !a.ss[tile.scene].run) return;
	const bool bad = a.ps[tile.scene].bad;
	if ((int)threadIdx.x < tile.count) {
		const int r = tile.first + threadIdx.x;
		for (int k = 0; k < 6; ++k) {
			a.du[6 * r + k] = bad ? 0.0 : body_du(a, x, r, k);
			gdu[6 * r + k] = -a.ru[6 * r + k];
		}
	}
}

// Per contact, from du of its two bodies (formed here from the PCG solution,
// as du_tile forms them, so that both run in one launch).
__device__ __forceinline__ void increments_tile(const SArgs& a, const double* __restrict__ x, int tix) {
	const Tile tile = a.tc[tix];
	if (tile.scene < 0 || !a.ss[tile.scene].run) return;
	const bool bad = a.ps[tile.scene].bad;
	if ((int)threadIdx.x < tile.count) {
		const int i = tile.first + threadIdx.x;
		double du0[6], du1[6];
		for (int k = 0; k < 6; ++k) {
			du0[k] = bad ? 0.0 : body_du(a, x, a.li0[i], k);
			du1[k] = bad ? 0.0 : body_du(a, x, a.li1[i], k);
		}
		double u[3];
		to_local(a, i, du0, du1, u);
		const double l0 = a.lam[i];
		if (a.nsc) {
			const double l1 = a.lam[a.cs + i];
			double d0 = a.sf[i] - u[0], d1 = a.sf[a.cs + i] - u[1], d2 = a.sf[2 * a.cs + i] - u[2];
			d0 = a.W[i] * d0;
			const double w00 = a.W[a.cs + i], w01 = a.W[2 * a.cs + i], w11 = a.W[3 * a.cs + i];
			const double t0 = w00 * d1 + w01 * d2;
			const double t1 = w01 * d1 + w11 * d2;
			a.df[i] = d0; a.df[a.cs + i] = t0; a.df[2 * a.cs + i] = t1;
			a.dl[i] = -a.invC[i] * (-l0 * d0 + a.rl[i]);
			const double gd = a.g[i] * t0 + a.g[a.cs + i] * t1;
			a.dl[a.cs + i] = -a.invC[a.cs + i] * (l1 * gd + a.rl[a.cs + i]);
		} else {
			const double d = (a.sf[i] - u[0]) * a.W[i];
			a.df[i] = d;
			a.dl[i] = -a.invC[i] * (-l0 * d + a.rl[i]);
		}
	}
}

// du and -ru per body (blocks [0, ntb)) and the force and multiplier
// increments per contact (blocks [ntb, ntb + ntc)) in one launch.
__global__ __launch_bounds__(256) void k_du_increments(SArgs a, const double* __restrict__ x,
                                                       double* __restrict__ gdu, int ntb) {
	if ((int)blockIdx.x < ntb) du_tile(a, x, gdu, blockIdx.x);
	else increments_tile(a, x, blockIdx.x - ntb);
}

// ---- (e) line search as fused reductions (PrimalDualForceSolver.cpp:562-635) --
// The trials of one search share the direction, so they are independent:
// one pass evaluates nt step sizes 2^-k0 ... 2^-(k0+nt-1) at once (passes
// of kNewtonPass trials in the Newton direction, all kTrials at once in the
// gradient-descent one) and k_ls_final applies the reference's sequential
// acceptance rule to them, which selects exactly the step the
// one-trial-at-a-time loop would.

// Filtered trial point of contact i at `step` (NonSmoothContactForce.cpp:281-319,
// OnlyNormalForce.cpp:172-201): f' = f + 0.99 df_in, lambda' = lambda + 0.99 dl_in.
__device__ __forceinline__ void trial_point(const SArgs& a, int i, const double* __restrict__ dF,
                                            const double* __restrict__ dL, double step, double mu, double* fI,
                                            double* lI, double* dfi, double* dli) {
	const double cf[3] = {a.f[i], a.nsc ? a.f[a.cs + i] : 0.0, a.nsc ? a.f[2 * a.cs + i] : 0.0};
	const double lam[2] = {a.lam[i], a.nsc ? a.lam[a.cs + i] : 0.0};
	dfi[0] = step * dF[i];
	dfi[1] = a.nsc ? step * dF[a.cs + i] : 0.0;
	dfi[2] = a.nsc ? step * dF[2 * a.cs + i] : 0.0;
	dli[0] = step * dL[i];
	dli[1] = a.nsc ? step * dL[a.cs + i] : 0.0;
	if (a.nsc) {
		double tf[3] = {cf[0] + dfi[0], cf[1] + dfi[1], cf[2] + dfi[2]};
		if (tf[0] < 0) {
			tf[0] = 0;
			for (int k = 0; k < 3; ++k) dfi[k] = tf[k] - cf[k];
		}
		if (Lorentz::constraint(mu, tf[0], tf[1], tf[2]) > 0) {
			double pf[3];
			Lorentz::project(mu, tf, pf);
			for (int k = 0; k < 3; ++k) dfi[k] = pf[k] - cf[k];
		}
	} else {
		const double tf = cf[0] + dfi[0];
		if (tf < 0) dfi[0] = 0 - cf[0];
	}
#pragma unroll
	for (int k = 0; k < 2; ++k)
		if ((k == 0 || a.nsc) && dli[k] < 0 && lam[k] + dli[k] < 0) dli[k] = -lam[k];
#pragma unroll
	for (int k = 0; k < 3; ++k) fI[k] = cf[k] + 0.99 * dfi[k];
#pragma unroll
	for (int k = 0; k < 2; ++k) lI[k] = lam[k] + 0.99 * dli[k];
}

// Line-search trials of one pass, a trial per grid row: trial t = blockIdx.y
// of the pass is the point at step 2^-(k0 + t); its partial sums go to slots
// P_LS + 7 t.  A pass runs only for the scenes whose search is in state `want`.
//
// Contact part: partial |wrf'|^2, |rl'|^2, |ac'|^2, df_in.rf' and
// sum (dl_in / lambda') rl' (fresh constraint values).
__device__ __forceinline__ void ls_contact_tile(const SArgs& a, const double* __restrict__ du,
                                                const double* __restrict__ dF, const double* __restrict__ dL,
                                                int k0, int want, int tix) {
	const Tile tile = a.tc[tix];
	if (tile.scene < 0) return;
	const int scn = tile.scene;  // running and in state `want` (the gate)
	const int t = blockIdx.y;
	const double muc = a.ss[scn].muc, mu = a.ctl[scn].mu, cm = a.ss[scn].char_mass;
	const double step = ldexp(1.0, -(k0 + t));
	double acc[5] = {0, 0, 0, 0, 0};
	if ((int)threadIdx.x < tile.count) {
		const int i = tile.first + threadIdx.x;
		double fI[3], lI[2], dfi[3], dli[2];
		trial_point(a, i, dF, dL, step, mu, fI, lI, dfi, dli);
		const int l0 = a.li0[i], l1 = a.li1[i];
		double va[6], vb[6];
		for (int k = 0; k < 6; ++k) {
			va[k] = a.v[6 * l0 + k] + 0.99 * (step * du[6 * l0 + k]);
			vb[k] = a.v[6 * l1 + k] + 0.99 * (step * du[6 * l1 + k]);
		}
		double rf[3], rl[2], wrf[3], ac[3];
		contact_terms(a, i, va, vb, fI, lI, muc, mu, cm, rf, rl, wrf, ac);
		acc[0] = wrf[0] * wrf[0] + wrf[1] * wrf[1] + wrf[2] * wrf[2];
		acc[1] = rl[0] * rl[0] + (a.nsc ? rl[1] * rl[1] : 0.0);
		acc[2] = ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
		acc[3] = dfi[0] * rf[0] + (a.nsc ? dfi[1] * rf[1] + dfi[2] * rf[2] : 0.0);
		acc[4] = dli[0] / lI[0] * rl[0] + (a.nsc ? dli[1] / lI[1] * rl[1] : 0.0);
	}
	const int b = P_LS + 7 * t;
	const int slots[5] = {b, b + 1, b + 2, b + 3, b + 4};
	write_partials<5>(acc, a.part, a.pstride, slots, tix);
}

// Body part, one wavefront per body: ru' = M (v' - vold) + sum_j T_j f';
// partial |ru'|^2 and du_in . ru'.
__device__ __forceinline__ void ls_body_tile(const SArgs& a, const double* __restrict__ du,
                                             const double* __restrict__ dF, const double* __restrict__ dL,
                                             int k0, int want, int tix) {
	const Tile tile = a.tw[tix];
	if (tile.scene < 0) return;
	const int scn = tile.scene;  // running and in state `want` (the gate)
	const int t = blockIdx.y;
	const double mu = a.ctl[scn].mu;
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	const double step = ldexp(1.0, -(k0 + t));
	double acc[2] = {0, 0};
	if (wid < tile.count && !a.lstat[tile.first + wid]) {
		const int r = tile.first + wid;
		// the body's rows first (every lane, same address), then the gathers
		double d[6], dd[6], M[7];
		for (int q = 0; q < 6; ++q) {
			dd[q] = step * du[6 * r + q];
			d[q] = (a.v[6 * r + q] + 0.99 * dd[q]) - a.vold[6 * r + q];
		}
		for (int q = 0; q < 7; ++q) M[q] = a.Ms[7 * r + q];
		double add[6] = {0, 0, 0, 0, 0, 0};
		for (int k = a.inc_ptr[r] + lane; k < a.inc_ptr[r + 1]; k += 64) {
			const int e = a.inc[k];
			const int i = e >> 1, j = e & 1;
			double fI[3], lI[2], dfi[3], dli[2];
			trial_point(a, i, dF, dL, step, mu, fI, lI, dfi, dli);
			// fI[1] = fI[2] = 0 for the frictionless model
#pragma unroll
			for (int row = 0; row < 6; ++row)
				add[row] += Tc(a, i, j, row) * fI[0] + Tc(a, i, j, 6 + row) * fI[1] + Tc(a, i, j, 12 + row) * fI[2];
		}
#pragma unroll
		for (int q = 0; q < 6; ++q) add[q] = wave_sum_dpp(add[q]);
		if (lane == 0) {
			double ru[6];
			for (int q = 0; q < 3; ++q) ru[q] = M[0] * d[q];
			double tt[3];
			sym3_mv(M + 1, d + 3, tt);
			ru[3] = tt[0]; ru[4] = tt[1]; ru[5] = tt[2];
			for (int q = 0; q < 6; ++q) {
				ru[q] += add[q];
				acc[0] += ru[q] * ru[q];
				acc[1] += dd[q] * ru[q];
			}
		}
	}
	const int slots[2] = {P_LS + 7 * t + 5, P_LS + 7 * t + 6};
	write_partials<2>(acc, a.part, a.pstride, slots, tix);
}

// One pass of trials, both parts in one launch: they evaluate the same trial
// points independently (the body part forms its contacts' trial forces
// itself) and write different partial slots.  Blocks [0, ntc) take the
// contact tiles, blocks [ntc, ntc + ntw) the body tiles; blockIdx.y is the
// trial of the pass.  (Looping over a pass's trials inside a workgroup, which
// shares the tile's inputs and dispatches fewer workgroups, was slower at
// every split measured: the trials of a workgroup then run one after another.)
__global__ __launch_bounds__(256) void k_ls_eval(SArgs a, const double* __restrict__ du,
                                                 const double* __restrict__ dF, const double* __restrict__ dL,
                                                 int k0, int want, int ntc) {
	// the tile's gate: a workgroup of a scene that is not in this pass's
	// state returns after one load
	if (a.lsgate[blockIdx.x] != want) return;
	if ((int)blockIdx.x < ntc) ls_contact_tile(a, du, dF, dL, k0, want, blockIdx.x);
	else ls_body_tile(a, du, dF, dL, k0, want, blockIdx.x - ntc);
}

// Acceptance (PrimalDualForceSolver.cpp:605-631) over the nt trials of a pass
// (trials k0 .. k0 + nt - 1 of a search of maxLS trials), in order: the first
// trial with (accept and gErr > 0), or the search's last trial, decides.  The
// trials of one search share the direction, so evaluating a pass at once
// selects exactly the step the one-trial-at-a-time loop would.  A decided
// regular search is accepted (LS_ACCEPTED, ls_step) or falls back to gradient
// descent (LS_GD_PENDING, PrimalDualForceSolver.cpp:441-467); a decided
// gradient-descent search ends in LS_GD_DECIDED with ls_step 0 on failure.
__global__ __launch_bounds__(1024) void k_ls_final(SArgs a, int k0, int nt, int maxLS, int gd) {
	__shared__ double red[7 * kTrials];
	const int sc = a.act[blockIdx.x];
	if (!a.ctl[sc].act || !a.ss[sc].run || a.ss[sc].ls_state != (gd ? LS_GD_PENDING : LS_PENDING)) return;
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	for (int q = wid; q < 7 * nt; q += 16) {
		const int slot = P_LS + q;
		const int c = q - 7 * (q / 7);
		const int t0 = c < 5 ? a.tbc[sc] : a.tbw[sc], t1 = c < 5 ? a.tbc[sc + 1] : a.tbw[sc + 1];
		const double v = wave_sum_dpp(partial_lane_sum(a.part + slot * (size_t)a.pstride + t0, t1 - t0));
		if (lane == 0) red[q] = v;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		SceneSc* S = &a.ss[sc];
		const double sErr = S->sErr, sacErr = S->sacErr;
		for (int t = 0; t < nt && k0 + t < maxLS; ++t) {
			const double* R = &red[7 * t];
			const double inErr = R[5] + R[0] + R[1];
			const double sacIn = R[5] + R[2] + R[1];
			const double gErr = -R[6] + R[3] + R[4];
			const bool acpt = (inErr < sErr) | (sacIn < sacErr);
			S->trials += 1;
			if ((acpt && gErr > 0) || k0 + t == maxLS - 1) {
				if (acpt) {
					S->ls_step = ldexp(1.0, -(k0 + t));
					S->t_inErr = inErr;
					S->t_sacIn = sacIn;
					S->t_gErr = gErr;
				} else {
					S->ls_step = 0;
				}
				if (gd) {
					S->ls_state = LS_GD_DECIDED;
				} else if (acpt) {
					S->ls_state = LS_ACCEPTED;
				} else {
					S->ls_state = LS_GD_PENDING;
					S->gd_count += 1;
				}
				break;
			}
		}
	}
	__syncthreads();
	write_ls_gates(a, sc, a.ss[sc].ls_state);
}

// Commit the accepted trial point of the scene's search (the regular
// direction, or the gradient-descent one after a fallback): f, lambda in
// place here, v in commit_v_tile.
__device__ __forceinline__ void ls_commit_tile(const SArgs& a, const double* __restrict__ dF,
                                               const double* __restrict__ dL, const double* __restrict__ gF,
                                               const double* __restrict__ gL, int tix) {
	const Tile tile = a.tc[tix];
	if (tile.scene < 0) return;
	const int scn = tile.scene;
	if (!a.ss[scn].run) return;
	const double step = a.ss[scn].ls_step, mu = a.ctl[scn].mu;
	if (!(step > 0)) return;
	const bool g = a.ss[scn].ls_state == LS_GD_DECIDED;
	if ((int)threadIdx.x < tile.count) {
		const int i = tile.first + threadIdx.x;
		double fI[3], lI[2], dfi[3], dli[2];
		trial_point(a, i, g ? gF : dF, g ? gL : dL, step, mu, fI, lI, dfi, dli);
		a.f[i] = fI[0];
		a.lam[i] = lI[0];
		if (a.nsc) {
			a.f[a.cs + i] = fI[1];
			a.f[2 * a.cs + i] = fI[2];
			a.lam[a.cs + i] = lI[1];
		}
	}
}

// After a gradient-descent fallback the increment becomes -ru, as in the
// reference (:451), and so does the next PCG's start.
__device__ __forceinline__ void commit_v_tile(const SArgs& a, const double* __restrict__ gdu, int tix) {
	const Tile tile = a.tb[tix];
	if (tile.scene < 0) return;
	const int scn = tile.scene;
	if (!a.ss[scn].run) return;
	const double step = a.ss[scn].ls_step;
	const bool gd = a.ss[scn].ls_state == LS_GD_DECIDED;
	if ((int)threadIdx.x < tile.count) {
		const int r = tile.first + threadIdx.x;
		for (int k = 0; k < 6; ++k) {
			const double d = gd ? gdu[6 * r + k] : a.du[6 * r + k];
			if (step > 0) a.v[6 * r + k] = a.v[6 * r + k] + 0.99 * (step * d);
			if (gd) a.du[6 * r + k] = d;
		}
	}
}

// Both commits in one launch (the contact part writes f and lambda, the body
// part v and du; neither reads what the other writes): blocks [0, ntc) take
// the contact tiles, blocks [ntc, ntc + ntb) the body tiles.
__global__ __launch_bounds__(256) void k_commit(SArgs a, const double* __restrict__ dF,
                                                const double* __restrict__ dL, const double* __restrict__ gF,
                                                const double* __restrict__ gL, const double* __restrict__ gdu,
                                                int ntc) {
	if ((int)blockIdx.x < ntc) ls_commit_tile(a, dF, dL, gF, gL, blockIdx.x);
	else commit_v_tile(a, gdu, blockIdx.x - ntc);
}

// End of a Newton iteration per scene (PrimalDualForceSolver.cpp:468-490):
// the Hu schedule, the PCG counters, the iteration count and the cap.
__global__ void k_newton_end(SArgs a, int maxit, int iterations) {
	const int sc = blockIdx.x * blockDim.x + threadIdx.x;
	if (sc >= a.S || !a.ctl[sc].act || !a.ss[sc].run) return;
	SceneSc* S = &a.ss[sc];
	const PcgScene P = a.ps[sc];
	const int iters = P.done ? P.iters : maxit;
	S->pcg_total += iters;
	if (iters == maxit) S->pcg_hits += 1;
	if (P.done && P.bad) S->nonfinite += 1;
	// The x1.3 rule follows a capped iterative solve, which the reference runs
	// only for 2000 or more bodies in contact; below that it factorises, its
	// iteration count stays 0 and a full step always gives Hu / 4
	// (PrimalDualForceSolver.cpp:379-394, 473-479).
	const bool ref_iterative = a.lbeg[sc + 1] - a.lbeg[sc] >= 2000;
	double Hu = S->Hu;
	const bool useGD = S->ls_state == LS_GD_DECIDED;
	if (!useGD && S->ls_step == 1.0) Hu = (iters == maxit && ref_iterative) ? Hu * 1.3 : Hu / 4;
	else Hu *= 4;
	S->Hu = fmax(Hu, 1e-6);
	S->it += 1;
	if (S->it >= iterations) S->run = 0;
	a.ps[sc].run = S->run;
}

// Start of a solve per scene: the Newton state of the scenes with contacts.
__global__ void k_scene_init(SArgs a) {
	const int sc = blockIdx.x * blockDim.x + threadIdx.x;
	if (sc >= a.S) return;
	SceneSc* S = &a.ss[sc];
	S->run = a.ctl[sc].act;
	S->it = 0;
	S->Hu = 1e-4;
	S->ls_state = LS_PENDING;
	S->ls_step = 0;
	S->gd_count = 0;
	S->trials = 0;
	S->pcg_total = 0;
	S->pcg_hits = 0;
	S->nonfinite = 0;
	S->exitErr = 0;
	a.ps[sc].run = S->run;
}

// ---- write-back (PrimalDualForceSolver.cpp:491-510) --------------------------
__global__ __launch_bounds__(256) void k_writeback_bodies(SArgs a, double* __restrict__ acc) {
	PDC_TILE(a.tb);
	const double s = a.ss[scn].char_speed;
	if ((int)threadIdx.x < tile.count) {
		const int r = tile.first + threadIdx.x;
		const int g = a.l2g[r];
		for (int k = 0; k < 6; ++k) acc[6 * g + k] = a.v[6 * r + k] * s - a.vphys[6 * r + k];
	}
}
__global__ __launch_bounds__(256) void k_writeback_contacts(SArgs a, double* __restrict__ lf) {
	PDC_TILE(a.tc);
	const double s = a.ss[scn].char_speed, m = a.ss[scn].char_mass;
	if ((int)threadIdx.x < tile.count) {
		const int i = tile.first + threadIdx.x;
		lf[i] = a.f[i] * s * m;
		lf[a.cs + i] = a.nsc ? a.f[a.cs + i] * s * m : 0.0;
		lf[2 * a.cs + i] = a.nsc ? a.f[2 * a.cs + i] * s * m : 0.0;
	}
}

__global__ __launch_bounds__(256) void k_acc_default(int N, double* __restrict__ acc, double g0, double g1, double g2) {
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= N) return;
	acc[6 * b] = g0; acc[6 * b + 1] = g1; acc[6 * b + 2] = g2;
	acc[6 * b + 3] = 0; acc[6 * b + 4] = 0; acc[6 * b + 5] = 0;
}

// ---- integration (DynamicSystem.cpp:137-161) ---------------------------------
__global__ __launch_bounds__(256) void k_integrate(int N, const int* __restrict__ stat, const int* __restrict__ bscene,
                            const int* __restrict__ stepped, const double* __restrict__ acc,
                            double* __restrict__ pos, double* __restrict__ quat,
                            double* __restrict__ vel, double* __restrict__ ang, double dt) {
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= N || !stepped[bscene[b]]) return;
	double v[3] = {vel[3 * b], vel[3 * b + 1], vel[3 * b + 2]};
	double w[3] = {ang[3 * b], ang[3 * b + 1], ang[3 * b + 2]};
	if (!stat[b]) {
		for (int k = 0; k < 3; ++k) {
			w[k] += acc[6 * b + 3 + k];
			v[k] += acc[6 * b + k];
		}
	}
	const double ri[3] = {w[0] * dt, w[1] * dt, w[2] * dt};
	const double rsq = dot3(ri, ri);
	double qi[4] = {0, 0, 0, 1};
	if (rsq > 0) {
		const double an = sqrt(rsq);
		const double ax[3] = {ri[0] / an, ri[1] / an, ri[2] / an};
		const double sh = sin(0.5 * an);
		qi[0] = sh * ax[0]; qi[1] = sh * ax[1]; qi[2] = sh * ax[2];
		qi[3] = cos(0.5 * an);
	}
	for (int k = 0; k < 3; ++k) pos[3 * b + k] += v[k] * dt;
	double q[4] = {quat[4 * b], quat[4 * b + 1], quat[4 * b + 2], quat[4 * b + 3]};
	double o[4];
	quat_mul(qi, q, o);
	for (int k = 0; k < 4; ++k) quat[4 * b + k] = o[k];
	for (int k = 0; k < 3; ++k) { vel[3 * b + k] = v[k]; ang[3 * b + k] = w[k]; }
}


// ===========================================================================
// host orchestration
// ===========================================================================
static SArgs make_args(pdc_context* ctx, double dt, const double* g) {
	SArgs a;
	const pdc_params& p = ctx->p;
	a.C = ctx->C; a.cs = ctx->c_stride; a.NL = ctx->NL; a.S = ctx->S;
	a.act = ctx->d_act.p;
	a.lsgate = ctx->ls_gate.p;
	a.ntc = ctx->n_tc;
	a.pstride = std::max(1, std::max(ctx->n_tc, std::max(ctx->n_tb, ctx->n_tw)));
	a.nsc = p.force_model == PDC_NON_SMOOTH_CONTACT;
	a.restitution = p.restitution;
	a.spring = p.spring_based_force;
	const double damp = p.spring_damp * p.spring_k;
	const double tk = dt * p.spring_k;
	a.compliance = 1. / (dt * (tk + damp));
	a.reduction = tk / (tk + damp);
	a.dt = dt;
	for (int k = 0; k < 3; ++k) a.gdt[k] = g[k] * dt;
	a.tc = ctx->t_c.p; a.tb = ctx->t_b.p; a.tw = ctx->t_w.p; a.tp = ctx->t_p.p;
	a.tbc = ctx->tb_c.p; a.tbb = ctx->tb_b.p; a.tbw = ctx->tb_w.p;
	a.cbeg = ctx->r_cbeg.p; a.lbeg = ctx->r_lbeg.p;
	a.ss = ctx->d_ss.p; a.ctl = ctx->d_sctl.p; a.ps = ctx->d_ps.p; a.prog = ctx->d_prog;
	a.li0 = ctx->li0.p; a.li1 = ctx->li1.p; a.T = ctx->c_T.p; a.depth = ctx->c_depth.p;
	a.pstart = ctx->ap_start.p; a.pa = ctx->ap_a.p; a.g2l = ctx->g2l.p;
	a.lstat = ctx->l_static.p; a.l2g = ctx->l2g.p;
	a.inc_ptr = ctx->inc_ptr.p; a.inc = ctx->inc.p; a.pair_adj = ctx->pair_adj.p; a.adj_blk = ctx->adj_blk.p;
	a.slot = ctx->l_slot.p;
	a.v = ctx->l_v.p; a.vold = ctx->l_vold.p; a.vphys = ctx->l_vphys.p; a.Ms = ctx->l_Ms.p;
	a.Minv = ctx->l_Minv.p; a.speed = ctx->l_speed.p; a.ru = ctx->l_ru.p; a.rhs = ctx->l_rhs.p;
	a.D = ctx->l_D.p; a.du = ctx->l_du.p; a.xs = ctx->x64.p; a.ys = ctx->y64.p;
	a.K = ctx->s_K.p; a.f = ctx->s_f.p; a.lam = ctx->s_lam.p; a.c = ctx->s_c.p; a.g = ctx->s_g.p;
	a.H = ctx->s_H.p; a.rf = ctx->s_rf.p; a.rl = ctx->s_rl.p; a.sf = ctx->s_sf.p;
	a.invC = ctx->s_invC.p; a.W = ctx->s_W.p; a.kf = ctx->s_kf.p; a.df = ctx->s_df.p;
	a.dl = ctx->s_dl.p;
	a.part = ctx->partials.p;
	return a;
}

// One recorded chunk of `chunk` PCG iterations (launches alternating the
// counter slots 0, 1; the chunk length, pdc_context::pcg_chunk(), is even).
template <class T>
static void pcg_chunk_launches(pdc_context* ctx, const PcgArgs<T>& pa, int nblk, bool two, int chunk, hipEvent_t* ev) {
	hipStream_t s = ctx->stream;
	const dim3 B(kPcgThreads);
	for (int k = 0; k < chunk; ++k) {
		if (ev) PDC_HIP(hipEventRecord(ev[2 * k], s));
		// the standard form keeps four 6-vectors per entry in flight: above two
		// entries per lane the generic width reads the rest after the barrier
		if (two && pa.one_W > 0) switch (pa.R / 8) {
			case 1: hipLaunchKernelGGL((k_pcgs_iter<T, 1, true>), dim3(nblk), B, 0, s, pa, k & 1); break;
			case 2: hipLaunchKernelGGL((k_pcgs_iter<T, 2, true>), dim3(nblk), B, 0, s, pa, k & 1); break;
			default: hipLaunchKernelGGL((k_pcgs_iter<T, 0, true>), dim3(nblk), B, 0, s, pa, k & 1); break;
		}
		else if (two) switch (pa.R / 8) {
			case 1: hipLaunchKernelGGL((k_pcgs_iter<T, 1, false>), dim3(nblk), B, 0, s, pa, k & 1); break;
			case 2: hipLaunchKernelGGL((k_pcgs_iter<T, 2, false>), dim3(nblk), B, 0, s, pa, k & 1); break;
			default: hipLaunchKernelGGL((k_pcgs_iter<T, 0, false>), dim3(nblk), B, 0, s, pa, k & 1); break;
		}
		else if (pa.one_W > 0) switch (pa.R / 8) {
			case 1: hipLaunchKernelGGL((k_pcg_iter<T, 1, true>), dim3(nblk), B, 0, s, pa, k & 1); break;
			case 2: hipLaunchKernelGGL((k_pcg_iter<T, 2, true>), dim3(nblk), B, 0, s, pa, k & 1); break;
			case 3: hipLaunchKernelGGL((k_pcg_iter<T, 3, true>), dim3(nblk), B, 0, s, pa, k & 1); break;
			case 4: hipLaunchKernelGGL((k_pcg_iter<T, 4, true>), dim3(nblk), B, 0, s, pa, k & 1); break;
			default: hipLaunchKernelGGL((k_pcg_iter<T, 0, true>), dim3(nblk), B, 0, s, pa, k & 1); break;
		}
		else switch (pa.R / 8) {
			case 1: hipLaunchKernelGGL((k_pcg_iter<T, 1, false>), dim3(nblk), B, 0, s, pa, k & 1); break;
			case 2: hipLaunchKernelGGL((k_pcg_iter<T, 2, false>), dim3(nblk), B, 0, s, pa, k & 1); break;
			case 3: hipLaunchKernelGGL((k_pcg_iter<T, 3, false>), dim3(nblk), B, 0, s, pa, k & 1); break;
			case 4: hipLaunchKernelGGL((k_pcg_iter<T, 4, false>), dim3(nblk), B, 0, s, pa, k & 1); break;
			default: hipLaunchKernelGGL((k_pcg_iter<T, 0, false>), dim3(nblk), B, 0, s, pa, k & 1); break;
		}
		if (ev) PDC_HIP(hipEventRecord(ev[2 * k + 1], s));
	}
	PDC_CHECK_LAUNCH();
}


// Arguments of the PCG kernels for the current scene set (grid fixed per
// scene set: the capacity of PCG tiles, so that a recorded chunk stays valid).
// Workgroups of a PCG launch: the PCG tiles of the scenes in this solve (the
// scenes of one step call), not the tile capacity of the whole scene set, so
// that stepping a few of many resident scenes launches only their tiles.
static int pcg_grid(const pdc_context* ctx) { return std::max(1, ctx->n_tr_act); }
// Workgroups of a launch of the PCG kernels: those tiles rounded up to a
// multiple of the eight XCDs (pcg_tile_id).
static int pcg_launch_grid(const pdc_context* ctx) { return (pcg_grid(ctx) + 7) / 8 * 8; }

template <class T>
static PcgArgs<T> pcg_args(pdc_context* ctx, DBuf<T>& Bm, DBuf<T>& P, DBuf<T>& b, DBuf<T>& G) {
	const int S = ctx->S;
	const int nblk = std::max(1, ctx->cap_tr);
	const size_t ncap = 6 * (size_t)kPcgRowsPerWG * nblk, pcap = (size_t)nblk;
	DBuf<double>& work = ctx->pcg64;
	work.ensure(8 * ncap + 13 * pcap + 4 * (size_t)S + 64);
	PcgArgs<T> pa;
	std::memset(&pa, 0, sizeof(pa));  // compared bytewise with the recorded arguments
	pa.ecol = ctx->e_col.p;
	pa.R = ctx->ell_R;
	pa.B = Bm.p;
	pa.P = P.p;
	pa.G = G.p;
	pa.gsz = ctx->l_gsz.p;
	pa.b = b.p;
	pa.x = ctx->x64.p;
	double* w = work.p;
	pa.r = w; pa.w[0] = w + ncap; pa.w[1] = w + 2 * ncap;
	pa.z = w + 3 * ncap; pa.q = w + 4 * ncap; pa.s = w + 5 * ncap; pa.p = w + 6 * ncap;
	double* parts = w + 8 * ncap;
	pa.gam[0] = parts; pa.gam[1] = parts + pcap;
	pa.del[0] = parts + 2 * pcap; pa.del[1] = parts + 3 * pcap;
	pa.rr[0] = parts + 4 * pcap; pa.rr[1] = parts + 5 * pcap;
	pa.bb = parts + 6 * pcap;
	pa.e1[0] = parts + 7 * pcap; pa.e1[1] = parts + 8 * pcap;
	pa.e2[0] = parts + 9 * pcap; pa.e2[1] = parts + 10 * pcap;
	pa.e1b[0] = parts + 11 * pcap; pa.e1b[1] = parts + 12 * pcap;
	pa.hist = parts + 13 * pcap;
	pa.pstride = (int)pcap;
	pa.ntiles = pcg_grid(ctx);
	pa.one_t0 = ctx->one_t0;
	pa.one_W = ctx->one_W;
	pa.one_scn = ctx->one_scn;
	// the standard form keeps r and q (own rows) in the recurrence form's r
	// and q, and z, w, p by parity in its z, s / w[0], w[1] / p and an eighth vector
	pa.zv[0] = pa.z; pa.zv[1] = pa.s;
	pa.wv[0] = pa.w[0]; pa.wv[1] = pa.w[1];
	pa.pv[0] = pa.p; pa.pv[1] = w + 7 * ncap;
	pa.tiles = ctx->t_pcg.p;
	pa.ss = ctx->d_ss.p;
	pa.ps = ctx->d_ps.p;
	pa.prog = ctx->d_prog;
	if (!pa.prog) throw HipError("PCG progress record not allocated (no scene table)", PDC_ERR_STATE);
	pa.tol2 = ctx->p.pcg_tol * ctx->p.pcg_tol;
	pa.maxit = ctx->p.pcg_max_iter;
	return pa;
}

// Start of the PCG of every scene still in its Newton loop: initial residual
// and the first partial sums.
template <class T>
static void pcg_begin(pdc_context* ctx, const PcgArgs<T>& pa) {
	const int nblk = pcg_launch_grid(ctx);
	hipStream_t s = ctx->stream;
	if (ctx->p.pcg_tol < kSingleReductionMinTol) {
		hipLaunchKernelGGL(k_pcgs_start<T>, dim3(nblk), dim3(kPcgThreads), 0, s, pa, (const double*)ctx->y64.p);
		hipLaunchKernelGGL(k_pcgs_init<T>, dim3(nblk), dim3(kPcgThreads), 0, s, pa);
	} else {
		hipLaunchKernelGGL(k_pcg_start<T>, dim3(nblk), dim3(kPcgThreads), 0, s, pa, (const double*)ctx->y64.p);
		hipLaunchKernelGGL(k_pcg_init<T>, dim3(nblk), dim3(kPcgThreads), 0, s, pa);
	}
	PDC_CHECK_LAUNCH();
}

// Algorithmic bytes of the scenes' active PCG iterations for the kernel
// timing hooks: per iteration of a scene (one launch of k_pcg_iter, or of
// k_pcgs_iter for the standard form), each input read once and each output
// written once, FP64 vectors:
//   k_pcg_iter   B blocks and column indices nnzb (36 T + 4), row pointers
//                4 (NL + 1), P blocks 36 T NL, w, z, s, p, x, r read
//                6 x 8 n6, z, s, p, x, r, w' written 6 x 8 n6
//   k_pcgs_iter  A blocks, column indices, row pointers, x, r, q, z, w, p
//                read 6 x 8 n6, written 6 x 8 n6, and the groups' inverse
//                blocks: 36 T g^2 per group of g bodies
template <class T>
static void pcg_account_bytes(pdc_context* ctx, const std::vector<int>& iters, bool two) {
	const double Tb = sizeof(T);
	for (int sc = 0; sc < ctx->S; ++sc) {
		if (iters[sc] <= 0) continue;
		const double NLs = ctx->h_lbeg[sc + 1] - ctx->h_lbeg[sc], n6 = 6.0 * NLs, nnzb = ctx->h_nnzb[sc];
		double pblocks = NLs;
		if (two && ctx->h_gcnt.size() >= 4 * (size_t)ctx->S) {
			const int* c = &ctx->h_gcnt[4 * (size_t)sc];
			pblocks = 64.0 * c[0] + 16.0 * c[1] + 4.0 * c[2] + c[3];
		}
		const double mat = nnzb * (36.0 * Tb + 4.0) + 4.0 * (NLs + 1) + 36.0 * Tb * pblocks;
		ctx->kt.bytes += iters[sc] * (mat + 96.0 * n6);
	}
}

// The recorded chunk of PCG launches for these arguments (recorded again
// when a buffer moved, the grid changed or the chunk length changed).
template <class T>
static PcgGraph& pcg_graph(pdc_context* ctx, const PcgArgs<T>& pa) {
	const int nblk = pcg_launch_grid(ctx);
	hipStream_t s = ctx->stream;
	const bool two = ctx->p.pcg_tol < kSingleReductionMinTol;
	const int chunk = ctx->pcg_chunk();
	PcgGraph& G = ctx->pcg_graph[(sizeof(T) == 8 ? 1 : 0) + (two ? 2 : 0)];
	static_assert(sizeof(PcgArgs<T>) <= sizeof(G.args), "recorded arguments fit");
	if (ctx->use_graphs && (G.exec == nullptr || G.nblk != nblk || G.chunk != chunk ||
	                        std::memcmp(G.args, &pa, sizeof(pa)) != 0)) {
		if (G.exec) { (void)hipGraphExecDestroy(G.exec); G.exec = nullptr; }
		if (G.graph) { (void)hipGraphDestroy(G.graph); G.graph = nullptr; }
		PDC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
		pcg_chunk_launches<T>(ctx, pa, nblk, two, chunk, nullptr);
		PDC_HIP(hipStreamEndCapture(s, &G.graph));
		PDC_HIP(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
		G.nblk = nblk;
		G.chunk = chunk;
		std::memcpy(G.args, &pa, sizeof(pa));
	}
	return G;
}

// The PCG iterations of every scene still in its Newton loop, after
// pcg_begin; returns whether any scene is (the scenes' run flags, set by
// k_err_final, arrive with the PCG done flags).  The host waits for the
// stream after every group of chunks.  Used with the kernel timing hooks
// and without graphs; pcg_solve_async below otherwise.
template <class T>
static bool pcg_solve(pdc_context* ctx, const PcgArgs<T>& pa) {
	const int S = ctx->S;
	const int nblk = pcg_launch_grid(ctx);
	hipStream_t s = ctx->stream;
	const bool two = ctx->p.pcg_tol < kSingleReductionMinTol;
	const int chunk = ctx->pcg_chunk();
	PcgGraph& G = pcg_graph<T>(ctx, pa);
	// Kernel timing of k_pcg_iter.  SPMV: launched directly between an event
	// pair per launch (events cannot be timed inside a replayed graph), only
	// launches of active iterations counted.  PCG: an event pair around every
	// group of replayed chunks, every launch counted (the launches after
	// convergence, which return at once, included).
	const bool kt_spmv = ctx->kt.which == PDC_KT_SPMV;
	const bool kt_iter = ctx->kt.which == PDC_KT_PCG;
	const bool graph = ctx->use_graphs && !kt_spmv;
	hipEvent_t* ev = kt_spmv ? ctx->kt_direct_ev(2 * chunk) : kt_iter ? ctx->kt_direct_ev(2) : nullptr;
	const int maxit = ctx->p.pcg_max_iter;
	int launched = 0;
	// The iteration count of a solve scatters widely around the previous
	// one's: the first group covers 0.8 of it and every further group a
	// quarter of what was launched, which balances the launches that find
	// the solve converged (a few us each) against the host round trips.
	auto chunks_for = [&](double launches) { return std::max(1, (int)std::ceil(launches / chunk)); };
	int want = chunks_for(0.8 * std::min(ctx->pcg_last_iters, maxit));
	ctx->h_ps.resize(S);
	double kt_ms = 0;
	bool all_done = false;
	while (launched < maxit) {
		if (kt_spmv) want = 1;  // the event pool holds one chunk
		if (kt_iter) PDC_HIP(hipEventRecord(ev[0], s));
		const int before = launched;
		for (int c = 0; c < want && launched < maxit; ++c) {
			if (graph) PDC_HIP(hipGraphLaunch(G.exec, s));
			else pcg_chunk_launches<T>(ctx, pa, nblk, two, chunk, kt_spmv ? ev : nullptr);
			launched += chunk;
		}
		if (kt_iter) PDC_HIP(hipEventRecord(ev[1], s));
		PDC_HIP(hipMemcpyAsync(ctx->h_ps.data(), ctx->d_ps.p, S * sizeof(PcgScene), hipMemcpyDeviceToHost, s));
		PDC_HIP(hipStreamSynchronize(s));
		all_done = true;
		for (int sc = 0; sc < S; ++sc)
			if (ctx->h_ps[sc].run && !ctx->h_ps[sc].done) all_done = false;
		if (kt_iter) {
			float e = 0;
			PDC_HIP(hipEventElapsedTime(&e, ev[0], ev[1]));
			kt_ms += e;
			ctx->kt.launches += launched - before;
		} else if (kt_spmv) {
			// launches up to the last active iteration of the scene that ran longest
			int active = 0;
			for (int sc = 0; sc < S; ++sc)
				if (ctx->h_ps[sc].run)
					active = std::max(active, ctx->h_ps[sc].done ? ctx->h_ps[sc].iters - before : chunk);
			active = std::min(std::max(active, 0), chunk);
			for (int k = 0; k < active; ++k) {
				float e = 0;
				PDC_HIP(hipEventElapsedTime(&e, ev[2 * k], ev[2 * k + 1]));
				kt_ms += e;
			}
			ctx->kt.launches += active;
		}
		if (all_done) break;  // every running scene converged
		want = chunks_for(0.25 * launched);
	}
	std::vector<int> iters(S, 0);
	int last = 0;
	bool any = false;
	for (int sc = 0; sc < S; ++sc) {
		if (!ctx->h_ps[sc].run) continue;
		any = true;
		iters[sc] = ctx->h_ps[sc].done ? ctx->h_ps[sc].iters : maxit;
		last = std::max(last, iters[sc]);
	}
	if (any) ctx->pcg_last_iters = last;
	if (kt_spmv || kt_iter) {
		ctx->kt.total_ms += kt_ms;
		for (int sc = 0; sc < S; ++sc)
			if (!ctx->h_ps[sc].run) iters[sc] = 0;
		pcg_account_bytes<T>(ctx, iters, two);
	}
	return any;
}

// The PCG of one Newton iteration without waiting for the stream.  The
// kernels write each scene's progress into pinned host memory (PcgProgress:
// its Newton state after k_err_final, the iterations started, |r|^2/|b|^2,
// the outcome), and the host keeps groups of launches queued a little ahead
// of the device's position: up to two groups, and no more than the iterations
// the slowest solving scene still needs (the previous solve's count at
// first, then the residual's decay since the solve's first reading,
// extrapolated).  The queue does not drain while a scene is solving, so
// there is no round trip between groups, and once every running scene has
// converged the caller queues the next phase at once, behind the few queued
// launches that find the solve converged.  `running` (per scene, in/out):
// the scenes still in their Newton loop; a scene leaves when its Newton
// state for iteration `iter` says so.  Returns whether any scene still runs.
//
// Termination does not depend on the progress records arriving: when the
// stream is idle and a scene has not reported an outcome, another group is
// queued, and at the iteration cap the host waits for the stream, after
// which every record is final.
template <class T>
static bool pcg_solve_async(pdc_context* ctx, const PcgArgs<T>& pa, int iter, std::vector<char>& running) {
	const int S = ctx->S;
	hipStream_t s = ctx->stream;
	const int chunk = ctx->pcg_chunk();
	const int nblk = pcg_launch_grid(ctx);
	const int maxit = ctx->p.pcg_max_iter;
	const bool two = ctx->p.pcg_tol < kSingleReductionMinTol;
	const bool kt_iter = ctx->kt.which == PDC_KT_PCG;
	const int ahead_max = 2 * chunk, ahead_min = 4;
	volatile PcgProgress* P = ctx->h_prog;
	int launched = 0;
	// direct launches, a group of `chunk` at a time: replaying recorded
	// chunks was slower (a boundary between replayed graphs costs more than
	// a launch, and graphs long enough to amortise it queue more launches
	// past the end of the solve)
	// (groups of an even length, so that the launches keep alternating the
	// counter slots; the last group of a target is cut to it, so that fewer
	// launches run past the end of the solve)
	auto enqueue_to = [&](int target) {
		target = std::min(target, maxit);
		while (launched < target) {
			const int n = std::min(chunk, (target - launched + 1) & ~1);
			pcg_chunk_launches<T>(ctx, pa, nblk, two, n, nullptr);
			launched += n;
		}
	};
	// kernel timing: one event pair around all the solve's launches (the
	// events add no gap between them)
	hipEvent_t kt_e0 = nullptr;
	if (kt_iter) kt_begin(ctx, PDC_KT_PCG, &kt_e0);
	const int last = std::min(ctx->pcg_last_iters, maxit);
	enqueue_to(std::max(ahead_min, std::min(last, ahead_max)));
	std::vector<int> iters(S, -1), a_it(S, -1);
	std::vector<double> a_rel(S, 0.0);
	int last_pos = -1, idle_polls = 0;
	auto moved = std::chrono::steady_clock::now();
	bool synced = false;
	for (;;) {
		bool pending = false, any = false;
		int pos = 0, total = 0;
		for (int sc = 0; sc < S; ++sc) {
			if (!running[sc]) continue;
			const long long tag = P[sc].runtag;
			const bool fresh = (int)(tag >> 32) == iter;
			if (fresh && !(tag & 1)) {
				running[sc] = 0;  // left the Newton loop at this iteration's exit test
				continue;
			}
			any = true;
			const long long oc = P[sc].outcome;
			if (oc >> 32) {
				iters[sc] = (int)(oc & 0xffffffffLL);
				continue;
			}
			pending = true;
			const int it = P[sc].it;
			const double rel = P[sc].rel2;
			pos = std::max(pos, it);
			// iterations this solve needs in all: the previous solve's count
			// until the residual's decay can be extrapolated
			int need = std::max(last, it + ahead_min);
			if (it > 0 && a_it[sc] < 0) { a_it[sc] = it; a_rel[sc] = rel; }
			if (a_it[sc] >= 0 && it >= a_it[sc] + 8 && rel > 0 && rel < a_rel[sc]) {
				const double rate = std::log(rel / a_rel[sc]) / (it - a_it[sc]);
				const double goal = std::log(std::max(ctx->p.pcg_tol * ctx->p.pcg_tol, 1e-300) / rel);
				need = it + (goal >= 0 ? 1 : (int)std::ceil(1.05 * goal / rate)) + 2;
			}
			total = std::max(total, need);
		}
		if (!any) {
			if (kt_iter) kt_end_span(ctx, kt_e0, launched);
			return false;
		}
		if (!pending) break;
		if (launched >= maxit) {
			if (synced) break;  // every record is final: scenes without an outcome reached the cap
			PDC_HIP(hipStreamSynchronize(s));
			synced = true;
			continue;
		}
		// queue up to ahead_max launches past the device, but not past the
		// estimated end (at least ahead_min past the device)
		const int target = std::max(pos + ahead_min, std::min(pos + ahead_max, total));
		if (launched < target) enqueue_to(target);
		// the device's position has not moved for a while: queue more if the
		// stream has drained; end the call if the stream failed or stayed
		// silent past the limit
		if (pos == last_pos) {
			if (++idle_polls >= 256) {
				idle_polls = 0;
				const hipError_t q = hipStreamQuery(s);
				const StreamState st = q == hipSuccess ? StreamState::Idle
				                       : q == hipErrorNotReady ? StreamState::Busy : StreamState::Failed;
				const double idle = std::chrono::duration<double>(std::chrono::steady_clock::now() - moved).count();
				switch (stall_action(st, idle, kStallLimitSeconds)) {
				case StallAction::QueueMore: enqueue_to(launched + chunk); break;
				case StallAction::Fail:
					PDC_HIP(q);
					throw HipError("PCG: no progress from the device for " + std::to_string((int)idle) + " s",
					               PDC_ERR_DEVICE);
				case StallAction::Wait: break;
				}
			}
		} else {
			last_pos = pos;
			idle_polls = 0;
			moved = std::chrono::steady_clock::now();
		}
		__builtin_ia32_pause();
	}
	int longest = 0;
	for (int sc = 0; sc < S; ++sc) {
		if (!running[sc]) { iters[sc] = 0; continue; }
		if (iters[sc] < 0) iters[sc] = maxit;
		longest = std::max(longest, iters[sc]);
	}
	ctx->pcg_last_iters = longest;
	if (kt_iter) {
		kt_end_span(ctx, kt_e0, launched);
		pcg_account_bytes<T>(ctx, iters, two);
	}
	return true;
}

// Algorithmic bytes (each input read once, each output written once).
// Constraint + contact-residual kernels of one Newton iteration, per contact:
//   f 24, lambda 16 in; c 16, g 16, H 24 out (constraints);
//   T 288, local ids 8, f 24, lambda 16, K 8 in; rf 24, rl 16 out (residuals);
// per body (gathers counted once): v 48, Minv 56, static flag 4.
double contact_terms_bytes(int C, int NL) { return 480.0 * C + 108.0 * NL; }

void write_acc_default(pdc_context* ctx, double dt, const double* g) {
	if (ctx->N == 0) return;
	hipLaunchKernelGGL(k_acc_default, dim3(grid_for(ctx->N)), dim3(kBlock), 0, ctx->stream, ctx->N,
	                   ctx->b_acc.p, g[0] * dt, g[1] * dt, g[2] * dt);
	PDC_CHECK_LAUNCH();
}

void integrate(pdc_context* ctx, double dt) {
	if (ctx->N == 0) return;
	hipLaunchKernelGGL(k_integrate, dim3(grid_for(ctx->N)), dim3(kBlock), 0, ctx->stream, ctx->N,
	                   ctx->b_static.p, ctx->b_scene.p, ctx->d_step.p, ctx->b_acc.p, ctx->b_pos.p,
	                   ctx->b_quat.p, ctx->b_vel.p, ctx->b_ang.p, dt);
	PDC_CHECK_LAUNCH();
}

// Records the launches `fn` makes on the context's stream as a graph and
// makes G.exec replay them.  The launch sequence keeps its shape from solve
// to solve (only sizes and arguments change), so the executable graph is
// updated in place; it is instantiated again when the update is refused.
template <class F>
static void record_phase(pdc_context* ctx, PcgGraph& G, F&& fn) {
	hipStream_t s = ctx->stream;
	hipGraph_t g = nullptr;
	PDC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
	fn();
	PDC_HIP(hipStreamEndCapture(s, &g));
	// The graph an executable was instantiated from stays alive as long as
	// the executable: an update takes the new capture's parameters and the
	// capture is then discarded; only a new instantiation replaces the graph.
	if (G.exec) {
		hipGraphExecUpdateResult res = hipGraphExecUpdateError;
		hipGraphNode_t bad = nullptr;
		if (hipGraphExecUpdate(G.exec, g, &bad, &res) == hipSuccess && res == hipGraphExecUpdateSuccess) {
			(void)hipGraphDestroy(g);
			return;
		}
		(void)hipGetLastError();
		(void)hipGraphExecDestroy(G.exec);
		G.exec = nullptr;
	}
	PDC_HIP(hipGraphInstantiate(&G.exec, g, nullptr, nullptr, 0));
	if (G.graph) (void)hipGraphDestroy(G.graph);
	G.graph = g;
}

// Per-scene scalars to the pinned host mirror (blocking).
static void read_scene_scalars(pdc_context* ctx) {
	PDC_HIP(hipMemcpyAsync(ctx->h_ss, ctx->d_ss.p, ctx->S * sizeof(SceneSc), hipMemcpyDeviceToHost, ctx->stream));
	PDC_HIP(hipStreamSynchronize(ctx->stream));
}

// The Newton solve of every scene that has contacts, all scenes advancing
// together: each phase runs for the scenes still taking part (SceneCtl.act),
// a scene drops out when it converges or reaches the iteration cap, and every
// decision of the reference's loop (PrimalDualForceSolver.cpp:278-490) is
// taken per scene.  A scene's arithmetic does not depend on the other scenes.
void solve(pdc_context* ctx, double dt, const double* g) {
	hipStream_t s = ctx->stream;
	const int S = ctx->S;
	ctx->scene_stats.assign(S, pdc_solve_stats{});
	write_acc_default(ctx, dt, g);
	const int C = ctx->C;
	if (C == 0) return;
	build_topology(ctx);
	const int NL = ctx->NL;
	const size_t nb = (size_t)NL;
	ctx->l_v.ensure(6 * nb); ctx->l_vold.ensure(6 * nb); ctx->l_vphys.ensure(6 * nb);
	ctx->l_Ms.ensure(7 * nb); ctx->l_Minv.ensure(7 * nb); ctx->l_speed.ensure(nb);
	ctx->l_ru.ensure(6 * nb); ctx->l_rhs.ensure(6 * nb); ctx->l_D.ensure(6 * nb);
	ctx->l_du.ensure(6 * nb); ctx->l_gdu.ensure(6 * nb);
	// PCG storage by slot (kPcgRowsPerWG per PCG tile), the operator ELL
	const size_t nslot = (size_t)kPcgRowsPerWG * std::max(1, ctx->cap_tr), nell = nslot * ctx->ell_R;
	ctx->x64.ensure(6 * nslot); ctx->y64.ensure(6 * nslot);
	const bool fp32 = ctx->p.pcg_precision == PDC_PCG_FP32;
	const bool plain = ctx->p.pcg_tol < kSingleReductionMinTol;  // the standard form's operator
	if (fp32) {
		ctx->A32.ensure(36 * nell); ctx->P32.ensure(36 * nslot); ctx->b32.ensure(6 * nslot);
		if (plain) ctx->G32.ensure(kGroupBlock * nslot);
	} else {
		ctx->A64.ensure(36 * nell); ctx->P64.ensure(36 * nslot); ctx->b64.ensure(6 * nslot);
		if (plain) ctx->G64.ensure(kGroupBlock * nslot);
	}
	SArgs a0 = make_args(ctx, dt, g);
	ctx->partials.ensure((size_t)P_COUNT * a0.pstride);
	ctx->ls_gate.ensure((size_t)std::max(1, ctx->n_tc + ctx->n_tw));
	SArgs a = make_args(ctx, dt, g);
	const int ntc = ctx->n_tc, ntb = ctx->n_tb, ntw = ctx->n_tw, ntp = ctx->n_tp;

	std::vector<char> in(S);
	for (int sc = 0; sc < S; ++sc) in[sc] = ctx->h_cbeg[sc + 1] > ctx->h_cbeg[sc];
	ctx->h_sctl.resize(S);
	std::vector<int>& act = ctx->h_act;
	act.clear();
	for (int sc = 0; sc < S; ++sc) {
		ctx->h_sctl[sc].act = in[sc] ? 1 : 0;
		ctx->h_sctl[sc].mu = ctx->h_mu[sc];
		if (in[sc]) act.push_back(sc);
	}
	PDC_HIP(hipMemcpyAsync(ctx->d_sctl.p, ctx->h_sctl.data(), S * sizeof(SceneCtl), hipMemcpyHostToDevice, s));
	// the per-scene kernels run one workgroup per scene of this solve
	const int nact = (int)act.size();
	if (nact > 0)
		PDC_HIP(hipMemcpyAsync(ctx->d_act.p, act.data(), nact * sizeof(int), hipMemcpyHostToDevice, s));
	auto launch = [&](auto kernel, int grid, int threads, auto... args) {
		if (grid > 0) hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), 0, s, args...);
	};
	// a line-search pass: one grid row per trial
	auto launch_trials = [&](auto kernel, int tiles, int trials, auto... args) {
		if (tiles > 0) hipLaunchKernelGGL(kernel, dim3(tiles, trials), dim3(kBlock), 0, s, args...);
	};
	const int gS = (S + kBlock - 1) / kBlock;

	launch(k_scene_init, gS, kBlock, a);
	launch(k_setup_bodies, ntb, kBlock, a, ctx->b_vel.p, ctx->b_ang.p, ctx->b_quat.p, ctx->b_mass.p, ctx->b_I0.p);
	launch(k_nondim_speed, ntw, kBlock, a);
	const double gn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
	launch(k_nondim_final, nact, 1024, a, (const double*)ctx->b_mass.p, gn * dt);
	launch(k_scale_bodies, ntb, kBlock, a);
	launch(k_preprocess, ntc, kBlock, a);
	PDC_CHECK_LAUNCH();

	// The Newton loop of every scene (PrimalDualForceSolver.cpp:278-490) as a
	// fixed sequence of launches per iteration: each scene's decisions (exit
	// test, line-search acceptance, gradient-descent fallback, Hu schedule,
	// iteration cap) are taken on the device, and the launches of a scene that
	// is in another state return at once.  The host learns which scenes are
	// still iterating, and when their PCG has converged, from the progress
	// records the kernels write into pinned host memory (pcg_solve_async),
	// without waiting for the stream; without graphs or with per-launch PCG
	// timing it waits for the stream once per group of PCG chunks
	// (pcg_solve).
	const double* du = ctx->l_du.p;
	const double* gdu = ctx->l_gdu.p;
	// phase A: errors, assembly and the PCG start; phase B: the increments,
	// the line search and the end of the iteration
	PcgArgs<float> pa32;
	PcgArgs<double> pa64;
	if (fp32) pa32 = pcg_args<float>(ctx, ctx->A32, ctx->P32, ctx->b32, ctx->G32);
	else pa64 = pcg_args<double>(ctx, ctx->A64, ctx->P64, ctx->b64, ctx->G64);
	auto phase_a = [&]() {
		hipEvent_t e0;
		kt_begin(ctx, PDC_KT_CONTACT_TERMS, &e0);
		launch(k_constraints, ntc, kBlock, a);
		if (e0) {
			// timed on their own: the contact terms, then the body residuals
			launch(k_contact_res, ntc, kBlock, a);
			kt_end(ctx, PDC_KT_CONTACT_TERMS, e0);
			ctx->kt.bytes += contact_terms_bytes(C, NL);
			launch(k_body_res, ntw, kBlock, a);
		} else {
			launch(k_residuals, ntc + ntw, kBlock, a, ntc);
		}
		launch(k_err_final, nact, 256, a, ctx->p.tolerance);
		PDC_CHECK_LAUNCH();
		// assembly
		kt_begin(ctx, PDC_KT_ASSEMBLY, &e0);
		launch(k_asm_prep, ntc, kBlock, a);
		// the standard form (tight tolerances) stores A itself, the recurrences A P
		const int ntr = pcg_grid(ctx);
		if (fp32 && plain) {
			launch(k_asm_diag<float, true>, ntw, kBlock, a, (const int*)ctx->a_diag.p, ctx->A32.p, ctx->P32.p, ctx->b32.p);
			launch(k_asm_off<float, false, true>, ntp, kBlock, a, (const float*)ctx->P32.p, ctx->A32.p);
			launch(k_asm_off<float, true, true>, ntp, kBlock, a, (const float*)ctx->P32.p, ctx->A32.p);
			launch(k_grp_inv<float>, ntr, kPcgThreads, (const PcgTile*)ctx->t_pcg.p, (const SceneSc*)ctx->d_ss.p,
			       (const int*)ctx->e_col.p, ctx->ell_R, (const float*)ctx->A32.p, (const int*)ctx->l_gsz.p,
			       ctx->G32.p);
		} else if (fp32) {
			launch(k_asm_diag<float, false>, ntw, kBlock, a, (const int*)ctx->a_diag.p, ctx->A32.p, ctx->P32.p, ctx->b32.p);
			launch(k_asm_off<float, false, false>, ntp, kBlock, a, (const float*)ctx->P32.p, ctx->A32.p);
			launch(k_asm_off<float, true, false>, ntp, kBlock, a, (const float*)ctx->P32.p, ctx->A32.p);
		} else if (plain) {
			launch(k_asm_diag<double, true>, ntw, kBlock, a, (const int*)ctx->a_diag.p, ctx->A64.p, ctx->P64.p, ctx->b64.p);
			launch(k_asm_off<double, false, true>, ntp, kBlock, a, (const double*)ctx->P64.p, ctx->A64.p);
			launch(k_asm_off<double, true, true>, ntp, kBlock, a, (const double*)ctx->P64.p, ctx->A64.p);
			launch(k_grp_inv<double>, ntr, kPcgThreads, (const PcgTile*)ctx->t_pcg.p, (const SceneSc*)ctx->d_ss.p,
			       (const int*)ctx->e_col.p, ctx->ell_R, (const double*)ctx->A64.p, (const int*)ctx->l_gsz.p,
			       ctx->G64.p);
		} else {
			launch(k_asm_diag<double, false>, ntw, kBlock, a, (const int*)ctx->a_diag.p, ctx->A64.p, ctx->P64.p, ctx->b64.p);
			launch(k_asm_off<double, false, false>, ntp, kBlock, a, (const double*)ctx->P64.p, ctx->A64.p);
			launch(k_asm_off<double, true, false>, ntp, kBlock, a, (const double*)ctx->P64.p, ctx->A64.p);
		}
		PDC_CHECK_LAUNCH();
		kt_end(ctx, PDC_KT_ASSEMBLY, e0);
		if (fp32) pcg_begin<float>(ctx, pa32);
		else pcg_begin<double>(ctx, pa64);
	};
	auto phase_b = [&]() {
		launch(k_du_increments, ntb + ntc, kBlock, a, (const double*)ctx->x64.p, ctx->l_gdu.p, ntb);
		// line search: passes of 2 and 8 trials in the Newton direction, then
		// the gradient-descent fallback (PrimalDualForceSolver.cpp:441-467) in
		// one pass of its 20 trials; each pass runs for the scenes still
		// searching (a pass of a decided search returns at once)
		hipEvent_t e1;
		for (int pass = 0, k0 = 0; pass < kNewtonPasses; k0 += kNewtonPass[pass], ++pass) {
			const int nt = kNewtonPass[pass];
			kt_begin(ctx, PDC_KT_LINE_SEARCH, &e1);
			launch_trials(k_ls_eval, ntc + ntw, nt, a, du, (const double*)ctx->s_df.p, (const double*)ctx->s_dl.p,
			              k0, (int)LS_PENDING, ntc);
			kt_end(ctx, PDC_KT_LINE_SEARCH, e1);
			launch(k_ls_final, nact, 1024, a, k0, nt, 10, 0);
		}
		kt_begin(ctx, PDC_KT_LINE_SEARCH, &e1);
		launch_trials(k_ls_eval, ntc + ntw, kTrials, a, gdu, (const double*)ctx->s_rf.p, (const double*)ctx->s_rl.p,
		              0, (int)LS_GD_PENDING, ntc);
		kt_end(ctx, PDC_KT_LINE_SEARCH, e1);
		launch(k_ls_final, nact, 1024, a, 0, kTrials, 20, 1);
		launch(k_commit, ntc + ntb, kBlock, a, (const double*)ctx->s_df.p, (const double*)ctx->s_dl.p,
		       (const double*)ctx->s_rf.p, (const double*)ctx->s_rl.p, gdu, ntc);
		launch(k_newton_end, gS, kBlock, a, ctx->p.pcg_max_iter, ctx->p.iterations);
		PDC_CHECK_LAUNCH();
	};
	// With graphs enabled and no per-launch timing, the host never waits for
	// the stream in the loop (pcg_solve_async) and launches the PCG
	// iterations directly.  The phases around the PCG, recorded once per
	// solve, are replayed from graphs for a set of scenes; a single scene
	// launches them directly too (replaying them measured slower there).  With
	// the per-launch PCG timing (direct PCG launches between events) the host
	// waits for each group of PCG launches (pcg_solve) and replays the phases;
	// without graphs it waits likewise and launches everything directly.
	// Kernel timing of a phase's own kernels needs direct launches.
	const int kw = ctx->kt.which;
	const bool graphs = ctx->use_graphs && kw != PDC_KT_CONTACT_TERMS && kw != PDC_KT_ASSEMBLY &&
	                    kw != PDC_KT_LINE_SEARCH;
	const bool async = graphs && (kw == PDC_KT_NONE || kw == PDC_KT_PCG);
	const bool phase_graphs = graphs && (!async || S > 1);
	// The host always queues phase B of one iteration and phase A of the next
	// together, so they are one graph (a boundary between replayed graphs
	// costs more than a launch inside one).  After the last iteration's
	// phase B the next phase A finds every scene out of its loop and its
	// launches return at once.
	if (phase_graphs) {
		record_phase(ctx, ctx->phase_graph[0], phase_a);
		record_phase(ctx, ctx->phase_graph[1], [&]() {
			phase_b();
			phase_a();
		});
	}
	std::vector<char> running(in);
	// No Newton-state record of an earlier solve (nor the zeroed record of a
	// scene never solved) may read as this solve's iteration 0: the host would
	// take the scene for one that left its loop and stop waiting for its PCG.
	// The stream is idle here (the previous solve ended with a wait).
	if (async)
		for (int sc = 0; sc < S; ++sc) ctx->h_prog[sc].runtag = -1;
	// the PCG records of the iteration about to be queued
	auto reset_records = [&]() {
		if (async)
			for (int sc = 0; sc < S; ++sc)
				if (running[sc]) {
					ctx->h_prog[sc].outcome = 0;
					ctx->h_prog[sc].it = 0;
				}
	};
	reset_records();
	if (phase_graphs) PDC_HIP(hipGraphLaunch(ctx->phase_graph[0].exec, s));
	else phase_a();
	for (int iter = 0; iter < ctx->p.iterations; ++iter) {
		bool any;
		if (async) any = fp32 ? pcg_solve_async<float>(ctx, pa32, iter, running)
		                      : pcg_solve_async<double>(ctx, pa64, iter, running);
		else any = fp32 ? pcg_solve<float>(ctx, pa32) : pcg_solve<double>(ctx, pa64);
		if (!any) break;  // every scene converged or reached the cap
		// phase B of this iteration, then phase A of the next (no PCG record
		// is written by phase B)
		reset_records();
		if (phase_graphs) {
			PDC_HIP(hipGraphLaunch(ctx->phase_graph[1].exec, s));
		} else {
			phase_b();
			if (iter + 1 < ctx->p.iterations) phase_a();
		}
	}
	launch(k_writeback_bodies, ntb, kBlock, a, ctx->b_acc.p);
	launch(k_writeback_contacts, ntc, kBlock, a, ctx->c_lf.p);
	PDC_CHECK_LAUNCH();
	read_scene_scalars(ctx);
	int itmax = 0;
	for (int sc = 0; sc < S; ++sc) {
		pdc_solve_stats& st = ctx->scene_stats[sc];
		st.n_contacts = ctx->h_cbeg[sc + 1] - ctx->h_cbeg[sc];
		if (!in[sc]) continue;
		const SceneSc& H = ctx->h_ss[sc];
		st.n_local_bodies = ctx->h_lbeg[sc + 1] - ctx->h_lbeg[sc];
		st.newton_iterations = H.it;
		st.pcg_iterations = H.pcg_total;
		st.pcg_maxiter_hits = H.pcg_hits;
		st.gd_fallbacks = H.gd_count;
		st.line_search_trials = H.trials;
		st.exit_error = H.exitErr;
		st.char_mass = H.char_mass;
		st.char_speed = H.char_speed;
		st.n_blocks = ctx->h_nnzb[sc];
		st.nonfinite = H.nonfinite;
		itmax = std::max(itmax, H.it);
	}
	if (S == 1) ctx->scene_stats[0].ties = ctx->n_ties;
	ctx->timings.newton_iterations += itmax;
	ctx->timings.contacts += C;
}

// Statistics of the last solve over the scenes: counts summed, Newton
// iterations and exit error the maximum over scenes, characteristic scales of
// the first scene with contacts.
pdc_solve_stats aggregate_stats(const pdc_context* ctx) {
	pdc_solve_stats t{};
	bool first = true;
	for (const pdc_solve_stats& st : ctx->scene_stats) {
		t.n_contacts += st.n_contacts;
		t.n_local_bodies += st.n_local_bodies;
		t.newton_iterations = std::max(t.newton_iterations, st.newton_iterations);
		t.pcg_iterations += st.pcg_iterations;
		t.pcg_maxiter_hits += st.pcg_maxiter_hits;
		t.gd_fallbacks += st.gd_fallbacks;
		t.line_search_trials += st.line_search_trials;
		t.n_blocks += st.n_blocks;
		t.nonfinite += st.nonfinite;
		t.exit_error = (t.exit_error < st.exit_error || st.exit_error != st.exit_error) ? st.exit_error : t.exit_error;
		if (first && st.n_contacts > 0) {
			t.char_mass = st.char_mass;
			t.char_speed = st.char_speed;
			first = false;
		}
	}
	t.ties = ctx->n_ties;
	return t;
}

}  // namespace pdc
