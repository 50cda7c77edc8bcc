// pdc_solver.hip — (b)-(e): the primal-dual friction Newton solve on the device.
//
// Restates PrimalDualForceSolver::fillForcesIntl (reference
// src/Solver/PrimalDual/PrimalDualForceSolver.cpp:167-561) with the
// NonSmoothContactForce / OnlyNormalForce plugins
// (src/Solver/PrimalDual/NonSmoothForces/*.cpp).  Everything is FP64 except the
// inner PCG, which runs in FP32 or FP64 (pdc_params.pcg_precision).
//
// Layout: bodies in the local solve set are numbered by ascending global id;
// per-body vectors are AoS (6 doubles per body), per-contact arrays are SoA
// with row stride ctx->c_stride.  Contact-to-body sums are gathers over a
// per-body pair adjacency (each pair's contacts are contiguous), so no
// floating-point atomics are used and every sum has a fixed order.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <cmath>

#include "pdc_internal.h"
#include "pdc_reduce.h"

namespace cg = cooperative_groups;

namespace pdc {

constexpr int kMaxWG = 1024;  // workgroups of the reducing 1-D kernels
enum {
	P_GAP = 0, P_WRF, P_RL, P_AC, P_RU,
	P_T_WRF, P_T_RL, P_T_AC, P_T_GF, P_T_GL, P_T_RU, P_T_GU,
	P_ND_TRACE, P_ND_CNT, P_COUNT
};

inline int wg_for(int64_t n) { return std::min(grid_for(n), kMaxWG); }

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

__global__ __launch_bounds__(256) void k_adj_fill(int NP, const int* __restrict__ pa, const int* __restrict__ pb,
                           const int* __restrict__ g2l, int* __restrict__ cur, int* __restrict__ nb,
                           int* __restrict__ pair) {
	const int p = blockIdx.x * blockDim.x + threadIdx.x;
	if (p >= NP) return;
	const int la = g2l[pa[p]], lb = g2l[pb[p]];
	int k = atomicAdd(&cur[la], 1);
	nb[k] = lb;
	pair[k] = p;
	k = atomicAdd(&cur[lb], 1);
	nb[k] = la;
	pair[k] = p;
}

// Sorts each adjacency row by partner (partners are unique per row), which
// makes the row order independent of the atomic fill order.
__global__ __launch_bounds__(256) void k_adj_sort(int NL, const int* __restrict__ ptr, int* __restrict__ nb,
                           int* __restrict__ pair) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= NL) return;
	const int s = ptr[r], e = ptr[r + 1];
	for (int i = s + 1; i < e; ++i) {
		const int kn = nb[i], kp = pair[i];
		int j = i - 1;
		while (j >= s && nb[j] > kn) {
			nb[j + 1] = nb[j];
			pair[j + 1] = pair[j];
			--j;
		}
		nb[j + 1] = kn;
		pair[j + 1] = kp;
	}
}

__global__ __launch_bounds__(256) void k_a_count(int NL, const int* __restrict__ l2g, const int* __restrict__ bstat,
                          const int* __restrict__ ptr, const int* __restrict__ nb,
                          int* __restrict__ lstat, int* __restrict__ cnt) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= NL) return;
	const int st = bstat[l2g[r]];
	lstat[r] = st;
	int c = 1;
	if (!st)
		for (int k = ptr[r]; k < ptr[r + 1]; ++k) c += bstat[l2g[nb[k]]] ? 0 : 1;
	cnt[r] = c;
}

// Block-CSR pattern of the Newton matrix (linsysReserve, NonSmoothForceBase.h:151-187):
// the diagonal block of every body plus one block per active-active pair.
__global__ __launch_bounds__(256) void k_a_fill(int NL, const int* __restrict__ lstat, const int* __restrict__ ptr,
                         const int* __restrict__ nb, const int* __restrict__ pair,
                         const int* __restrict__ aptr, int* __restrict__ acol,
                         int* __restrict__ apair, int* __restrict__ adiag) {
	const int r = blockIdx.x * blockDim.x + threadIdx.x;
	if (r >= NL) return;
	int o = aptr[r];
	bool diag_done = false;
	if (!lstat[r]) {
		for (int k = ptr[r]; k < ptr[r + 1]; ++k) {
			const int c = nb[k];
			if (lstat[c]) continue;
			if (!diag_done && c > r) {
				adiag[r] = o;
				acol[o] = r;
				apair[o] = -1;
				++o;
				diag_done = true;
			}
			acol[o] = c;
			apair[o] = pair[k];
			++o;
		}
	}
	if (!diag_done) {
		adiag[r] = o;
		acol[o] = r;
		apair[o] = -1;
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
	re(ctx->s_W, 4); re(ctx->s_kf, 3); re(ctx->s_df, 3); re(ctx->s_dl, 2); re(ctx->s_ft, 3);
	re(ctx->s_lt, 2);
	ctx->c_stride = (int)cs;
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
	ctx->adj_cnt.ensure(NL + 1); ctx->adj_ptr.ensure(NL + 1); ctx->adj_cur.ensure(NL + 1);
	ctx->adj_nb.ensure(2 * (size_t)NP); ctx->adj_pair.ensure(2 * (size_t)NP);
	PDC_HIP(hipMemsetAsync(ctx->adj_cnt.p, 0, (NL + 1) * sizeof(int), s));
	hipLaunchKernelGGL(k_adj_count, dim3(grid_for(NP)), dim3(kBlock), 0, s, NP, ctx->ap_a.p,
	                   ctx->ap_b.p, ctx->g2l.p, ctx->adj_cnt.p);
	PDC_CHECK_LAUNCH();
	scan_exclusive(ctx, ctx->adj_cnt.p, ctx->adj_ptr.p, NL);
	PDC_HIP(hipMemcpyAsync(ctx->adj_cur.p, ctx->adj_ptr.p, (NL + 1) * sizeof(int),
	                       hipMemcpyDeviceToDevice, s));
	hipLaunchKernelGGL(k_adj_fill, dim3(grid_for(NP)), dim3(kBlock), 0, s, NP, ctx->ap_a.p,
	                   ctx->ap_b.p, ctx->g2l.p, ctx->adj_cur.p, ctx->adj_nb.p, ctx->adj_pair.p);
	PDC_CHECK_LAUNCH();
	hipLaunchKernelGGL(k_adj_sort, dim3(grid_for(NL)), dim3(kBlock), 0, s, NL, ctx->adj_ptr.p,
	                   ctx->adj_nb.p, ctx->adj_pair.p);
	PDC_CHECK_LAUNCH();
	ctx->l_static.ensure(NL + 1); ctx->a_cnt.ensure(NL + 1); ctx->a_ptr.ensure(NL + 1);
	ctx->a_diag.ensure(NL + 1);
	hipLaunchKernelGGL(k_a_count, dim3(grid_for(NL)), dim3(kBlock), 0, s, NL, ctx->l2g.p,
	                   ctx->b_static.p, ctx->adj_ptr.p, ctx->adj_nb.p, ctx->l_static.p, ctx->a_cnt.p);
	PDC_CHECK_LAUNCH();
	scan_exclusive(ctx, ctx->a_cnt.p, ctx->a_ptr.p, NL);
	const int NNZB = read_int(ctx, ctx->a_ptr.p + NL);
	ctx->NNZB = NNZB;
	ctx->a_col.ensure(NNZB); ctx->a_pair.ensure(NNZB);
	hipLaunchKernelGGL(k_a_fill, dim3(grid_for(NL)), dim3(kBlock), 0, s, NL, ctx->l_static.p,
	                   ctx->adj_ptr.p, ctx->adj_nb.p, ctx->adj_pair.p, ctx->a_ptr.p, ctx->a_col.p,
	                   ctx->a_pair.p, ctx->a_diag.p);
	PDC_CHECK_LAUNCH();
}

// ===========================================================================
// solver kernels
// ===========================================================================
struct SArgs {
	int C, cs, NL, nsc;
	double mu, restitution, compliance, reduction, dt, gdt[3];
	int spring;
	// contacts
	const int* li0;
	const int* li1;
	const double* T;        // [36][cs]
	const double* depth;
	// topology
	const int* adj_ptr;
	const int* adj_pair;
	const int* pstart;
	const int* lstat;
	const int* l2g;
	// bodies
	double* v;
	double* vold;
	double* vphys;
	double* Ms;
	double* Minv;
	unsigned long long* speed;
	double* ru;
	double* rhs;
	double* D;
	double* du;
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
	double* ft;
	double* lt;
	double* part;
	Scalars* sc;
};

__device__ __forceinline__ double Tc(const SArgs& a, int i, int j, int k) {
	return a.T[(size_t)(j * 18 + k) * a.cs + i];
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
__device__ __forceinline__ void write_partials(double (&v)[K], double* part, const int* slots) {
	__shared__ double lds[K * 16];
	double out[K];
	block_sum<K, double>(v, lds, out);
	if (threadIdx.x == 0)
		for (int k = 0; k < K; ++k) part[slots[k] * kMaxWG + blockIdx.x] = out[k];
}

// ---- setup ----------------------------------------------------------------
__global__ __launch_bounds__(256) void k_setup_bodies(SArgs a, const double* __restrict__ vel, const double* __restrict__ ang,
                               const double* __restrict__ quat, const double* __restrict__ mass,
                               const double* __restrict__ I0) {
	double acc[2] = {0, 0};
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.NL; r += gridDim.x * blockDim.x) {
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
		a.speed[r] = 0ull;
		if (!st) {
			acc[0] += 3 * m + (Iw[0] + Iw[3] + Iw[5]);
			acc[1] += 1.0;
		}
	}
	const int slots[2] = {P_ND_TRACE, P_ND_CNT};
	write_partials<2>(acc, a.part, slots);
}

// objSpeed per body (PrimalDualForceSolver.cpp:38-63); max via ordered bits.
__global__ __launch_bounds__(256) void k_nondim_contacts(SArgs a) {
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.C; i += gridDim.x * blockDim.x) {
		const int l[2] = {a.li0[i], a.li1[i]};
		bool hasStatic = false;
		double rel[3] = {0, 0, 0};
		for (int j = 0; j < 2; ++j) {
			double ve[6];
			for (int k = 0; k < 6; ++k) ve[k] = a.vphys[6 * l[j] + k];
			if (!a.lstat[l[j]]) {
				for (int k = 0; k < 3; ++k) ve[k] += a.gdt[k];
			} else {
				hasStatic = true;
			}
			double half[3];
			for (int col = 0; col < 3; ++col) {
				double s = 0;
				for (int r = 0; r < 6; ++r) s += Tc(a, i, j, col * 6 + r) * ve[r];
				half[col] = -s;
			}
			for (int k = 0; k < 3; ++k) rel[k] += half[k];
			const double hn = sqrt(dot3(half, half));
			atomicMax(&a.speed[l[j]], (unsigned long long)__double_as_longlong(hn));
		}
		if (hasStatic) {
			const double rn = sqrt(dot3(rel, rel));
			for (int j = 0; j < 2; ++j)
				atomicMax(&a.speed[l[j]], (unsigned long long)__double_as_longlong(rn));
		}
	}
}

// charMass, charSpeed (PrimalDualForceSolver.cpp:23-78); one workgroup.
__global__ __launch_bounds__(1024) void k_nondim_final(SArgs a, int nwg, const double* __restrict__ mass, double gnorm_dt) {
	__shared__ double slot;
	__shared__ double lds[2 * 16];
	const double tr = block_reduce_array(&a.part[P_ND_TRACE * kMaxWG], nwg, &slot);
	const double cnt = block_reduce_array(&a.part[P_ND_CNT * kMaxWG], nwg, &slot);
	double v[2] = {0, 0};
	for (int r = threadIdx.x; r < a.NL; r += blockDim.x) {
		if (a.lstat[r]) continue;
		const double m = mass[a.l2g[r]];
		v[0] += m;
		v[1] += __longlong_as_double((long long)a.speed[r]) * m;
	}
	double out[2];
	block_sum<2, double>(v, lds, out);
	if (threadIdx.x == 0) {
		const double avg = cnt > 0 ? tr / cnt / 6 : 1.0;
		double s = out[1] / out[0];
		s = s > gnorm_dt * 0.25 ? s : gnorm_dt * 0.25;
		s = s > 1e-4 ? s : 1e-4;
		a.sc->char_mass = avg;
		a.sc->char_speed = s;
	}
}

// Scale velocities by charSpeed and masses by charMass; K, f, lambda init
// (NonSmoothContactForce.cpp:8-52, PrimalDualForceSolver.cpp:216-252).
__global__ __launch_bounds__(256) void k_scale_bodies(SArgs a) {
	const double s = a.sc->char_speed, m = a.sc->char_mass;
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.NL; r += gridDim.x * blockDim.x) {
		for (int k = 0; k < 6; ++k) {
			a.v[6 * r + k] /= s;
			a.vold[6 * r + k] /= s;
		}
		if (!a.lstat[r])
			for (int k = 0; k < 7; ++k) a.Ms[7 * r + k] /= m;
	}
}

__global__ __launch_bounds__(256) void k_preprocess(SArgs a) {
	const double s = a.sc->char_speed;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.C; i += gridDim.x * blockDim.x) {
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
	double acc[1] = {0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.C; i += gridDim.x * blockDim.x) {
		const double f0 = a.f[i];
		if (a.nsc) {
			const double f1 = a.f[a.cs + i], f2 = a.f[2 * a.cs + i];
			const double c0 = -f0, c1 = Lorentz::constraint(a.mu, f0, f1, f2);
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
	write_partials<1>(acc, a.part, slots);
}

// Residual terms of one contact at (v', f', lambda') (NonSmoothContactForce.cpp:53-172).
// Returns rf (3), rl (2), wrf (3), ac (3) with fresh constraint values.
__device__ __forceinline__ void contact_terms(const SArgs& a, int i, const double* va, const double* vb,
                                              const double* f, const double* lam, double muc,
                                              double* rf, double* rl, double* wrf, double* ac) {
	double u[3];
	to_local(a, i, va, vb, u);
	const double Ki = a.K[i];
	const double spring_term = a.spring ? a.compliance * a.sc->char_mass * f[0] : 0.0;
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
		rl[1] = lam[1] * Lorentz::constraint(a.mu, f[0], f[1], f[2]) + muc;
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
			fw += q * a.sc->char_mass;
		}
		const double rad = Lorentz::radius(a.mu, f[0]);
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
__global__ __launch_bounds__(256) void k_contact_res(SArgs a, int nwg_gap, double inv_nL) {
	__shared__ double slot;
	const double gapsum = block_reduce_array(&a.part[P_GAP * kMaxWG], nwg_gap, &slot);
	const double gap = -gapsum * inv_nL;
	const double muc = gap * 0.1;
	if (blockIdx.x == 0 && threadIdx.x == 0) { a.sc->gap = gap; a.sc->muc = muc; }
	double acc[3] = {0, 0, 0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.C; i += gridDim.x * blockDim.x) {
		const double f[3] = {a.f[i], a.f[a.cs + i], a.f[2 * a.cs + i]};
		const double lam[2] = {a.lam[i], a.lam[a.cs + i]};
		double rf[3], rl[2], wrf[3], ac[3];
		contact_terms(a, i, &a.v[6 * a.li0[i]], &a.v[6 * a.li1[i]], f, lam, muc, rf, rl, wrf, ac);
		for (int k = 0; k < 3; ++k) a.rf[k * a.cs + i] = rf[k];
		a.rl[i] = rl[0];
		a.rl[a.cs + i] = rl[1];
		acc[0] += wrf[0] * wrf[0] + wrf[1] * wrf[1] + wrf[2] * wrf[2];
		acc[1] += rl[0] * rl[0] + rl[1] * rl[1];
		acc[2] += ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
	}
	const int slots[3] = {P_WRF, P_RL, P_AC};
	write_partials<3>(acc, a.part, slots);
}

// Iterates the contacts incident to local body r: fn(i, side).
template <class F>
__device__ __forceinline__ void for_incident(const SArgs& a, int r, F fn) {
	for (int k = a.adj_ptr[r]; k < a.adj_ptr[r + 1]; ++k) {
		const int p = a.adj_pair[k];
		for (int i = a.pstart[p]; i < a.pstart[p + 1]; ++i) fn(i, a.li0[i] == r ? 0 : 1);
	}
}

// ru = M (v - vold) + sum_j T_j f (static rows 0)  (PrimalDualForceSolver.cpp:151-163).
// TRIAL: at v + 0.99 step du and the trial forces; also partial du_in . ru.
template <bool TRIAL>
__global__ __launch_bounds__(256) void k_body_res(SArgs a, const double* __restrict__ dudir, double step) {
	double acc[2] = {0, 0};
	const double* F = TRIAL ? a.ft : a.f;
	const int fd = a.nsc ? 3 : 1;
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.NL; r += gridDim.x * blockDim.x) {
		double ru[6] = {0, 0, 0, 0, 0, 0};
		if (!a.lstat[r]) {
			double d[6];
			for (int k = 0; k < 6; ++k) {
				double vk = a.v[6 * r + k];
				if (TRIAL) vk = vk + 0.99 * (step * dudir[6 * r + k]);
				d[k] = vk - a.vold[6 * r + k];
			}
			const double* M = &a.Ms[7 * r];
			for (int k = 0; k < 3; ++k) ru[k] = M[0] * d[k];
			double t[3];
			sym3_mv(M + 1, d + 3, t);
			ru[3] = t[0]; ru[4] = t[1]; ru[5] = t[2];
			double add[6] = {0, 0, 0, 0, 0, 0};
			for_incident(a, r, [&](int i, int j) {
				for (int row = 0; row < 6; ++row) {
					double s = 0;
					for (int col = 0; col < fd; ++col) s += Tc(a, i, j, col * 6 + row) * F[col * a.cs + i];
					add[row] += s;
				}
			});
			for (int k = 0; k < 6; ++k) ru[k] += add[k];
		}
		if (!TRIAL)
			for (int k = 0; k < 6; ++k) a.ru[6 * r + k] = ru[k];
		for (int k = 0; k < 6; ++k) {
			acc[0] += ru[k] * ru[k];
			if (TRIAL) acc[1] += (step * dudir[6 * r + k]) * ru[k];
		}
	}
	const int slots[2] = {TRIAL ? P_T_RU : P_RU, P_T_GU};
	write_partials<2>(acc, a.part, slots);
}

// err, eps and the exit test (PrimalDualForceSolver.cpp:295-320); one workgroup.
__global__ __launch_bounds__(256) void k_err_final(SArgs a, int nwg_c, int nwg_b, double Hu, double nR, double nF,
                            double nLc, double nC) {
	__shared__ double slot;
	const double wrf2 = block_reduce_array(&a.part[P_WRF * kMaxWG], nwg_c, &slot);
	const double rl2 = block_reduce_array(&a.part[P_RL * kMaxWG], nwg_c, &slot);
	const double ac2 = block_reduce_array(&a.part[P_AC * kMaxWG], nwg_c, &slot);
	const double ru2 = block_reduce_array(&a.part[P_RU * kMaxWG], nwg_b, &slot);
	if (threadIdx.x == 0) {
		Scalars* S = a.sc;
		S->ru2 = ru2; S->wrf2 = wrf2; S->rl2 = rl2; S->ac2 = ac2;
		const double err = ru2 + wrf2 + rl2;
		S->err = err;
		S->sErr = err;
		S->sacErr = ru2 + ac2 + rl2;
		S->Hu = Hu;
		S->eps = Hu * pow(err, 0.66667 * 0.5);
		const double momen = sqrt(ru2) / nR;
		double e = fmax(momen, sqrt(wrf2) / nF);
		e = fmax(e, sqrt(rl2) / nLc);
		e = fmax(e, S->gap);
		e = fmin(e, fmax(momen, sqrt(ac2) / nC));
		S->exitErr = e;
	}
}

// ---- (c) assembly ----------------------------------------------------------
// Per contact: invC, sf, Kn^-1, Kt^-1, Kt^-1 sf (NonSmoothContactForce.cpp:184-219).
__global__ __launch_bounds__(256) void k_asm_prep(SArgs a) {
	const double eps = a.sc->eps;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.C; i += gridDim.x * blockDim.x) {
		const double l0 = a.lam[i];
		if (a.nsc) {
			const double l1 = a.lam[a.cs + i];
			const double ic0 = 1. / (a.c[i] - eps * l0);
			const double ic1 = 1. / (a.c[a.cs + i] - eps * l1);
			const double g0 = a.g[i], g1 = a.g[a.cs + i];
			double lN = -l0 * ic0;
			if (a.spring) lN += a.compliance * a.sc->char_mass;
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
			if (a.spring) reg += a.compliance * a.sc->char_mass;
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

// acc(6x6) += T_j W T_k^T for contact i (upper triangle only when SYM).
__device__ __forceinline__ void add_tWt(const SArgs& a, int i, int j, int k, double* B) {
	const double kn = a.W[i];
	const double w00 = a.W[a.cs + i], w01 = a.W[2 * a.cs + i], w11 = a.W[3 * a.cs + i];
	double tk0[6], tk1[6], tk2[6];
	for (int c = 0; c < 6; ++c) {
		tk0[c] = Tc(a, i, k, c);
		tk1[c] = Tc(a, i, k, 6 + c);
		tk2[c] = Tc(a, i, k, 12 + c);
	}
	for (int r = 0; r < 6; ++r) {
		const double a0 = Tc(a, i, j, r) * kn;
		const double t1 = Tc(a, i, j, 6 + r), t2 = Tc(a, i, j, 12 + r);
		const double at0 = t1 * w00 + t2 * w01;
		const double at1 = t1 * w01 + t2 * w11;
		for (int c = 0; c < 6; ++c) B[r * 6 + c] += a0 * tk0[c] + at0 * tk1[c] + at1 * tk2[c];
	}
}

// In-place inverse of an SPD 6x6 through its Cholesky factor; false if not SPD.
__device__ __forceinline__ bool spd_inverse6(const double* A, double* Ai) {
	double L[36];
	for (int k = 0; k < 36; ++k) L[k] = 0;
	for (int j = 0; j < 6; ++j) {
		double d = A[j * 6 + j];
		for (int k = 0; k < j; ++k) d -= L[j * 6 + k] * L[j * 6 + k];
		if (!(d > 0)) return false;
		const double ljj = sqrt(d);
		L[j * 6 + j] = ljj;
		for (int i = j + 1; i < 6; ++i) {
			double s = A[i * 6 + j];
			for (int k = 0; k < j; ++k) s -= L[i * 6 + k] * L[j * 6 + k];
			L[i * 6 + j] = s / ljj;
		}
	}
	for (int c = 0; c < 6; ++c) {
		double e[6] = {0, 0, 0, 0, 0, 0};
		e[c] = 1;
		for (int i = 0; i < 6; ++i) {
			double s = e[i];
			for (int k = 0; k < i; ++k) s -= L[i * 6 + k] * e[k];
			e[i] = s / L[i * 6 + i];
		}
		for (int i = 5; i >= 0; --i) {
			double s = e[i];
			for (int k = i + 1; k < 6; ++k) s -= L[k * 6 + i] * e[k];
			e[i] = s / L[i * 6 + i];
		}
		for (int i = 0; i < 6; ++i) Ai[i * 6 + c] = e[i];
	}
	return true;
}

// Diagonal blocks, rhs, Jacobi scaling D, block-Jacobi preconditioner, scaled b.
template <class T>
__global__ __launch_bounds__(256) void k_asm_diag(SArgs a, const int* __restrict__ adiag, T* __restrict__ Aval,
                           T* __restrict__ Pval, T* __restrict__ bvec) {
	const double eps = a.sc->eps;
	const int fd = a.nsc ? 3 : 1;
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.NL; r += gridDim.x * blockDim.x) {
		double B[36];
		for (int k = 0; k < 36; ++k) B[k] = 0;
		double y[6] = {0, 0, 0, 0, 0, 0};
		const bool st = a.lstat[r] != 0;
		if (st) {
			for (int k = 0; k < 6; ++k) B[k * 6 + k] = 1.0 + eps;
		} else {
			const double* M = &a.Ms[7 * r];
			for (int k = 0; k < 3; ++k) B[k * 6 + k] = M[0];
			B[3 * 6 + 3] = M[1]; B[3 * 6 + 4] = M[2]; B[3 * 6 + 5] = M[3];
			B[4 * 6 + 3] = M[2]; B[4 * 6 + 4] = M[4]; B[4 * 6 + 5] = M[5];
			B[5 * 6 + 3] = M[3]; B[5 * 6 + 4] = M[5]; B[5 * 6 + 5] = M[6];
			for (int k = 0; k < 6; ++k) B[k * 6 + k] += eps;
			for_incident(a, r, [&](int i, int j) {
				// y += getTrans(j) kInvSf = -T_j kInvSf
				for (int row = 0; row < 6; ++row) {
					double s = 0;
					for (int col = 0; col < fd; ++col) s += Tc(a, i, j, col * 6 + row) * a.kf[col * a.cs + i];
					y[row] -= s;
				}
				add_tWt(a, i, j, j, B);
			});
		}
		double Dr[6], rhs[6];
		for (int k = 0; k < 6; ++k) {
			Dr[k] = 1 / sqrt(B[k * 6 + k]);
			rhs[k] = st ? 0.0 : y[k] - a.ru[6 * r + k];
			a.D[6 * r + k] = Dr[k];
			a.rhs[6 * r + k] = rhs[k];
			bvec[6 * r + k] = (T)(rhs[k] * Dr[k]);
		}
		double S[36], Si[36];
		for (int i = 0; i < 6; ++i)
			for (int j = 0; j < 6; ++j) S[i * 6 + j] = B[i * 6 + j] * (Dr[i] * Dr[j]);
		const int d = adiag[r];
		for (int k = 0; k < 36; ++k) Aval[36 * (size_t)d + k] = (T)S[k];
		if (!spd_inverse6(S, Si)) {
			for (int k = 0; k < 36; ++k) Si[k] = 0;
			for (int k = 0; k < 6; ++k) Si[k * 6 + k] = 1.0 / S[k * 6 + k];
		}
		for (int k = 0; k < 36; ++k) Pval[36 * (size_t)r + k] = (T)Si[k];
	}
}

// Off-diagonal blocks: sum over the pair's contacts of T_side(r) W T_side(c)^T, scaled.
template <class T>
__global__ __launch_bounds__(256) void k_asm_off(SArgs a, int NL, const int* __restrict__ aptr, const int* __restrict__ acol,
                          const int* __restrict__ apair, T* __restrict__ Aval) {
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < NL; r += gridDim.x * blockDim.x) {
		for (int blk = aptr[r]; blk < aptr[r + 1]; ++blk) {
			const int p = apair[blk];
			if (p < 0) continue;
			const int c = acol[blk];
			double B[36];
			for (int k = 0; k < 36; ++k) B[k] = 0;
			for (int i = a.pstart[p]; i < a.pstart[p + 1]; ++i) {
				const int j = a.li0[i] == r ? 0 : 1;
				add_tWt(a, i, j, 1 - j, B);
			}
			for (int ii = 0; ii < 6; ++ii)
				for (int jj = 0; jj < 6; ++jj)
					Aval[36 * (size_t)blk + ii * 6 + jj] =
						(T)(B[ii * 6 + jj] * (a.D[6 * r + ii] * a.D[6 * c + jj]));
		}
	}
}

// ---- (d) PCG ----------------------------------------------------------------
template <class T>
struct PcgArgs {
	int nb;
	const int* ptr;
	const int* col;
	const T* A;
	const T* P;
	const T* b;
	T* x;
	T* r0;
	T* r1;
	T* z;
	T* p;
	T* q;
	T* part;  // 3 * npart
	int npart;
	int maxit;
	T tol2;
	int* iters;
};

template <class T>
__device__ __forceinline__ T spmv_row(const PcgArgs<T>& a, const T* __restrict__ xv, int t) {
	const int br = t / 6, i = t - 6 * br;
	T s = 0;
	for (int k = a.ptr[br]; k < a.ptr[br + 1]; ++k) {
		const T* blk = a.A + 36 * (size_t)k + 6 * i;
		const T* xc = xv + 6 * a.col[k];
#pragma unroll
		for (int c = 0; c < 6; ++c) s += blk[c] * xc[c];
	}
	return s;
}

// Block-Jacobi PCG (zero initial guess) in one cooperative launch; iterates
// while it < maxit and |r|^2 > tol^2 |b|^2.  Uses q = A p = A z + beta q_old so
// the SpMV reads z, which is final at the grid barrier.
template <class T>
__global__ __launch_bounds__(256) void k_pcg(PcgArgs<T> a) {
	cg::grid_group grid = cg::this_grid();
	__shared__ T lds[3 * 16];
	__shared__ T slot;
	const int n = 6 * a.nb;
	const int gt = blockIdx.x * blockDim.x + threadIdx.x;
	const int gs = gridDim.x * blockDim.x;
	const int W = gridDim.x;
	T v2[2] = {0, 0};
	for (int t = gt; t < n; t += gs) {
		const int br = t / 6, i = t - 6 * br;
		const T bt = a.b[t];
		a.r0[t] = bt;
		a.x[t] = 0;
		a.p[t] = 0;
		a.q[t] = 0;
		T zt = 0;
		for (int c = 0; c < 6; ++c) zt += a.P[36 * (size_t)br + 6 * i + c] * a.b[6 * br + c];
		a.z[t] = zt;
		v2[0] += bt * zt;
		v2[1] += bt * bt;
	}
	{
		T out[2];
		block_sum<2, T>(v2, lds, out);
		if (threadIdx.x == 0) { a.part[1 * a.npart + blockIdx.x] = out[0]; a.part[2 * a.npart + blockIdx.x] = out[1]; }
	}
	grid.sync();
	T rz = block_reduce_array(&a.part[1 * a.npart], W, &slot);
	const T bb = block_reduce_array(&a.part[2 * a.npart], W, &slot);
	const T eps2 = a.tol2 * bb;
	T rr = bb;
	T beta = 0;
	T* rc = a.r0;
	T* rn = a.r1;
	int it = 0;
	if (rr > eps2) {
		while (it < a.maxit) {
			T v1[1] = {0};
			for (int t = gt; t < n; t += gs) {
				const T pt = a.z[t] + beta * a.p[t];
				const T qt = spmv_row(a, a.z, t) + beta * a.q[t];
				a.p[t] = pt;
				a.q[t] = qt;
				v1[0] += pt * qt;
			}
			{
				T out[1];
				block_sum<1, T>(v1, lds, out);
				if (threadIdx.x == 0) a.part[blockIdx.x] = out[0];
			}
			grid.sync();
			const T pq = block_reduce_array(&a.part[0], W, &slot);
			const T alpha = rz / pq;
			T v3[2] = {0, 0};
			for (int t = gt; t < n; t += gs) {
				const int br = t / 6, i = t - 6 * br;
				a.x[t] += alpha * a.p[t];
				T rv[6];
				for (int c = 0; c < 6; ++c) rv[c] = rc[6 * br + c] - alpha * a.q[6 * br + c];
				rn[t] = rv[i];
				T zt = 0;
				for (int c = 0; c < 6; ++c) zt += a.P[36 * (size_t)br + 6 * i + c] * rv[c];
				a.z[t] = zt;
				v3[0] += rv[i] * zt;
				v3[1] += rv[i] * rv[i];
			}
			{
				T out[2];
				block_sum<2, T>(v3, lds, out);
				if (threadIdx.x == 0) { a.part[1 * a.npart + blockIdx.x] = out[0]; a.part[2 * a.npart + blockIdx.x] = out[1]; }
			}
			grid.sync();
			const T rzn = block_reduce_array(&a.part[1 * a.npart], W, &slot);
			rr = block_reduce_array(&a.part[2 * a.npart], W, &slot);
			++it;
			if (!(rr > eps2)) break;
			beta = rzn / rz;
			rz = rzn;
			T* tmp = rc; rc = rn; rn = tmp;
		}
	}
	if (gt == 0) *a.iters = it;
}

template <class T>
__global__ __launch_bounds__(256) void k_du(int NL, const double* __restrict__ D, const T* __restrict__ x, double* __restrict__ du) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < 6 * NL) du[t] = (double)x[t] * D[t];
}

// ---- increments (NonSmoothContactForce.cpp:254-280) -------------------------
__global__ __launch_bounds__(256) void k_increments(SArgs a) {
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.C; i += gridDim.x * blockDim.x) {
		double u[3];
		to_local(a, i, &a.du[6 * a.li0[i]], &a.du[6 * a.li1[i]], u);
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

__global__ __launch_bounds__(256) void k_neg(int n, const double* __restrict__ x, double* __restrict__ y) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) y[t] = -x[t];
}

// ---- (e) line search trial (PrimalDualForceSolver.cpp:588-613) --------------
// Filter (NonSmoothContactForce.cpp:281-319), trial point, contact residuals
// with fresh constraints; partial |wrf|^2, |rl|^2, |ac|^2, df.rf, (dl/l).rl.
__global__ __launch_bounds__(256) void k_ls_contact(SArgs a, const double* __restrict__ du, const double* __restrict__ dF,
                             const double* __restrict__ dL, double step) {
	const double muc = a.sc->muc;
	double acc[5] = {0, 0, 0, 0, 0};
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.C; i += gridDim.x * blockDim.x) {
		const double cf[3] = {a.f[i], a.nsc ? a.f[a.cs + i] : 0.0, a.nsc ? a.f[2 * a.cs + i] : 0.0};
		const double lam[2] = {a.lam[i], a.nsc ? a.lam[a.cs + i] : 0.0};
		double dfi[3] = {step * dF[i], a.nsc ? step * dF[a.cs + i] : 0.0, a.nsc ? step * dF[2 * a.cs + i] : 0.0};
		double dli[2] = {step * dL[i], a.nsc ? step * dL[a.cs + i] : 0.0};
		if (a.nsc) {
			double tf[3] = {cf[0] + dfi[0], cf[1] + dfi[1], cf[2] + dfi[2]};
			if (tf[0] < 0) {
				tf[0] = 0;
				for (int k = 0; k < 3; ++k) dfi[k] = tf[k] - cf[k];
			}
			if (Lorentz::constraint(a.mu, tf[0], tf[1], tf[2]) > 0) {
				double pf[3];
				Lorentz::project(a.mu, tf, pf);
				for (int k = 0; k < 3; ++k) dfi[k] = pf[k] - cf[k];
			}
		} else {
			const double tf = cf[0] + dfi[0];
			if (tf < 0) dfi[0] = 0 - cf[0];
		}
		const int nl = a.nsc ? 2 : 1;
		for (int k = 0; k < nl; ++k)
			if (dli[k] < 0 && lam[k] + dli[k] < 0) dli[k] = -lam[k];
		double fI[3], lI[2];
		for (int k = 0; k < 3; ++k) fI[k] = cf[k] + 0.99 * dfi[k];
		for (int k = 0; k < 2; ++k) lI[k] = lam[k] + 0.99 * dli[k];
		const int l0 = a.li0[i], l1 = a.li1[i];
		double va[6], vb[6];
		for (int k = 0; k < 6; ++k) {
			va[k] = a.v[6 * l0 + k] + 0.99 * (step * du[6 * l0 + k]);
			vb[k] = a.v[6 * l1 + k] + 0.99 * (step * du[6 * l1 + k]);
		}
		double rf[3], rl[2], wrf[3], ac[3];
		contact_terms(a, i, va, vb, fI, lI, muc, rf, rl, wrf, ac);
		a.ft[i] = fI[0];
		a.lt[i] = lI[0];
		if (a.nsc) {
			a.ft[a.cs + i] = fI[1];
			a.ft[2 * a.cs + i] = fI[2];
			a.lt[a.cs + i] = lI[1];
		}
		acc[0] += wrf[0] * wrf[0] + wrf[1] * wrf[1] + wrf[2] * wrf[2];
		acc[1] += rl[0] * rl[0] + (a.nsc ? rl[1] * rl[1] : 0.0);
		acc[2] += ac[0] * ac[0] + ac[1] * ac[1] + ac[2] * ac[2];
		acc[3] += dfi[0] * rf[0] + (a.nsc ? dfi[1] * rf[1] + dfi[2] * rf[2] : 0.0);
		acc[4] += dli[0] / lI[0] * rl[0] + (a.nsc ? dli[1] / lI[1] * rl[1] : 0.0);
	}
	const int slots[5] = {P_T_WRF, P_T_RL, P_T_AC, P_T_GF, P_T_GL};
	write_partials<5>(acc, a.part, slots);
}

__global__ __launch_bounds__(256) void k_ls_final(SArgs a, int nwg_c, int nwg_b) {
	__shared__ double slot;
	const double wrf2 = block_reduce_array(&a.part[P_T_WRF * kMaxWG], nwg_c, &slot);
	const double rl2 = block_reduce_array(&a.part[P_T_RL * kMaxWG], nwg_c, &slot);
	const double ac2 = block_reduce_array(&a.part[P_T_AC * kMaxWG], nwg_c, &slot);
	const double gf = block_reduce_array(&a.part[P_T_GF * kMaxWG], nwg_c, &slot);
	const double gl = block_reduce_array(&a.part[P_T_GL * kMaxWG], nwg_c, &slot);
	const double ru2 = block_reduce_array(&a.part[P_T_RU * kMaxWG], nwg_b, &slot);
	const double gu = block_reduce_array(&a.part[P_T_GU * kMaxWG], nwg_b, &slot);
	if (threadIdx.x == 0) {
		Scalars* S = a.sc;
		S->t_inErr = ru2 + wrf2 + rl2;
		S->t_sacIn = ru2 + ac2 + rl2;
		S->t_gErr = -gu + gf + gl;
	}
}

__global__ __launch_bounds__(256) void k_commit_v(int n, double* __restrict__ v, const double* __restrict__ du, double step) {
	const int t = blockIdx.x * blockDim.x + threadIdx.x;
	if (t < n) v[t] = v[t] + 0.99 * (step * du[t]);
}

// ---- write-back (PrimalDualForceSolver.cpp:491-510) --------------------------
__global__ __launch_bounds__(256) void k_writeback_bodies(SArgs a, double* __restrict__ acc) {
	const double s = a.sc->char_speed;
	for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < a.NL; r += gridDim.x * blockDim.x) {
		const int g = a.l2g[r];
		for (int k = 0; k < 6; ++k) acc[6 * g + k] = a.v[6 * r + k] * s - a.vphys[6 * r + k];
	}
}
__global__ __launch_bounds__(256) void k_writeback_contacts(SArgs a, double* __restrict__ lf) {
	const double s = a.sc->char_speed, m = a.sc->char_mass;
	for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.C; i += gridDim.x * blockDim.x) {
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
__global__ __launch_bounds__(256) void k_integrate(int N, const int* __restrict__ stat, const double* __restrict__ acc,
                            double* __restrict__ pos, double* __restrict__ quat,
                            double* __restrict__ vel, double* __restrict__ ang, double dt) {
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= N) return;
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
	a.C = ctx->C; a.cs = ctx->c_stride; a.NL = ctx->NL;
	a.nsc = p.force_model == PDC_NON_SMOOTH_CONTACT;
	a.mu = p.friction_coeff; a.restitution = p.restitution;
	a.spring = p.spring_based_force;
	const double damp = p.spring_damp * p.spring_k;
	const double tk = dt * p.spring_k;
	a.compliance = 1. / (dt * (tk + damp));
	a.reduction = tk / (tk + damp);
	a.dt = dt;
	for (int k = 0; k < 3; ++k) a.gdt[k] = g[k] * dt;
	a.li0 = ctx->li0.p; a.li1 = ctx->li1.p; a.T = ctx->c_T.p; a.depth = ctx->c_depth.p;
	a.adj_ptr = ctx->adj_ptr.p; a.adj_pair = ctx->adj_pair.p; a.pstart = ctx->ap_start.p;
	a.lstat = ctx->l_static.p; a.l2g = ctx->l2g.p;
	a.v = ctx->l_v.p; a.vold = ctx->l_vold.p; a.vphys = ctx->l_vphys.p; a.Ms = ctx->l_Ms.p;
	a.Minv = ctx->l_Minv.p; a.speed = ctx->l_speed.p; a.ru = ctx->l_ru.p; a.rhs = ctx->l_rhs.p;
	a.D = ctx->l_D.p; a.du = ctx->l_du.p;
	a.K = ctx->s_K.p; a.f = ctx->s_f.p; a.lam = ctx->s_lam.p; a.c = ctx->s_c.p; a.g = ctx->s_g.p;
	a.H = ctx->s_H.p; a.rf = ctx->s_rf.p; a.rl = ctx->s_rl.p; a.sf = ctx->s_sf.p;
	a.invC = ctx->s_invC.p; a.W = ctx->s_W.p; a.kf = ctx->s_kf.p; a.df = ctx->s_df.p;
	a.dl = ctx->s_dl.p; a.ft = ctx->s_ft.p; a.lt = ctx->s_lt.p;
	a.part = ctx->partials.p; a.sc = ctx->d_sc.p;
	return a;
}

template <class T>
static int run_pcg(pdc_context* ctx, DBuf<T>& A, DBuf<T>& P, DBuf<T>& b, DBuf<T>& x, DBuf<T>& work) {
	const int NL = ctx->NL;
	const int n = 6 * NL;
	static int max_blocks_cache[2] = {0, 0};
	int& maxb = max_blocks_cache[sizeof(T) == 8];
	if (maxb == 0) {
		int per_cu = 0;
		PDC_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_pcg<T>, 256, 0));
		maxb = std::max(1, per_cu) * ctx->cus;
	}
	const int nblk = std::max(1, std::min({grid_for(n, 256), maxb, 1024}));
	work.ensure(5 * (size_t)n + 8);
	ctx->pcg_part.ensure(4096);
	PcgArgs<T> pa;
	pa.nb = NL;
	pa.ptr = ctx->a_ptr.p;
	pa.col = ctx->a_col.p;
	pa.A = A.p;
	pa.P = P.p;
	pa.b = b.p;
	pa.x = x.p;
	pa.r0 = work.p;
	pa.r1 = work.p + n;
	pa.z = work.p + 2 * n;
	pa.p = work.p + 3 * n;
	pa.q = work.p + 4 * n;
	pa.part = reinterpret_cast<T*>(ctx->pcg_part.p);
	pa.npart = 1024;
	pa.maxit = ctx->p.pcg_max_iter;
	pa.tol2 = (T)(ctx->p.pcg_tol * ctx->p.pcg_tol);
	pa.iters = &ctx->d_sc.p->pcg_iters;
	void* args[] = {&pa};
	hipEvent_t e0;
	kt_begin(ctx, PDC_KT_PCG, &e0);
	PDC_HIP(hipLaunchCooperativeKernel((const void*)k_pcg<T>, dim3(nblk), dim3(256), args, 0,
	                                   ctx->stream));
	kt_end(ctx, PDC_KT_PCG, e0);
	hipLaunchKernelGGL(k_du<T>, dim3(grid_for(n)), dim3(kBlock), 0, ctx->stream, NL, ctx->l_D.p,
	                   x.p, ctx->l_du.p);
	PDC_CHECK_LAUNCH();
	return nblk;
}

// Algorithmic bytes (each input read once, each output written once).
// PCG launch with `it` iterations, scalar size T, n6 = 6 NL:
//   set-up   b read, r x p q z written (6 T n6) + block-Jacobi blocks (36 T NL)
//   per it   A blocks and column indices nnzb (36 T + 4), row pointers 4 (NL+1),
//            block-Jacobi blocks 36 T NL, twelve n6-vector reads/writes (12 T n6)
double pcg_bytes(int NL, int nnzb, int it, int T) {
	const double n6 = 6.0 * NL;
	const double init = 6.0 * T * n6 + 36.0 * T * NL;
	const double per = (double)nnzb * (36.0 * T + 4.0) + 4.0 * (NL + 1) + 36.0 * T * NL + 12.0 * T * n6;
	return init + (double)it * per;
}
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
	                   ctx->b_static.p, ctx->b_acc.p, ctx->b_pos.p, ctx->b_quat.p, ctx->b_vel.p,
	                   ctx->b_ang.p, dt);
	PDC_CHECK_LAUNCH();
}

void solve(pdc_context* ctx, double dt, const double* g, pdc_solve_stats* st) {
	hipStream_t s = ctx->stream;
	pdc_solve_stats stats{};
	write_acc_default(ctx, dt, g);
	const int C = ctx->C;
	if (C == 0) {
		if (st) *st = stats;
		return;
	}
	build_topology(ctx);
	const int NL = ctx->NL;
	const size_t nb = (size_t)NL;
	ctx->l_v.ensure(6 * nb); ctx->l_vold.ensure(6 * nb); ctx->l_vphys.ensure(6 * nb);
	ctx->l_Ms.ensure(7 * nb); ctx->l_Minv.ensure(7 * nb); ctx->l_speed.ensure(nb);
	ctx->l_ru.ensure(6 * nb); ctx->l_rhs.ensure(6 * nb); ctx->l_D.ensure(6 * nb);
	ctx->l_du.ensure(6 * nb); ctx->l_gdu.ensure(6 * nb);
	ctx->partials.ensure((size_t)P_COUNT * kMaxWG);
	const bool fp32 = ctx->p.pcg_precision == PDC_PCG_FP32;
	const size_t nnzb = (size_t)ctx->NNZB;
	if (fp32) {
		ctx->A32.ensure(36 * nnzb); ctx->P32.ensure(36 * nb); ctx->b32.ensure(6 * nb); ctx->x32.ensure(6 * nb);
	} else {
		ctx->A64.ensure(36 * nnzb); ctx->P64.ensure(36 * nb); ctx->b64.ensure(6 * nb); ctx->x64.ensure(6 * nb);
	}
	SArgs a = make_args(ctx, dt, g);
	const int wc = wg_for(C), wb = wg_for(NL);
	const dim3 B(kBlock);
	const int nsc = a.nsc;
	const int fd = nsc ? 3 : 1, cd = nsc ? 2 : 1;
	const double nR = 6.0 * NL, nF = (double)fd * C, nLc = (double)cd * C;

	hipEvent_t t0;
	kt_begin(ctx, PDC_KT_NONE, &t0);
	hipLaunchKernelGGL(k_setup_bodies, dim3(wb), B, 0, s, a, ctx->b_vel.p, ctx->b_ang.p, ctx->b_quat.p,
	                   ctx->b_mass.p, ctx->b_I0.p);
	PDC_CHECK_LAUNCH();
	hipLaunchKernelGGL(k_nondim_contacts, dim3(wc), B, 0, s, a);
	PDC_CHECK_LAUNCH();
	const double gn = std::sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
	hipLaunchKernelGGL(k_nondim_final, dim3(1), dim3(1024), 0, s, a, wb, ctx->b_mass.p, gn * dt);
	PDC_CHECK_LAUNCH();
	hipLaunchKernelGGL(k_scale_bodies, dim3(wb), B, 0, s, a);
	PDC_CHECK_LAUNCH();
	hipLaunchKernelGGL(k_preprocess, dim3(wc), B, 0, s, a);
	PDC_CHECK_LAUNCH();

	double Hu = 1e-4;
	int it = 0, pcg_total = 0, pcg_hits = 0, gd = 0, trials = 0;
	double exitErr = 0;
	Scalars* H = ctx->h_sc;
	for (it = 0; it < ctx->p.iterations; ++it) {
		hipEvent_t e0;
		kt_begin(ctx, PDC_KT_CONTACT_TERMS, &e0);
		hipLaunchKernelGGL(k_constraints, dim3(wc), B, 0, s, a);
		PDC_CHECK_LAUNCH();
		hipLaunchKernelGGL(k_contact_res, dim3(wc), B, 0, s, a, wc, 1.0 / nLc);
		PDC_CHECK_LAUNCH();
		kt_end(ctx, PDC_KT_CONTACT_TERMS, e0);
		if (e0) ctx->kt.bytes += contact_terms_bytes(C, NL);
		hipLaunchKernelGGL(k_body_res<false>, dim3(wb), B, 0, s, a, (const double*)nullptr, 0.0);
		PDC_CHECK_LAUNCH();
		hipLaunchKernelGGL(k_err_final, dim3(1), dim3(256), 0, s, a, wc, wb, Hu, nR, nF, nLc, (double)C);
		PDC_CHECK_LAUNCH();
		sync_scalars(ctx);
		exitErr = H->exitErr;
		if (exitErr < ctx->p.tolerance) break;
		const double sErr = H->sErr, sacErr = H->sacErr;
		// assembly
		kt_begin(ctx, PDC_KT_ASSEMBLY, &e0);
		hipLaunchKernelGGL(k_asm_prep, dim3(wc), B, 0, s, a);
		PDC_CHECK_LAUNCH();
		if (fp32) {
			hipLaunchKernelGGL(k_asm_diag<float>, dim3(wb), B, 0, s, a, ctx->a_diag.p, ctx->A32.p,
			                   ctx->P32.p, ctx->b32.p);
			PDC_CHECK_LAUNCH();
			hipLaunchKernelGGL(k_asm_off<float>, dim3(wb), B, 0, s, a, NL, ctx->a_ptr.p, ctx->a_col.p,
			                   ctx->a_pair.p, ctx->A32.p);
		} else {
			hipLaunchKernelGGL(k_asm_diag<double>, dim3(wb), B, 0, s, a, ctx->a_diag.p, ctx->A64.p,
			                   ctx->P64.p, ctx->b64.p);
			PDC_CHECK_LAUNCH();
			hipLaunchKernelGGL(k_asm_off<double>, dim3(wb), B, 0, s, a, NL, ctx->a_ptr.p, ctx->a_col.p,
			                   ctx->a_pair.p, ctx->A64.p);
		}
		PDC_CHECK_LAUNCH();
		kt_end(ctx, PDC_KT_ASSEMBLY, e0);
		// solve
		if (fp32) run_pcg<float>(ctx, ctx->A32, ctx->P32, ctx->b32, ctx->x32, ctx->pcg32);
		else run_pcg<double>(ctx, ctx->A64, ctx->P64, ctx->b64, ctx->x64, ctx->pcg64);
		hipLaunchKernelGGL(k_increments, dim3(wc), B, 0, s, a);
		PDC_CHECK_LAUNCH();
		sync_scalars(ctx);
		const int iters = H->pcg_iters;
		pcg_total += iters;
		if (ctx->kt.which == PDC_KT_PCG) ctx->kt.bytes += pcg_bytes(NL, ctx->NNZB, iters, fp32 ? 4 : 8);
		if (iters == ctx->p.pcg_max_iter) pcg_hits++;
		// line search
		auto line_search = [&](const double* du, const double* dF, const double* dL, int maxLS) {
			double step = 1.0;
			for (int k = 0; k < maxLS; ++k) {
				hipEvent_t e1;
				kt_begin(ctx, PDC_KT_LINE_SEARCH, &e1);
				hipLaunchKernelGGL(k_ls_contact, dim3(wc), B, 0, s, a, du, dF, dL, step);
				PDC_CHECK_LAUNCH();
				kt_end(ctx, PDC_KT_LINE_SEARCH, e1);
				hipLaunchKernelGGL(k_body_res<true>, dim3(wb), B, 0, s, a, du, step);
				PDC_CHECK_LAUNCH();
				hipLaunchKernelGGL(k_ls_final, dim3(1), dim3(256), 0, s, a, wc, wb);
				PDC_CHECK_LAUNCH();
				sync_scalars(ctx);
				trials++;
				const bool acpt = (H->t_inErr < sErr) | (H->t_sacIn < sacErr);
				if ((acpt && H->t_gErr > 0) || k == maxLS - 1) {
					if (acpt) {
						hipLaunchKernelGGL(k_commit_v, dim3(grid_for(6 * NL)), B, 0, s, 6 * NL, ctx->l_v.p,
						                   du, step);
						PDC_CHECK_LAUNCH();
						std::swap(ctx->s_f, ctx->s_ft);
						std::swap(ctx->s_lam, ctx->s_lt);
						a = make_args(ctx, dt, g);
						return step;
					}
					break;
				}
				step *= 0.5;
			}
			return 0.0;
		};
		double step = line_search(ctx->l_du.p, ctx->s_df.p, ctx->s_dl.p, 10);
		bool useGD = false;
		if (step == 0) {
			useGD = true;
			gd++;
			hipLaunchKernelGGL(k_neg, dim3(grid_for(6 * NL)), B, 0, s, 6 * NL, ctx->l_ru.p, ctx->l_gdu.p);
			PDC_CHECK_LAUNCH();
			step = line_search(ctx->l_gdu.p, ctx->s_rf.p, ctx->s_rl.p, 20);
		}
		if (!useGD && step == 1) {
			if (iters == ctx->p.pcg_max_iter) Hu *= 1.3;
			else Hu /= 4;
		} else {
			Hu *= 4;
		}
		Hu = std::max(Hu, 1e-6);
	}
	hipLaunchKernelGGL(k_writeback_bodies, dim3(wb), B, 0, s, a, ctx->b_acc.p);
	PDC_CHECK_LAUNCH();
	hipLaunchKernelGGL(k_writeback_contacts, dim3(wc), B, 0, s, a, ctx->c_lf.p);
	PDC_CHECK_LAUNCH();
	sync_scalars(ctx);
	stats.n_contacts = C;
	stats.n_local_bodies = NL;
	stats.newton_iterations = it;
	stats.pcg_iterations = pcg_total;
	stats.pcg_maxiter_hits = pcg_hits;
	stats.gd_fallbacks = gd;
	stats.line_search_trials = trials;
	stats.exit_error = exitErr;
	stats.char_mass = H->char_mass;
	stats.char_speed = H->char_speed;
	stats.ties = ctx->n_ties;
	stats.n_blocks = ctx->NNZB;
	if (st) *st = stats;
	ctx->timings.newton_iterations += it;
	ctx->timings.contacts += C;
}

}  // namespace pdc
