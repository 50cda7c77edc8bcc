// pdc_capi.hip — the extern "C" boundary of libpdc (include/pdc.h).
//
// No exception crosses the boundary: every entry point catches, records the
// message for pdc_last_error and returns a status code.  Configuration errors
// are the ones the reference throws std::runtime_error for
// (PrimalDualForceSolver.cpp:655-661, DynamicSystem.cpp:394-400).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <mutex>
#include <numeric>

#include "pdc_internal.h"

using pdc::HipError;

namespace pdc {
thread_local Arena* t_arena = nullptr;
thread_local int64_t* t_dev_bytes = nullptr;
}

// ---- device scratch of the entry point that allocates nothing --------------
// pdc_fill_forces_noalloc takes every device buffer it needs beyond the body
// table from this array.  Each device that loads the library holds its own
// copy until the process ends.  In host code `scratch` is only a host-side
// name: its device address is taken once per device at context creation, with
// that device current, and only that address is used.
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kNoAllocMaxContacts = 32768;
constexpr int kNoAllocMaxBodies = 4096;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (taken once per device).
static void scratch_setup(int device) {
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (device < 0 || device >= kMaxDevices)
		throw HipError("device index beyond the scratch table", PDC_ERR_NO_DEVICE);
	if (g_scratch_ptr[device]) return;
	void* p = nullptr;
	PDC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
	if (!p) throw HipError("hipGetSymbolAddress returned no address for the scratch array", PDC_ERR_DEVICE);
	g_scratch_ptr[device] = p;
}

static void* scratch_address(int device) {
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

namespace {

struct ConfigError : std::runtime_error {
	using std::runtime_error::runtime_error;
};

bool is_gfx950(int dev) {
	hipDeviceProp_t prop;
	if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return false;
	return std::strncmp(prop.gcnArchName, "gfx950", 6) == 0;
}

void validate(const pdc_params* p) {
	if (!p) throw ConfigError("null parameters");
	if (p->force_model != PDC_NON_SMOOTH_CONTACT && p->force_model != PDC_ONLY_NORMAL)
		throw ConfigError("non_smooth_force.name must be non_smooth_contact or only_normal");
	if (!(p->timestep > 0) || !std::isfinite(p->timestep)) throw ConfigError("timestep must be > 0");
	if (p->iterations < 0) throw ConfigError("iterations must be >= 0");
	if (!(p->tolerance >= 0)) throw ConfigError("tolerance must be >= 0");
	if (p->pcg_precision != PDC_PCG_FP32 && p->pcg_precision != PDC_PCG_FP64)
		throw ConfigError("pcg_precision must be PDC_PCG_FP32 or PDC_PCG_FP64");
	if (!(p->pcg_tol > 0)) throw ConfigError("pcg_tol must be > 0");
	if (p->pcg_max_iter < 1) throw ConfigError("pcg_max_iter must be >= 1");
	if (p->grid_n < 4 || p->grid_n > 512) throw ConfigError("grid_n must be in [4, 512]");
	if (!(p->grid_pad_frac > 0) || p->grid_pad_frac > 10) throw ConfigError("grid_pad_frac must be in (0, 10]");
	if (!(p->sample_spacing > 0)) throw ConfigError("sample_spacing must be > 0");
	if (!(p->friction_coeff >= 0)) throw ConfigError("static_coeff must be >= 0");
	if (p->spring_based_force && !(p->spring_k > 0)) throw ConfigError("spring_k must be > 0");
}

// The buffers' byte counter of the context a call works on, for its duration.
struct ByteCounter {
	explicit ByteCounter(pdc_context* ctx) { pdc::t_dev_bytes = &ctx->dev_bytes; }
	~ByteCounter() { pdc::t_dev_bytes = nullptr; }
};

template <class F>
int guard(pdc_context* ctx, F fn) {
	if (!ctx) return PDC_ERR_INVALID_ARGUMENT;
	ByteCounter counter(ctx);
	try {
		ctx->err.clear();
		PDC_HIP(hipSetDevice(ctx->device));
		fn();
		return PDC_OK;
	} catch (const ConfigError& e) {
		ctx->err = e.what();
		return PDC_ERR_CONFIG;
	} catch (const HipError& e) {
		ctx->err = e.what();
		return e.code;
	} catch (const std::invalid_argument& e) {
		ctx->err = e.what();
		return PDC_ERR_INVALID_ARGUMENT;
	} catch (const std::logic_error& e) {
		ctx->err = e.what();
		return PDC_ERR_STATE;
	} catch (const std::bad_alloc&) {
		ctx->err = "host allocation failed";
		return PDC_ERR_CAPACITY;
	} catch (const std::exception& e) {
		ctx->err = e.what();
		return PDC_ERR_DEVICE;
	}
}

double now_s() {
	return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void free_all(pdc_context* c) {
	c->b_shape.release(); c->b_static.release(); c->b_emit.release(); c->b_dims.release();
	c->b_mass.release(); c->b_I0.release(); c->b_pos.release(); c->b_quat.release();
	c->b_vel.release(); c->b_ang.release(); c->b_acc.release(); c->b_R.release(); c->b_aabb.release();
	c->b_grid.release(); c->grid_pool.release(); c->b_soff.release(); c->samples.release();
	for (auto& m : c->meshes) {
		m->d_recs.release();
		m->grid.release();
	}
	c->cells.release(); c->b_ncell.release(); c->stage_idx.release(); c->pr_big.release(); c->pr_st.release();
	c->bp_cnt.release(); c->bp_off.release(); c->pr_a.release(); c->pr_b.release();
	c->pr_soff.release(); c->pr_nh.release(); c->pr_nc.release(); c->pr_coff.release();
	c->pr_tmp.release(); c->stage_raw.release(); c->stage_m.release();
	c->c_id0.release(); c->c_id1.release(); c->c_pos.release(); c->c_depth.release();
	c->c_T.release(); c->c_lf.release();
	c->t_key_flag.release(); c->t_pid.release(); c->ap_a.release(); c->ap_b.release();
	c->ap_start.release(); c->g_flag.release(); c->g2l.release(); c->l2g.release();
	c->li0.release(); c->li1.release(); c->adj_cnt.release(); c->adj_ptr.release();
	c->adj_cur.release(); c->adj_nb.release(); c->adj_pair.release(); c->a_cnt.release();
	c->a_ptr.release(); c->a_diag.release();
	c->adj_pos.release(); c->adj_blk.release(); c->adj_nb_u.release(); c->adj_pair_u.release();
	c->adj_row_u.release(); c->pair_adj.release(); c->inc_cnt.release(); c->inc_ptr.release();
	c->inc.release(); c->c_pid.release();
	c->l_static.release(); c->l_v.release(); c->l_vold.release(); c->l_Ms.release();
	c->l_Minv.release(); c->l_vphys.release(); c->l_speed.release(); c->l_ru.release();
	c->l_rhs.release(); c->l_D.release(); c->l_du.release(); c->l_gdu.release();
	c->s_K.release(); c->s_f.release(); c->s_lam.release(); c->s_c.release(); c->s_g.release();
	c->s_H.release(); c->s_rf.release(); c->s_rl.release(); c->s_sf.release(); c->s_invC.release();
	c->s_W.release(); c->s_kf.release(); c->s_df.release(); c->s_dl.release(); c->partials.release();
	c->A32.release(); c->P32.release(); c->b32.release();
	c->A64.release(); c->P64.release(); c->b64.release(); c->x64.release(); c->y64.release(); c->pcg64.release();
	c->d_sc.release();
	c->b_scene.release(); c->d_sbeg.release(); c->d_step.release(); c->d_act.release(); c->ls_gate.release();
	c->r_cbeg.release(); c->r_lbeg.release(); c->r_pbeg.release(); c->r_nnzb.release();
	c->t_c.release(); c->t_b.release(); c->t_w.release(); c->t_p.release(); c->t_pcg.release();
	c->tb_c.release(); c->tb_b.release(); c->tb_w.release(); c->tb_p.release();
	c->tb_r.release(); c->l_slot.release(); c->e_col.release(); c->r_amax.release();
	c->d_ss.release(); c->d_sctl.release(); c->d_ps.release();
	c->c_stride = 0;
}

void require_bodies(pdc_context* ctx) {
	if (ctx->N <= 0) throw std::logic_error("no bodies uploaded (call pdc_upload_bodies first)");
}

}  // namespace

extern "C" {

void pdc_default_params(pdc_params* p) {
	if (!p) return;
	std::memset(p, 0, sizeof(*p));
	p->timestep = 1e-2;
	p->gravity[0] = 0.0; p->gravity[1] = -9.81; p->gravity[2] = 0.0;
	p->force_model = PDC_NON_SMOOTH_CONTACT;
	p->spring_based_force = 0;
	p->friction_coeff = 0.5;
	p->restitution = 0.0;
	p->spring_k = 1e6;
	p->spring_damp = 0.01;
	p->tolerance = 1e-6;
	p->iterations = 100;
	p->pcg_precision = PDC_PCG_FP64;
	p->pcg_tol = 1e-8;
	p->pcg_max_iter = 1000;
	p->grid_n = 64;
	p->grid_pad_frac = 0.1;
	p->sample_spacing = 0.1;
}

int pdc_device_count(void) {
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess) return 0;
	int ok = 0;
	for (int d = 0; d < n; ++d) ok += is_gfx950(d) ? 1 : 0;
	return ok;
}

int pdc_create(int device, const pdc_params* params, pdc_context** out) {
	if (!out) return PDC_ERR_INVALID_ARGUMENT;
	*out = nullptr;
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n || !is_gfx950(device))
		return PDC_ERR_NO_DEVICE;
	pdc_context* ctx = new pdc_context();
	ctx->device = device;
	pdc_params p;
	if (params) p = *params;
	else pdc_default_params(&p);
	int rc = guard(ctx, [&] {
		validate(&p);
		ctx->p = p;
		PDC_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
		ctx->d_sc.ensure(1);
		PDC_HIP(hipHostMalloc((void**)&ctx->h_sc, sizeof(pdc::Scalars), hipHostMallocDefault));
		if (const char* g = std::getenv("PDC_NO_GRAPHS")) ctx->use_graphs = g[0] == '0';
		std::memset(ctx->h_sc, 0, sizeof(pdc::Scalars));
		PDC_HIP(hipMemsetAsync(ctx->d_sc.p, 0, sizeof(pdc::Scalars), ctx->stream));
		hipDeviceProp_t prop;
		PDC_HIP(hipGetDeviceProperties(&prop, device));
		ctx->cus = prop.multiProcessorCount;
		scratch_setup(device);
		PDC_HIP(hipStreamSynchronize(ctx->stream));
	});
	if (rc != PDC_OK) {
		pdc_destroy(ctx);
		return rc;
	}
	*out = ctx;
	return PDC_OK;
}

void pdc_destroy(pdc_context* ctx) {
	if (!ctx) return;
	ByteCounter counter(ctx);
	(void)hipSetDevice(ctx->device);
	if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
	for (auto& G : ctx->pcg_graph) {
		if (G.exec) (void)hipGraphExecDestroy(G.exec);
		if (G.graph) (void)hipGraphDestroy(G.graph);
	}
	for (auto& G : ctx->phase_graph) {
		if (G.exec) (void)hipGraphExecDestroy(G.exec);
		if (G.graph) (void)hipGraphDestroy(G.graph);
	}
	for (auto e : ctx->kt_direct) (void)hipEventDestroy(e);
	free_all(ctx);
	for (auto e : ctx->kt.pool) (void)hipEventDestroy(e);
	if (ctx->h_sc) (void)hipHostFree(ctx->h_sc);
	if (ctx->h_ss) (void)hipHostFree(ctx->h_ss);
	if (ctx->h_prog) (void)hipHostFree(ctx->h_prog);
	if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
	delete ctx;
}

int pdc_configure(pdc_context* ctx, const pdc_params* params) {
	return guard(ctx, [&] {
		validate(params);
		const bool regrid = ctx->N > 0 && (params->grid_n != ctx->p.grid_n ||
		                                   params->grid_pad_frac != ctx->p.grid_pad_frac ||
		                                   params->sample_spacing != ctx->p.sample_spacing);
		if (regrid)
			throw ConfigError("grid_n, grid_pad_frac and sample_spacing are fixed once bodies are uploaded");
		if (params->friction_coeff != ctx->p.friction_coeff) ctx->h_mu.assign(ctx->S, params->friction_coeff);
		ctx->p = *params;
	});
}

const char* pdc_last_error(const pdc_context* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

static void check_bodies(const pdc_context* ctx, const char* who, int n, const int32_t* shape_type,
                         const double* dims, const double* mass, const double* inertia_diag,
                         const int32_t* is_static) {
	const std::string w(who);
	if (n < 0 || (n > 0 && (!shape_type || !dims || !mass || !inertia_diag || !is_static)))
		throw std::invalid_argument(w + ": null array or negative count");
	for (int i = 0; i < n; ++i) {
		if (shape_type[i] != PDC_SHAPE_BOX && shape_type[i] != PDC_SHAPE_SPHERE &&
		    shape_type[i] != PDC_SHAPE_ELLIPSOID && shape_type[i] != PDC_SHAPE_MESH)
			throw std::invalid_argument(w + ": unknown shape type");
		if (shape_type[i] == PDC_SHAPE_MESH) {
			const double id = dims[3 * i];
			if (!(id >= 0 && id < (double)ctx->meshes.size() && id == std::floor(id)))
				throw std::invalid_argument(w + ": a mesh body's dims[0] must be a registered mesh id");
		} else {
			for (int a = 0; a < 3; ++a)
				if (!(dims[3 * i + a] > 0) && !(shape_type[i] == PDC_SHAPE_SPHERE && a > 0))
					throw std::invalid_argument(w + ": dimensions must be > 0");
		}
		if (!(mass[i] > 0)) throw std::invalid_argument(w + ": mass must be > 0");
		for (int a = 0; a < 3; ++a)
			if (!(inertia_diag[3 * i + a] > 0)) throw std::invalid_argument(w + ": inertia must be > 0");
	}
}

// State of bodies [first, first + n); a null array leaves that part (a body
// that is new gets the identity pose and zero velocities from the caller).
static void write_state(pdc_context* ctx, int first, int n, const double* pos, const double* quat,
                        const double* vel, const double* ang) {
	hipStream_t s = ctx->stream;
	const size_t f = (size_t)first, m = (size_t)n;
	if (pos) PDC_HIP(hipMemcpyAsync(ctx->b_pos.p + 3 * f, pos, 3 * m * sizeof(double), hipMemcpyHostToDevice, s));
	if (quat) PDC_HIP(hipMemcpyAsync(ctx->b_quat.p + 4 * f, quat, 4 * m * sizeof(double), hipMemcpyHostToDevice, s));
	if (vel) PDC_HIP(hipMemcpyAsync(ctx->b_vel.p + 3 * f, vel, 3 * m * sizeof(double), hipMemcpyHostToDevice, s));
	if (ang) PDC_HIP(hipMemcpyAsync(ctx->b_ang.p + 3 * f, ang, 3 * m * sizeof(double), hipMemcpyHostToDevice, s));
	PDC_HIP(hipStreamSynchronize(s));
}

static void identity_state(pdc_context* ctx, int first, int n) {
	std::vector<double> z3(3 * (size_t)n, 0.0), q(4 * (size_t)n, 0.0);
	for (int i = 0; i < n; ++i) q[4 * i + 3] = 1.0;
	write_state(ctx, first, n, z3.data(), q.data(), z3.data(), z3.data());
}

int pdc_register_mesh(pdc_context* ctx, int32_t nv, const double* v, int32_t nt, const int32_t* tri,
                      int32_t* mesh_id) {
	return guard(ctx, [&] {
		if (nv < 3 || nt < 1 || !v || !tri || !mesh_id)
			throw std::invalid_argument("pdc_register_mesh: null array, or fewer than 3 vertices or 1 triangle");
		for (int64_t i = 0; i < 3 * (int64_t)nv; ++i)
			if (!std::isfinite(v[i])) throw std::invalid_argument("pdc_register_mesh: non-finite vertex coordinate");
		for (int64_t i = 0; i < 3 * (int64_t)nt; ++i)
			if (tri[i] < 0 || tri[i] >= nv) throw std::invalid_argument("pdc_register_mesh: vertex index out of range");
		*mesh_id = pdc::register_mesh(ctx, nv, v, nt, tri);
	});
}

int pdc_get_grid_nodes(pdc_context* ctx, int32_t body, int64_t capacity, float* nodes) {
	return guard(ctx, [&] {
		if (body < 0 || body >= ctx->N || !nodes) throw std::invalid_argument("pdc_get_grid_nodes: bad body or null");
		const int64_t n3 = (int64_t)ctx->p.grid_n * ctx->p.grid_n * ctx->p.grid_n;
		if (capacity < n3) throw pdc::HipError("pdc_get_grid_nodes: capacity below grid_n^3", PDC_ERR_CAPACITY);
		PDC_HIP(hipMemcpyAsync(nodes, ctx->grid_pool.p + (size_t)body * (size_t)n3, (size_t)n3 * sizeof(float),
		                       hipMemcpyDeviceToHost, ctx->stream));
		PDC_HIP(hipStreamSynchronize(ctx->stream));
	});
}

int pdc_upload_bodies(pdc_context* ctx, int32_t n, const int32_t* shape_type, const double* dims,
                      const double* mass, const double* inertia_diag, const int32_t* is_static,
                      const int32_t* emit_samples) {
	return guard(ctx, [&] {
		check_bodies(ctx, "pdc_upload_bodies", n, shape_type, dims, mass, inertia_diag, is_static);
		pdc::upload_bodies(ctx, n, shape_type, dims, mass, inertia_diag, is_static, emit_samples);
		if (n > 0) identity_state(ctx, 0, n);  // until the caller sets one
		ctx->C = 0;
	});
}

int pdc_append_bodies(pdc_context* ctx, int32_t n, const int32_t* shape_type, const double* dims,
                      const double* mass, const double* inertia_diag, const int32_t* is_static,
                      const int32_t* emit_samples, const double* pos, const double* quat_xyzw,
                      const double* vel, const double* angvel, int32_t* first_body) {
	return guard(ctx, [&] {
		if (n < 1) throw std::invalid_argument("pdc_append_bodies: no bodies");
		check_bodies(ctx, "pdc_append_bodies", n, shape_type, dims, mass, inertia_diag, is_static);
		const int first = pdc::append_bodies(ctx, n, shape_type, dims, mass, inertia_diag, is_static, emit_samples);
		identity_state(ctx, first, n);
		write_state(ctx, first, n, pos, quat_xyzw, vel, angvel);
		ctx->C = 0;
		if (first_body) *first_body = first;
	});
}

int pdc_replace_bodies(pdc_context* ctx, int32_t first, int32_t n, const int32_t* shape_type, const double* dims,
                       const double* mass, const double* inertia_diag, const int32_t* is_static,
                       const int32_t* emit_samples, const double* pos, const double* quat_xyzw,
                       const double* vel, const double* angvel) {
	return guard(ctx, [&] {
		require_bodies(ctx);
		check_bodies(ctx, "pdc_replace_bodies", n, shape_type, dims, mass, inertia_diag, is_static);
		pdc::replace_bodies(ctx, first, n, shape_type, dims, mass, inertia_diag, is_static, emit_samples);
		identity_state(ctx, first, n);
		write_state(ctx, first, n, pos, quat_xyzw, vel, angvel);
		ctx->C = 0;
	});
}

int pdc_set_state_range(pdc_context* ctx, int32_t first, int32_t n, const double* pos, const double* quat_xyzw,
                        const double* vel, const double* angvel) {
	return guard(ctx, [&] {
		require_bodies(ctx);
		if (first < 0 || n < 0 || first + n > ctx->N) throw std::invalid_argument("pdc_set_state_range: bad range");
		if (n > 0) write_state(ctx, first, n, pos, quat_xyzw, vel, angvel);
	});
}

int pdc_get_state_range(pdc_context* ctx, int32_t first, int32_t n, double* pos, double* quat_xyzw, double* vel,
                        double* angvel) {
	return guard(ctx, [&] {
		require_bodies(ctx);
		if (first < 0 || n < 0 || first + n > ctx->N) throw std::invalid_argument("pdc_get_state_range: bad range");
		if (n == 0) return;
		hipStream_t s = ctx->stream;
		const size_t f = (size_t)first, m = (size_t)n;
		if (pos) PDC_HIP(hipMemcpyAsync(pos, ctx->b_pos.p + 3 * f, 3 * m * sizeof(double), hipMemcpyDeviceToHost, s));
		if (quat_xyzw)
			PDC_HIP(hipMemcpyAsync(quat_xyzw, ctx->b_quat.p + 4 * f, 4 * m * sizeof(double), hipMemcpyDeviceToHost, s));
		if (vel) PDC_HIP(hipMemcpyAsync(vel, ctx->b_vel.p + 3 * f, 3 * m * sizeof(double), hipMemcpyDeviceToHost, s));
		if (angvel) PDC_HIP(hipMemcpyAsync(angvel, ctx->b_ang.p + 3 * f, 3 * m * sizeof(double), hipMemcpyDeviceToHost, s));
		PDC_HIP(hipStreamSynchronize(s));
	});
}

int pdc_reserve_bodies(pdc_context* ctx, int64_t bodies, int64_t samples) {
	return guard(ctx, [&] {
		if (bodies < 0 || samples < 0) throw std::invalid_argument("pdc_reserve_bodies: negative capacity");
		pdc::reserve_bodies(ctx, bodies, samples);
	});
}

int pdc_device_bytes(const pdc_context* ctx, int64_t* bytes) {
	if (!ctx || !bytes) return PDC_ERR_INVALID_ARGUMENT;
	*bytes = ctx->dev_bytes;
	return PDC_OK;
}

int pdc_get_groups(pdc_context* ctx, int32_t n_bodies, int32_t* group_key) {
	return guard(ctx, [&] {
		require_bodies(ctx);
		if (!group_key || n_bodies != ctx->N)
			throw std::invalid_argument("pdc_get_groups: group_key must hold one entry per uploaded body");
		if (!ctx->groups_valid)
			throw std::logic_error("pdc_get_groups: no solve below pcg_tol 1e-9 has built body groups");
		hipStream_t s = ctx->stream;
		const int NL = ctx->NL;
		std::vector<int> rep(std::max(NL, 1)), l2g(std::max(NL, 1));
		if (NL > 0) {
			PDC_HIP(hipMemcpyAsync(rep.data(), ctx->g_rep.p, NL * sizeof(int), hipMemcpyDeviceToHost, s));
			PDC_HIP(hipMemcpyAsync(l2g.data(), ctx->l2g.p, NL * sizeof(int), hipMemcpyDeviceToHost, s));
		}
		PDC_HIP(hipStreamSynchronize(s));
		std::fill(group_key, group_key + n_bodies, -1);
		for (int r = 0; r < NL; ++r) group_key[l2g[r]] = l2g[rep[r]];
	});
}

int pdc_set_state(pdc_context* ctx, const double* pos, const double* quat_xyzw, const double* vel,
                  const double* angvel) {
	return guard(ctx, [&] {
		require_bodies(ctx);
		const size_t n = (size_t)ctx->N;
		hipStream_t s = ctx->stream;
		if (pos) PDC_HIP(hipMemcpyAsync(ctx->b_pos.p, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
		if (quat_xyzw) PDC_HIP(hipMemcpyAsync(ctx->b_quat.p, quat_xyzw, 4 * n * sizeof(double), hipMemcpyHostToDevice, s));
		if (vel) PDC_HIP(hipMemcpyAsync(ctx->b_vel.p, vel, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
		if (angvel) PDC_HIP(hipMemcpyAsync(ctx->b_ang.p, angvel, 3 * n * sizeof(double), hipMemcpyHostToDevice, s));
		PDC_HIP(hipStreamSynchronize(s));
	});
}

int pdc_get_state(pdc_context* ctx, double* pos, double* quat_xyzw, double* vel, double* angvel) {
	return guard(ctx, [&] {
		require_bodies(ctx);
		const size_t n = (size_t)ctx->N;
		hipStream_t s = ctx->stream;
		if (pos) PDC_HIP(hipMemcpyAsync(pos, ctx->b_pos.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
		if (quat_xyzw) PDC_HIP(hipMemcpyAsync(quat_xyzw, ctx->b_quat.p, 4 * n * sizeof(double), hipMemcpyDeviceToHost, s));
		if (vel) PDC_HIP(hipMemcpyAsync(vel, ctx->b_vel.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
		if (angvel) PDC_HIP(hipMemcpyAsync(angvel, ctx->b_ang.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, s));
		PDC_HIP(hipStreamSynchronize(s));
	});
}

int pdc_get_detection_stats(const pdc_context* ctx, pdc_detection_stats* out) {
	if (!ctx || !out) return PDC_ERR_INVALID_ARGUMENT;
	const pdc::DetectionStats& d = ctx->det;
	out->pairs = d.pairs;
	out->raw_hits = d.raw_hits;
	out->ties = d.ties;
	out->samples_all = d.samples_all;
	out->samples_in_grid = d.samples_in_grid;
	out->samples_read = d.samples_read;
	out->brick_nodes = d.brick_nodes;
	out->large_pairs = d.large_pairs;
	return PDC_OK;
}

int pdc_detect_contacts(pdc_context* ctx, int32_t* n_contacts, int32_t* n_ties) {
	return guard(ctx, [&] {
		require_bodies(ctx);
		if (std::count(ctx->h_step.begin(), ctx->h_step.end(), 0) > 0) pdc::set_stepped(ctx, 0, nullptr);
		const double t0 = now_s();
		pdc::detect_contacts(ctx);
		PDC_HIP(hipStreamSynchronize(ctx->stream));
		ctx->timings.contact_detection += now_s() - t0;
		if (n_contacts) *n_contacts = ctx->C;
		if (n_ties) *n_ties = ctx->n_ties;
	});
}

int pdc_get_contacts(pdc_context* ctx, int32_t capacity, int32_t* obj_ids, double* pos, double* depth,
                     double* transformation, double* local_force) {
	return guard(ctx, [&] {
		const int C = ctx->C;
		if (capacity < C) throw std::invalid_argument("pdc_get_contacts: capacity smaller than the contact count");
		if (C == 0) return;
		const size_t cs = (size_t)ctx->c_stride;
		hipStream_t s = ctx->stream;
		std::vector<int> i0(C), i1(C);
		std::vector<double> P(3 * (size_t)C), Dp(C), T(36 * (size_t)C), L(3 * (size_t)C);
		PDC_HIP(hipMemcpyAsync(i0.data(), ctx->c_id0.p, C * sizeof(int), hipMemcpyDeviceToHost, s));
		PDC_HIP(hipMemcpyAsync(i1.data(), ctx->c_id1.p, C * sizeof(int), hipMemcpyDeviceToHost, s));
		for (int k = 0; k < 3; ++k) {
			PDC_HIP(hipMemcpyAsync(&P[k * (size_t)C], ctx->c_pos.p + k * cs, C * sizeof(double), hipMemcpyDeviceToHost, s));
			PDC_HIP(hipMemcpyAsync(&L[k * (size_t)C], ctx->c_lf.p + k * cs, C * sizeof(double), hipMemcpyDeviceToHost, s));
		}
		PDC_HIP(hipMemcpyAsync(Dp.data(), ctx->c_depth.p, C * sizeof(double), hipMemcpyDeviceToHost, s));
		PDC_HIP(hipMemcpyAsync(T.data(), ctx->c_T.p, 36 * (size_t)C * sizeof(double), hipMemcpyDeviceToHost, s));
		PDC_HIP(hipStreamSynchronize(s));
		for (int i = 0; i < C; ++i) {
			if (obj_ids) { obj_ids[2 * i] = i0[i]; obj_ids[2 * i + 1] = i1[i]; }
			for (int k = 0; k < 3; ++k) {
				if (pos) pos[3 * i + k] = P[k * (size_t)C + i];
				if (local_force) local_force[3 * i + k] = L[k * (size_t)C + i];
			}
			if (depth) depth[i] = Dp[i];
			if (transformation)
				for (int k = 0; k < 36; ++k) transformation[36 * (size_t)i + k] = T[36 * (size_t)i + k];
		}
	});
}

static void fill_forces_impl(pdc_context* ctx, int32_t n_contacts, const int32_t* obj_ids,
                             const double* transformation, const double* depth, const double* quat_xyzw,
                             const double* vel, const double* angvel, double* out_acc, double* out_local_force,
                             pdc_solve_stats* stats) {
	{
		require_bodies(ctx);
		const int C = n_contacts;
		const int N = ctx->N;
		if (C < 0 || (C > 0 && (!obj_ids || !transformation)) || !quat_xyzw || !vel || !angvel)
			throw std::invalid_argument("pdc_fill_forces: null array or negative count");
		for (int i = 0; i < 2 * C; ++i)
			if (obj_ids[i] < 0 || obj_ids[i] >= N)
				throw std::invalid_argument("pdc_fill_forces: contact body id out of range");
		for (int i = 0; i < C; ++i)
			if (obj_ids[2 * i] == obj_ids[2 * i + 1])
				throw std::invalid_argument("pdc_fill_forces: contact between a body and itself");
		if (ctx->S > 1) {
			auto scene_of = [&](int b) {
				return (int)(std::upper_bound(ctx->h_sbeg.begin(), ctx->h_sbeg.end(), b) - ctx->h_sbeg.begin()) - 1;
			};
			for (int i = 0; i < C; ++i)
				if (scene_of(obj_ids[2 * i]) != scene_of(obj_ids[2 * i + 1]))
					throw std::invalid_argument("pdc_fill_forces: contact between bodies of different scenes");
		}
		const double t0 = now_s();
		hipStream_t s = ctx->stream;
		PDC_HIP(hipMemcpyAsync(ctx->b_quat.p, quat_xyzw, 4 * (size_t)N * sizeof(double), hipMemcpyHostToDevice, s));
		PDC_HIP(hipMemcpyAsync(ctx->b_vel.p, vel, 3 * (size_t)N * sizeof(double), hipMemcpyHostToDevice, s));
		PDC_HIP(hipMemcpyAsync(ctx->b_ang.p, angvel, 3 * (size_t)N * sizeof(double), hipMemcpyHostToDevice, s));
		// contacts grouped by unordered body pair (stable), as the topology needs
		std::vector<int> perm(C);
		std::iota(perm.begin(), perm.end(), 0);
		auto key = [&](int i) {
			const int a = std::min(obj_ids[2 * i], obj_ids[2 * i + 1]);
			const int b = std::max(obj_ids[2 * i], obj_ids[2 * i + 1]);
			return std::make_pair(a, b);
		};
		std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) { return key(x) < key(y); });
		ctx->C = 0;
		if (C > 0) {
			pdc::ensure_contact_capacity(ctx, C);
			std::vector<int> i0(C), i1(C);
			std::vector<double> Dp(C), T(36 * (size_t)C);
			for (int k = 0; k < C; ++k) {
				const int i = perm[k];
				i0[k] = obj_ids[2 * i];
				i1[k] = obj_ids[2 * i + 1];
				Dp[k] = depth ? depth[i] : 0.0;
				for (int c = 0; c < 36; ++c) T[36 * (size_t)k + c] = transformation[36 * (size_t)i + c];
			}
			PDC_HIP(hipMemcpyAsync(ctx->c_id0.p, i0.data(), C * sizeof(int), hipMemcpyHostToDevice, s));
			PDC_HIP(hipMemcpyAsync(ctx->c_id1.p, i1.data(), C * sizeof(int), hipMemcpyHostToDevice, s));
			PDC_HIP(hipMemcpyAsync(ctx->c_depth.p, Dp.data(), C * sizeof(double), hipMemcpyHostToDevice, s));
			PDC_HIP(hipMemcpyAsync(ctx->c_T.p, T.data(), 36 * (size_t)C * sizeof(double), hipMemcpyHostToDevice, s));
			PDC_HIP(hipStreamSynchronize(s));
			ctx->C = C;
		}
		pdc::solve(ctx, ctx->p.timestep, ctx->p.gravity);
		const pdc_solve_stats st = pdc::aggregate_stats(ctx);
		if (out_acc)
			PDC_HIP(hipMemcpyAsync(out_acc, ctx->b_acc.p, 6 * (size_t)N * sizeof(double), hipMemcpyDeviceToHost, s));
		std::vector<double> L(3 * (size_t)C);
		if (C > 0 && out_local_force) {
			const size_t cs = (size_t)ctx->c_stride;
			for (int k = 0; k < 3; ++k)
				PDC_HIP(hipMemcpyAsync(&L[k * (size_t)C], ctx->c_lf.p + k * cs, C * sizeof(double),
				                       hipMemcpyDeviceToHost, s));
		}
		PDC_HIP(hipStreamSynchronize(s));
		if (C > 0 && out_local_force)
			for (int k = 0; k < C; ++k)
				for (int c = 0; c < 3; ++c) out_local_force[3 * perm[k] + c] = L[c * (size_t)C + k];
		ctx->last = st;
		if (stats) *stats = st;
		ctx->timings.force_process += now_s() - t0;
		ctx->timings.steps += 1;
	}
}

int pdc_fill_forces(pdc_context* ctx, int32_t n_contacts, const int32_t* obj_ids,
                    const double* transformation, const double* depth, const double* quat_xyzw,
                    const double* vel, const double* angvel, double* out_acc, double* out_local_force,
                    pdc_solve_stats* stats) {
	return guard(ctx, [&] {
		fill_forces_impl(ctx, n_contacts, obj_ids, transformation, depth, quat_xyzw, vel, angvel, out_acc,
		                 out_local_force, stats);
	});
}

int pdc_fill_forces_noalloc(pdc_context* ctx, int32_t n_contacts, const int32_t* obj_ids,
                            const double* transformation, const double* depth, const double* quat_xyzw,
                            const double* vel, const double* angvel, double* out_acc, double* out_local_force,
                            pdc_solve_stats* stats) {
	return guard(ctx, [&] {
		require_bodies(ctx);
		if (ctx->S != 1)
			throw std::logic_error("pdc_fill_forces_noalloc: the context must hold one scene");
		if (n_contacts > kNoAllocMaxContacts)
			throw HipError("pdc_fill_forces_noalloc: " + std::to_string(n_contacts) + " contacts exceed the " +
			                   std::to_string(kNoAllocMaxContacts) + " of its scratch",
			               PDC_ERR_CAPACITY);
		if (ctx->N > kNoAllocMaxBodies)
			throw HipError("pdc_fill_forces_noalloc: " + std::to_string(ctx->N) + " bodies exceed the " +
			                   std::to_string(kNoAllocMaxBodies) + " of its scratch",
			               PDC_ERR_CAPACITY);
		void* base = scratch_address(ctx->device);
		if (!base) throw std::logic_error("pdc_fill_forces_noalloc: scratch address not taken for this device");
		pdc::Arena arena;
		arena.base = static_cast<char*>(base);
		arena.cap = kScratchBytes;
		// Everything this call grows comes from the arena: the contact
		// buffers start from zero capacity, kernel timing (which creates
		// events) is off, and afterwards every buffer, the contact stride and
		// the timing class are put back.  The contacts of this call are not
		// kept (they live in the scratch the next call overwrites).
		const int saved_stride = ctx->c_stride, saved_kt = ctx->kt.which;
		auto undo = [&]() {
			pdc::t_arena = nullptr;
			arena.restore();
			ctx->c_stride = saved_stride;
			ctx->kt.which = saved_kt;
			ctx->C = 0;
			ctx->groups_valid = false;
		};
		ctx->c_stride = 0;
		ctx->kt.which = PDC_KT_NONE;
		pdc::t_arena = &arena;
		try {
			fill_forces_impl(ctx, n_contacts, obj_ids, transformation, depth, quat_xyzw, vel, angvel, out_acc,
			                 out_local_force, stats);
		} catch (...) {
			(void)hipStreamSynchronize(ctx->stream);
			undo();
			throw;
		}
		undo();
	});
}

static void step_stepped_scenes(pdc_context* ctx, int n_steps) {
	for (int k = 0; k < n_steps; ++k) {
		const double t0 = now_s();
		pdc::detect_contacts(ctx);
		const double t1 = now_s();
		pdc::solve(ctx, ctx->p.timestep, ctx->p.gravity);
		pdc::integrate(ctx, ctx->p.timestep);
		PDC_HIP(hipStreamSynchronize(ctx->stream));
		const double t2 = now_s();
		ctx->timings.contact_detection += t1 - t0;
		ctx->timings.force_process += t2 - t1;
		ctx->timings.steps += 1;
		ctx->last = pdc::aggregate_stats(ctx);
	}
}

int pdc_step(pdc_context* ctx, int32_t n_steps, pdc_solve_stats* stats) {
	return guard(ctx, [&] {
		require_bodies(ctx);
		if (n_steps < 0) throw std::invalid_argument("pdc_step: negative step count");
		if (std::count(ctx->h_step.begin(), ctx->h_step.end(), 0) > 0) pdc::set_stepped(ctx, 0, nullptr);
		step_stepped_scenes(ctx, n_steps);
		if (stats) *stats = ctx->last;
	});
}

int pdc_set_scenes(pdc_context* ctx, int32_t n_scenes, const int32_t* body_begin) {
	return guard(ctx, [&] {
		require_bodies(ctx);
		if (n_scenes < 1 || !body_begin) throw std::invalid_argument("pdc_set_scenes: null offsets or no scene");
		pdc::set_scenes(ctx, n_scenes, body_begin);
	});
}

int pdc_scene_count(const pdc_context* ctx) { return ctx ? ctx->S : 0; }

int pdc_set_scene_friction(pdc_context* ctx, const double* mu) {
	return guard(ctx, [&] {
		if (!mu) throw std::invalid_argument("pdc_set_scene_friction: null array");
		for (int k = 0; k < ctx->S; ++k)
			if (!(mu[k] >= 0)) throw ConfigError("static_coeff must be >= 0");
		ctx->h_mu.assign(mu, mu + ctx->S);
	});
}

int pdc_step_scenes(pdc_context* ctx, int32_t n_steps, int32_t n_selected, const int32_t* scene_ids,
                    pdc_solve_stats* stats) {
	return guard(ctx, [&] {
		require_bodies(ctx);
		if (n_steps < 0 || n_selected < 0 || (n_selected > 0 && !scene_ids))
			throw std::invalid_argument("pdc_step_scenes: negative count or null scene ids");
		if (n_selected == 0) pdc::set_stepped(ctx, 0, nullptr);
		else pdc::set_stepped(ctx, n_selected, scene_ids);
		step_stepped_scenes(ctx, n_steps);
		if (stats) {
			if (n_selected == 0)
				for (int k = 0; k < ctx->S; ++k) stats[k] = ctx->scene_stats[k];
			else
				for (int k = 0; k < n_selected; ++k) stats[k] = ctx->scene_stats[scene_ids[k]];
		}
	});
}

static void check_shapes(const pdc_shapes& sh) {
	for (int c = 0; c < 4; ++c) {
		const int ct = sh.contact_tile[c], bw = sh.body_waves[c];
		if (ct != 64 && ct != 128 && ct != 256) throw ConfigError("contact_tile must be 64, 128 or 256");
		if (bw != 1 && bw != 2 && bw != 4) throw ConfigError("body_waves must be 1, 2 or 4");
	}
	if (sh.pcg_chunk != 0 && sh.pcg_chunk != 8 && sh.pcg_chunk != 16 && sh.pcg_chunk != 32)
		throw ConfigError("pcg_chunk must be 0 (automatic), 8, 16 or 32");
}

int pdc_get_shapes(const pdc_context* ctx, pdc_shapes* out) {
	if (!ctx || !out) return PDC_ERR_INVALID_ARGUMENT;
	*out = ctx->shapes;
	return PDC_OK;
}

int pdc_set_shapes(pdc_context* ctx, const pdc_shapes* shapes) {
	return guard(ctx, [&] {
		if (!shapes) throw std::invalid_argument("pdc_set_shapes: null shapes");
		check_shapes(*shapes);
		ctx->shapes = *shapes;
	});
}

// The autotuner: each trial restores the saved body state, steps the scenes
// under test and is timed on the host (a timestep ends with a wait); the
// solver's counters, timings and the PCG launch predictor are restored
// afterwards, so tuning leaves no trace but the shape table.
int pdc_autotune(pdc_context* ctx, int32_t steps, pdc_shapes* chosen) {
	return guard(ctx, [&] {
		require_bodies(ctx);
		if (steps < 1) throw std::invalid_argument("pdc_autotune: steps must be >= 1");
		const size_t n = (size_t)ctx->N;
		hipStream_t s = ctx->stream;
		pdc::DBuf<double> sp, sq, sv, sw;
		sp.ensure(3 * n); sq.ensure(4 * n); sv.ensure(3 * n); sw.ensure(3 * n);
		auto copy = [&](double* dst, const double* src, size_t k) {
			PDC_HIP(hipMemcpyAsync(dst, src, k * sizeof(double), hipMemcpyDeviceToDevice, s));
		};
		copy(sp.p, ctx->b_pos.p, 3 * n); copy(sq.p, ctx->b_quat.p, 4 * n);
		copy(sv.p, ctx->b_vel.p, 3 * n); copy(sw.p, ctx->b_ang.p, 3 * n);
		const pdc_timings tm = ctx->timings;
		const pdc_solve_stats last = ctx->last;
		const int last_iters = ctx->pcg_last_iters;
		const int kt_which = ctx->kt.which;
		ctx->kt.which = PDC_KT_NONE;
		auto restore = [&]() {
			copy(ctx->b_pos.p, sp.p, 3 * n); copy(ctx->b_quat.p, sq.p, 4 * n);
			copy(ctx->b_vel.p, sv.p, 3 * n); copy(ctx->b_ang.p, sw.p, 3 * n);
			PDC_HIP(hipStreamSynchronize(s));
			ctx->pcg_last_iters = last_iters;
		};
		// best of two runs: the first run after a shape change also pays for
		// re-recording graphs and growing buffers
		auto trial = [&](const std::vector<int>& ids) {
			if (ids.empty()) pdc::set_stepped(ctx, 0, nullptr);
			else pdc::set_stepped(ctx, (int)ids.size(), ids.data());
			double tmin = 1e300;
			for (int rep = 0; rep < 2; ++rep) {
				restore();
				const double t0 = now_s();
				step_stepped_scenes(ctx, steps);
				tmin = std::min(tmin, now_s() - t0);
			}
			return tmin;
		};
		// scenes per size class
		std::vector<int> by_class[4];
		for (int k = 0; k < ctx->S; ++k)
			by_class[pdc::size_class(ctx->h_sclass[k])].push_back(k);
		const pdc_shapes shapes0 = ctx->shapes;
		pdc_shapes best = ctx->shapes;
		const int tiles[3] = {64, 128, 256}, waves[3] = {1, 2, 4}, chunks[3] = {8, 16, 32};
		for (int c = 0; c < 4; ++c) {
			if (by_class[c].empty()) continue;
			double tbest = 1e300;
			int ct_best = best.contact_tile[c], bw_best = best.body_waves[c];
			for (int ct : tiles)
				for (int bw : waves) {
					ctx->shapes = best;
					ctx->shapes.contact_tile[c] = ct;
					ctx->shapes.body_waves[c] = bw;
					const double t = trial(by_class[c]);
					if (t < tbest) { tbest = t; ct_best = ct; bw_best = bw; }
				}
			best.contact_tile[c] = ct_best;
			best.body_waves[c] = bw_best;
		}
		double tbest = 1e300;
		int ch_best = best.pcg_chunk;
		for (int ch : chunks) {
			ctx->shapes = best;
			ctx->shapes.pcg_chunk = ch;
			const double t = trial({});
			if (t < tbest) { tbest = t; ch_best = ch; }
		}
		best.pcg_chunk = ch_best;
		// the per-class choices are made on each class alone: keep the table
		// only if all scenes together also run faster with it than with the
		// table the tuning started from
		auto timed = [&](const pdc_shapes& sh) {
			ctx->shapes = sh;
			return trial({});
		};
		if (timed(shapes0) <= timed(best)) best = shapes0;
		ctx->shapes = best;
		restore();
		pdc::set_stepped(ctx, 0, nullptr);
		ctx->timings = tm;
		ctx->last = last;
		ctx->kt.which = kt_which;
		if (chosen) *chosen = best;
	});
}

int pdc_get_timings(const pdc_context* ctx, pdc_timings* out) {
	if (!ctx || !out) return PDC_ERR_INVALID_ARGUMENT;
	*out = ctx->timings;
	return PDC_OK;
}

void pdc_reset_timings(pdc_context* ctx) {
	if (ctx) std::memset(&ctx->timings, 0, sizeof(ctx->timings));
}

int pdc_enable_kernel_timing(pdc_context* ctx, int32_t which) {
	return guard(ctx, [&] {
		if (which < PDC_KT_NONE || which > PDC_KT_PCG) throw std::invalid_argument("unknown kernel class");
		pdc::kt_collect(ctx);
		ctx->kt.which = which;
		ctx->kt.total_ms = 0;
		ctx->kt.launches = 0;
		ctx->kt.bytes = 0;
	});
}

int pdc_kernel_timing(pdc_context* ctx, double* total_ms, int64_t* launches, double* algorithmic_bytes) {
	return guard(ctx, [&] {
		pdc::kt_collect(ctx);
		if (total_ms) *total_ms = ctx->kt.total_ms;
		if (launches) *launches = ctx->kt.launches;
		if (algorithmic_bytes) *algorithmic_bytes = ctx->kt.bytes;
	});
}

}  // extern "C"
