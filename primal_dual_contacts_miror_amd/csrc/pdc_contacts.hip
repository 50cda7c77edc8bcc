// pdc_contacts.hip — (a) grid-SDF contact generation on the device.
//
// Replaces the reference's FCL broadphase + narrowphase and its per-pair merge
// (src/Dynamics/DynamicSystem.cpp:219-228, 247-298, 317-383;
// src/Contact/ContactGenerator/ContactGeneratorBase.h:21-68).  Per step:
//   1. per body: rotation and world AABB of its grid box;
//   2. broadphase: every pair a < b (not both static) whose AABBs overlap, in
//      (a, b) order (two passes + scan, so the order is deterministic);
//   3. one workgroup per pair: the surface samples of a against the grid of b
//      and of b against the grid of a; hits (signed distance < 0) are compacted
//      in sample order with wavefront ballots; the hits are sorted by position
//      (lexicographic) in LDS and chain-merged (|dp|^2 <= 1e-8: summed normal,
//      minimum depth), the reference's merge rule;
//   4. scan of per-pair counts; one workgroup per pair writes the final
//      contact list (ids, position, depth, both 6x3 frames) in pair order.
// Normal convention (MeshMeshContactGenerator.cpp:43-55): a sample of body a
// inside b gives normal -grad(sd_b); a sample of b inside a gives +grad(sd_a).
#include <algorithm>
#include <array>
#include <climits>
#include <cstdint>
#include <cmath>
#include <cstring>
#include <map>
#include <string>
#include <utility>

#include "pdc_internal.h"
#include "pdc_reduce.h"

namespace pdc {

// ---------------------------------------------------------------------------
// host: surface samples (body frame)
//   box:    each face (axis a, side -/+) carries a cell-centred nu x nw grid,
//           n = max(1, round(edge / spacing)), in the order axis, side, iu, iw
//   sphere: Fibonacci lattice of max(8, round(4 pi r^2 / spacing^2)) points
//   ellipsoid: the unit Fibonacci lattice scaled by the semi-axes (a, b, c),
//           max(8, round(4 pi (ab + ac + bc) / 3 / spacing^2)) points (the
//           area of the ellipsoid by the p = 1 case of Thomsen's formula,
//           exact for a sphere)
//   mesh:   its vertices, as the reference tests a mesh's vertices against
//           the other mesh (MeshMeshContactGenerator.cpp:27-63)
// ---------------------------------------------------------------------------
static void make_samples(const pdc_context* ctx, int type, const double* dims, double spacing,
                         std::vector<double>& out) {
	if (type == PDC_SHAPE_MESH) {
		const MeshShape& m = *ctx->meshes[(size_t)dims[0]];
		out.insert(out.end(), m.verts.begin(), m.verts.end());
		return;
	}
	if (type == PDC_SHAPE_SPHERE || type == PDC_SHAPE_ELLIPSOID) {
		const double PI = 3.14159265358979323846;
		double sx = dims[0], sy = dims[0], sz = dims[0];
		double area4 = sx * sx;
		if (type == PDC_SHAPE_ELLIPSOID) {
			sy = dims[1];
			sz = dims[2];
			area4 = (sx * sy + sx * sz + sy * sz) / 3.0;
		}
		long n = std::lround(4.0 * PI * area4 / (spacing * spacing));
		if (n < 8) n = 8;
		const double ga = PI * (3.0 - std::sqrt(5.0));
		for (long i = 0; i < n; ++i) {
			const double y = 1.0 - (2.0 * (double)i + 1.0) / (double)n;
			const double rho = std::sqrt(std::max(0.0, 1.0 - y * y));
			const double phi = (double)i * ga;
			out.push_back(sx * (rho * std::cos(phi)));
			out.push_back(sy * y);
			out.push_back(sz * (rho * std::sin(phi)));
		}
		return;
	}
	for (int a = 0; a < 3; ++a) {
		const int u = (a + 1) % 3, w = (a + 2) % 3;
		long nu = std::lround(2.0 * dims[u] / spacing);
		long nw = std::lround(2.0 * dims[w] / spacing);
		if (nu < 1) nu = 1;
		if (nw < 1) nw = 1;
		const double du = 2.0 * dims[u] / (double)nu;
		const double dw = 2.0 * dims[w] / (double)nw;
		for (int s = -1; s <= 1; s += 2)
			for (long iu = 0; iu < nu; ++iu)
				for (long iw = 0; iw < nw; ++iw) {
					double p[3];
					p[a] = (double)s * dims[a];
					p[u] = -dims[u] + ((double)iu + 0.5) * du;
					p[w] = -dims[w] + ((double)iw + 0.5) * dw;
					out.push_back(p[0]);
					out.push_back(p[1]);
					out.push_back(p[2]);
				}
	}
}

// One thread per grid node; blockIdx.y = body.
__global__ __launch_bounds__(256) void k_build_grids(const GridDesc* __restrict__ grids, const int* __restrict__ shape,
                              const double* __restrict__ dims, float* __restrict__ pool) {
	const int b = blockIdx.y;
	const GridDesc g = grids[b];
	const int64_t n3 = (int64_t)g.n * g.n * g.n;
	const double d[3] = {dims[3 * b], dims[3 * b + 1], dims[3 * b + 2]};
	const int type = shape[b];
	if (type == PDC_SHAPE_MESH) return;  // copied from the mesh's own grid
	for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n3;
	     t += (int64_t)gridDim.x * blockDim.x) {
		const int i = (int)(t % g.n);
		const int j = (int)((t / g.n) % g.n);
		const int k = (int)(t / ((int64_t)g.n * g.n));
		const double px = g.lo[0] + (double)i * g.sp[0];
		const double py = g.lo[1] + (double)j * g.sp[1];
		const double pz = g.lo[2] + (double)k * g.sp[2];
		pool[g.offset + t] = (float)analytic_sdf(type, d, px, py, pz);
	}
}

// ---------------------------------------------------------------------------
// Triangle meshes (Mesh::initializeFCLObject, src/CollisionObject/Mesh.cpp:18-110)
// ---------------------------------------------------------------------------
// Triangle records of a mesh (pdc_math.h, kTriRec doubles each).  Vertices
// with equal coordinates are one vertex for the adjacency (glTF files repeat
// vertices per face); triangles of zero area are left out of the distance
// (their vertices stay samples).  Face normals are normalised cross products;
// the pseudonormal of an edge is the normalised sum of the normals of the
// faces sharing it (libigl's uniform edge weighting, Mesh.cpp:66-68), of a
// vertex the normalised sum of its faces' normals weighted by the face angle
// at the vertex (angle weighting, Mesh.cpp:64-65).  Sums run in face order.
static void mesh_records(int nv, const double* v, int nt, const int* tri, std::vector<double>& rec) {
	std::vector<int> canon((size_t)nv);
	std::map<std::array<double, 3>, int> first;
	for (int i = 0; i < nv; ++i) {
		const std::array<double, 3> key{v[3 * i], v[3 * i + 1], v[3 * i + 2]};
		canon[i] = first.emplace(key, i).first->second;
	}
	std::vector<double> fn(3 * (size_t)nt, 0.0), vn(3 * (size_t)nv, 0.0);
	std::vector<char> ok((size_t)nt, 0);
	std::map<std::pair<int, int>, std::array<double, 3>> en;
	for (int f = 0; f < nt; ++f) {
		const double* P[3] = {&v[3 * tri[3 * f]], &v[3 * tri[3 * f + 1]], &v[3 * tri[3 * f + 2]]};
		const double e1[3] = {P[1][0] - P[0][0], P[1][1] - P[0][1], P[1][2] - P[0][2]};
		const double e2[3] = {P[2][0] - P[0][0], P[2][1] - P[0][1], P[2][2] - P[0][2]};
		double n[3];
		cross3(e1, e2, n);
		const double len = std::sqrt(dot3(n, n));
		if (!(len > 0.0)) continue;
		ok[f] = 1;
		for (int k = 0; k < 3; ++k) fn[3 * (size_t)f + k] = n[k] / len;
		const double* nf = &fn[3 * (size_t)f];
		for (int k = 0; k < 3; ++k) {
			const double* o = P[k];
			const double* u = P[(k + 1) % 3];
			const double* w = P[(k + 2) % 3];
			const double du[3] = {u[0] - o[0], u[1] - o[1], u[2] - o[2]};
			const double dw[3] = {w[0] - o[0], w[1] - o[1], w[2] - o[2]};
			double cs = dot3(du, dw) / (std::sqrt(dot3(du, du)) * std::sqrt(dot3(dw, dw)));
			cs = cs < -1.0 ? -1.0 : cs > 1.0 ? 1.0 : cs;
			const double ang = std::acos(cs);
			double* vk = &vn[3 * (size_t)canon[tri[3 * f + k]]];
			for (int c = 0; c < 3; ++c) vk[c] += ang * nf[c];
			const int i0 = canon[tri[3 * f + k]], i1 = canon[tri[3 * f + (k + 1) % 3]];
			auto it = en.emplace(std::make_pair(std::min(i0, i1), std::max(i0, i1)),
			                     std::array<double, 3>{0.0, 0.0, 0.0}).first;
			for (int c = 0; c < 3; ++c) it->second[c] += nf[c];
		}
	}
	auto unit = [](double* x) {
		const double l = std::sqrt(dot3(x, x));
		if (l > 0.0)
			for (int c = 0; c < 3; ++c) x[c] /= l;
	};
	for (int i = 0; i < nv; ++i) unit(&vn[3 * (size_t)i]);
	for (auto& kv : en) unit(kv.second.data());
	rec.clear();
	for (int f = 0; f < nt; ++f) {
		if (!ok[f]) continue;
		const int c3[3] = {canon[tri[3 * f]], canon[tri[3 * f + 1]], canon[tri[3 * f + 2]]};
		for (int k = 0; k < 3; ++k)
			for (int c = 0; c < 3; ++c) rec.push_back(v[3 * (size_t)tri[3 * f + k] + c]);
		for (int c = 0; c < 3; ++c) rec.push_back(fn[3 * (size_t)f + c]);
		for (int k = 0; k < 3; ++k) {
			const auto& e = en.at(std::make_pair(std::min(c3[k], c3[(k + 1) % 3]), std::max(c3[k], c3[(k + 1) % 3])));
			for (int c = 0; c < 3; ++c) rec.push_back(e[c]);
		}
		for (int k = 0; k < 3; ++k)
			for (int c = 0; c < 3; ++c) rec.push_back(vn[3 * (size_t)c3[k] + c]);
	}
}

int register_mesh(pdc_context* ctx, int nv, const double* v, int nt, const int* tri) {
	auto m = std::make_unique<MeshShape>();
	m->verts.assign(v, v + 3 * (size_t)nv);
	mesh_records(nv, v, nt, tri, m->recs);
	if (m->recs.empty()) throw std::invalid_argument("pdc_register_mesh: every triangle has zero area");
	// grid box: the largest |coordinate| per axis, at least 1e-3 of the largest
	// (so that a flat mesh still gets a box of nonzero thickness)
	double emax = 0.0;
	for (int a = 0; a < 3; ++a) {
		for (int i = 0; i < nv; ++i) m->ext[a] = std::max(m->ext[a], std::fabs(v[3 * (size_t)i + a]));
		emax = std::max(emax, m->ext[a]);
	}
	if (!(emax > 0.0)) throw std::invalid_argument("pdc_register_mesh: all vertices at the origin");
	for (int a = 0; a < 3; ++a) m->ext[a] = std::max(m->ext[a], 1e-3 * emax);
	ctx->meshes.push_back(std::move(m));
	return (int)ctx->meshes.size() - 1;
}

// Half extents of a body's shape before padding.
static void shape_extent(const pdc_context* ctx, int shape, const double* dims, double* e) {
	if (shape == PDC_SHAPE_SPHERE) e[0] = e[1] = e[2] = dims[0];
	else if (shape == PDC_SHAPE_MESH) {
		const MeshShape& m = *ctx->meshes[(size_t)dims[0]];
		e[0] = m.ext[0]; e[1] = m.ext[1]; e[2] = m.ext[2];
	} else { e[0] = dims[0]; e[1] = dims[1]; e[2] = dims[2]; }
}

// Signed distance of every node of one mesh grid; one thread per node, the
// triangle records staged through LDS 32 at a time.  The grid is launched
// with exactly one thread per node rounded up to the block, so every thread
// of a block reaches each barrier.
constexpr int kMeshChunk = 32;
__global__ __launch_bounds__(256) void k_mesh_grid(const double* __restrict__ recs, int nt, GridDesc g,
                                                   float* __restrict__ out) {
	__shared__ double st[kMeshChunk * kTriRec];
	const int64_t n3 = (int64_t)g.n * g.n * g.n;
	const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
	double p[3] = {0.0, 0.0, 0.0};
	if (t < n3) {
		const int i = (int)(t % g.n);
		const int j = (int)((t / g.n) % g.n);
		const int k = (int)(t / ((int64_t)g.n * g.n));
		p[0] = g.lo[0] + (double)i * g.sp[0];
		p[1] = g.lo[1] + (double)j * g.sp[1];
		p[2] = g.lo[2] + (double)k * g.sp[2];
	}
	double best = INFINITY, sgn = 1.0;
	for (int c0 = 0; c0 < nt; c0 += kMeshChunk) {
		const int m = min(kMeshChunk, nt - c0);
		__syncthreads();
		for (int x = threadIdx.x; x < m * kTriRec; x += blockDim.x) st[x] = recs[(size_t)c0 * kTriRec + x];
		__syncthreads();
		if (t < n3)
			for (int j = 0; j < m; ++j) mesh_sdf_step(&st[j * kTriRec], p, &best, &sgn);
	}
	if (t < n3) out[t] = (float)(sgn * sqrt(best));
}

// Grid box of a body: the shape's box padded by grid_pad_frac, its node block
// at body * grid_n^3 in the pool.
static GridDesc grid_desc(const pdc_context* ctx, int shape, const double* dims, int64_t body) {
	const pdc_params& p = ctx->p;
	GridDesc g{};
	double e[3];
	shape_extent(ctx, shape, dims, e);
	const int gn = p.grid_n;
	for (int a = 0; a < 3; ++a) {
		const double h = e[a] * (1.0 + p.grid_pad_frac);
		g.half[a] = h;
		g.lo[a] = -h;
		g.sp[a] = (2.0 * h) / (double)(gn - 1);
	}
	g.n = gn;
	g.offset = body * (int64_t)gn * gn * gn;
	return g;
}

// Device grid of mesh `id` at the context's grid_n and padding (built on
// first use, rebuilt when those change).
static const float* mesh_grid(pdc_context* ctx, int id) {
	MeshShape& m = *ctx->meshes[(size_t)id];
	const int gn = ctx->p.grid_n;
	if (m.grid_n == gn && m.grid_pad == ctx->p.grid_pad_frac) return m.grid.p;
	hipStream_t s = ctx->stream;
	const int nt = (int)(m.recs.size() / kTriRec);
	m.d_recs.ensure(m.recs.size());
	PDC_HIP(hipMemcpyAsync(m.d_recs.p, m.recs.data(), m.recs.size() * sizeof(double), hipMemcpyHostToDevice, s));
	GridDesc g = grid_desc(ctx, PDC_SHAPE_MESH, std::array<double, 3>{(double)id, 0.0, 0.0}.data(), 0);
	const int64_t n3 = (int64_t)gn * gn * gn;
	m.grid.ensure((size_t)n3);
	hipLaunchKernelGGL(k_mesh_grid, dim3(grid_for(n3)), dim3(kBlock), 0, s, m.d_recs.p, nt, g, m.grid.p);
	PDC_CHECK_LAUNCH();
	PDC_HIP(hipStreamSynchronize(s));
	m.grid_n = gn;
	m.grid_pad = ctx->p.grid_pad_frac;
	return m.grid.p;
}

// Flags, dimensions, masses, inertia and grid descriptors of bodies
// [first, first + n), their grids built; the tie band narrowed to the finest
// spacing among them (`reset`: start it afresh).
static void write_bodies(pdc_context* ctx, int first, int n, const int* shape, const double* dims,
                         const double* mass, const double* inertia, const int* is_static, const int* emit,
                         bool reset) {
	if (n <= 0) return;
	const size_t N = (size_t)n, f = (size_t)first;
	std::vector<int> em(N);
	for (size_t i = 0; i < N; ++i) em[i] = emit ? (emit[i] != 0) : 1;
	hipStream_t s = ctx->stream;
	PDC_HIP(hipMemcpyAsync(ctx->b_shape.p + f, shape, N * sizeof(int), hipMemcpyHostToDevice, s));
	PDC_HIP(hipMemcpyAsync(ctx->b_static.p + f, is_static, N * sizeof(int), hipMemcpyHostToDevice, s));
	PDC_HIP(hipMemcpyAsync(ctx->b_emit.p + f, em.data(), N * sizeof(int), hipMemcpyHostToDevice, s));
	PDC_HIP(hipMemcpyAsync(ctx->b_dims.p + 3 * f, dims, 3 * N * sizeof(double), hipMemcpyHostToDevice, s));
	PDC_HIP(hipMemcpyAsync(ctx->b_mass.p + f, mass, N * sizeof(double), hipMemcpyHostToDevice, s));
	PDC_HIP(hipMemcpyAsync(ctx->b_I0.p + 3 * f, inertia, 3 * N * sizeof(double), hipMemcpyHostToDevice, s));
	if (ctx->h_static.size() < f + N) ctx->h_static.resize(f + N);
	if (ctx->h_emit.size() < f + N) ctx->h_emit.resize(f + N);
	std::copy(is_static, is_static + N, ctx->h_static.begin() + f);
	std::copy(em.begin(), em.end(), ctx->h_emit.begin() + f);
	std::vector<GridDesc> gd(N);
	double finest = 1e300;
	for (size_t b = 0; b < N; ++b) {
		gd[b] = grid_desc(ctx, shape[b], &dims[3 * b], (int64_t)(f + b));
		for (int a = 0; a < 3; ++a) finest = std::min(finest, gd[b].sp[a]);
	}
	// tie band: 1e-7 of the finest grid spacing of the table
	ctx->tie_band = (reset || ctx->tie_band <= 0) ? 1e-7 * finest : std::min(ctx->tie_band, 1e-7 * finest);
	PDC_HIP(hipMemcpyAsync(ctx->b_grid.p + f, gd.data(), N * sizeof(GridDesc), hipMemcpyHostToDevice, s));
	const int gn = ctx->p.grid_n;
	const int64_t n3 = (int64_t)gn * gn * gn;
	const int gx = (int)std::min<int64_t>((n3 + kBlock - 1) / kBlock, 4096);
	for (size_t b0 = 0; b0 < N; b0 += 65535) {
		const int nb = (int)std::min<size_t>(65535, N - b0);
		hipLaunchKernelGGL(k_build_grids, dim3(gx, nb), dim3(kBlock), 0, s, ctx->b_grid.p + f + b0,
		                   ctx->b_shape.p + f + b0, ctx->b_dims.p + 3 * (f + b0), ctx->grid_pool.p);
		PDC_CHECK_LAUNCH();
	}
	for (size_t b = 0; b < N; ++b) {
		if (shape[b] != PDC_SHAPE_MESH) continue;
		const float* mg = mesh_grid(ctx, (int)dims[3 * b]);
		PDC_HIP(hipMemcpyAsync(ctx->grid_pool.p + gd[b].offset, mg, (size_t)n3 * sizeof(float),
		                       hipMemcpyDeviceToDevice, s));
	}
}

// Cells of one body's samples: a kc^3 grid over the box of its FP32
// samples, kc the largest of 4, 3, 2, 1 with kc^3 <= S / 8 (so a body's cells
// fit in the S / 8 table entries from floor(soff / 8)), none below 8 samples.
// The samples are reordered by cell (stably) and appended to `out`; each
// non-empty cell, in cell order, gets the box of its samples.
static void sort_into_cells(const double* smp, int S, std::vector<float>& out, std::vector<SampleCell>& cells) {
	std::vector<float> f(3 * (size_t)S);
	for (size_t i = 0; i < f.size(); ++i) f[i] = (float)smp[i];
	const int kc = S >= 512 ? 4 : S >= 216 ? 3 : S >= 64 ? 2 : S >= 8 ? 1 : 0;
	if (kc == 0) {
		out.insert(out.end(), f.begin(), f.end());
		return;
	}
	float lo[3] = {f[0], f[1], f[2]}, hi[3] = {f[0], f[1], f[2]};
	for (int i = 0; i < S; ++i)
		for (int a = 0; a < 3; ++a) {
			lo[a] = std::min(lo[a], f[3 * (size_t)i + a]);
			hi[a] = std::max(hi[a], f[3 * (size_t)i + a]);
		}
	std::vector<int> cell(S);
	const int nc = kc * kc * kc;
	std::vector<int> cnt(nc + 1, 0);
	for (int i = 0; i < S; ++i) {
		int c[3];
		for (int a = 0; a < 3; ++a) {
			const double w = (double)hi[a] - (double)lo[a];
			int k = w > 0.0 ? (int)(((double)f[3 * (size_t)i + a] - (double)lo[a]) / w * kc) : 0;
			c[a] = std::min(std::max(k, 0), kc - 1);
		}
		cell[i] = (c[2] * kc + c[1]) * kc + c[0];
		cnt[cell[i] + 1] += 1;
	}
	for (int k = 0; k < nc; ++k) cnt[k + 1] += cnt[k];
	std::vector<int> order(S), cur(cnt.begin(), cnt.end() - 1);
	for (int i = 0; i < S; ++i) order[cur[cell[i]]++] = i;
	for (int k = 0; k < nc; ++k) {
		if (cnt[k + 1] == cnt[k]) continue;
		SampleCell sc{};
		for (int a = 0; a < 3; ++a) { sc.lo[a] = INFINITY; sc.hi[a] = -INFINITY; }
		for (int j = cnt[k]; j < cnt[k + 1]; ++j)
			for (int a = 0; a < 3; ++a) {
				const float v = f[3 * (size_t)order[j] + a];
				sc.lo[a] = std::min(sc.lo[a], v);
				sc.hi[a] = std::max(sc.hi[a], v);
			}
		sc.beg = cnt[k];
		sc.end = cnt[k + 1];
		cells.push_back(sc);
	}
	for (int j = 0; j < S; ++j)
		for (int a = 0; a < 3; ++a) out.push_back(f[3 * (size_t)order[j] + a]);
}

// FP32 body-frame surface samples of n bodies, ordered by cell; counts[b]
// samples and ncell[b] cells per body, the cells of all n bodies in `cells`.
// A body that emits no samples (emit[b] == 0) stores none: its samples would
// never be tested.
static void samples_of(const pdc_context* ctx, int n, const int* shape, const double* dims, const int* emit,
                       std::vector<float>& smp, std::vector<int>& counts, std::vector<SampleCell>& cells,
                       std::vector<int>& ncell) {
	counts.assign(n, 0);
	ncell.assign(n, 0);
	std::vector<double> one;
	for (int b = 0; b < n; ++b) {
		if (emit && !emit[b]) continue;
		one.clear();
		make_samples(ctx, shape[b], &dims[3 * b], ctx->p.sample_spacing, one);
		const int S = (int)(one.size() / 3);
		const size_t c0 = cells.size();
		sort_into_cells(one.data(), S, smp, cells);
		counts[b] = S;
		ncell[b] = (int)(cells.size() - c0);
	}
}

// Cells of bodies [first, first + n) to their places in the cell table (the
// sample offsets of those bodies already set in h_soff), in one copy: body
// b's cells start at floor(soff[b] / kCellStride) and end before the next
// body's (its kc^3 <= S / 8 cells), so the span from the first body's start
// to the last body's end holds only these bodies' cells and unused slots.
static void upload_cells(pdc_context* ctx, int first, int n, const std::vector<SampleCell>& cells,
                         const std::vector<int>& ncell) {
	hipStream_t s = ctx->stream;
	size_t lo = SIZE_MAX, hi = 0, k = 0;
	for (int b = 0; b < n; ++b)
		if (ncell[b] > 0) {
			const size_t at = (size_t)ctx->h_soff[first + b] / kCellStride;
			lo = std::min(lo, at);
			hi = std::max(hi, at + (size_t)ncell[b]);
		}
	if (hi > lo) {
		std::vector<SampleCell> span(hi - lo, SampleCell{});
		for (int b = 0; b < n; ++b) {
			if (ncell[b] > 0) {
				const size_t at = (size_t)ctx->h_soff[first + b] / kCellStride;
				std::copy(cells.begin() + k, cells.begin() + k + ncell[b], span.begin() + (at - lo));
			}
			k += (size_t)ncell[b];
		}
		PDC_HIP(hipMemcpyAsync(ctx->cells.p + lo, span.data(), span.size() * sizeof(SampleCell),
		                       hipMemcpyHostToDevice, s));
	}
	if (n > 0)
		PDC_HIP(hipMemcpyAsync(ctx->b_ncell.p + first, ncell.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice,
		                       s));
	PDC_HIP(hipStreamSynchronize(s));
}

static size_t cell_slots(int64_t samples) { return (size_t)(samples / kCellStride) + 1; }

void upload_bodies(pdc_context* ctx, int n, const int* shape, const double* dims, const double* mass,
                   const double* inertia, const int* is_static, const int* emit) {
	ctx->N = n;
	const size_t N = (size_t)n;
	ctx->b_shape.ensure(N); ctx->b_static.ensure(N); ctx->b_emit.ensure(N);
	ctx->b_dims.ensure(3 * N); ctx->b_mass.ensure(N); ctx->b_I0.ensure(3 * N);
	ctx->b_pos.ensure(3 * N); ctx->b_quat.ensure(4 * N); ctx->b_vel.ensure(3 * N);
	ctx->b_ang.ensure(3 * N); ctx->b_acc.ensure(6 * N);
	ctx->b_R.ensure(9 * N); ctx->b_aabb.ensure(6 * N);
	ctx->b_grid.ensure(N); ctx->b_soff.ensure(N + 1); ctx->b_ncell.ensure(N);
	ctx->h_static.assign(N, 0);
	ctx->h_emit.assign(N, 0);
	const int gn = ctx->p.grid_n;
	ctx->grid_pool.ensure(std::max<size_t>(N * (size_t)gn * gn * gn, 1));
	write_bodies(ctx, 0, n, shape, dims, mass, inertia, is_static, emit, true);
	if (n == 0) ctx->tie_band = 0.0;
	// samples
	std::vector<float> smp;
	std::vector<int> cnt, ncell;
	std::vector<SampleCell> cells;
	samples_of(ctx, n, shape, dims, emit, smp, cnt, cells, ncell);
	std::vector<int64_t> total(N + 1, 0);
	for (size_t b = 0; b < N; ++b) total[b + 1] = total[b] + cnt[b];
	if (total[N] > INT32_MAX) throw HipError("sample table beyond 2^31 samples", PDC_ERR_CAPACITY);
	std::vector<int> soff(N + 1, 0);
	for (size_t b = 0; b <= N; ++b) soff[b] = (int)total[b];
	ctx->n_samples = soff[N];
	ctx->h_soff = soff;
	hipStream_t s = ctx->stream;
	ctx->samples.ensure(std::max<size_t>(smp.size(), 3));
	ctx->cells.ensure(cell_slots(soff[N]));
	if (!smp.empty())
		PDC_HIP(hipMemcpyAsync(ctx->samples.p, smp.data(), smp.size() * sizeof(float),
		                       hipMemcpyHostToDevice, s));
	PDC_HIP(hipMemcpyAsync(ctx->b_soff.p, soff.data(), (N + 1) * sizeof(int), hipMemcpyHostToDevice, s));
	upload_cells(ctx, 0, n, cells, ncell);
	const int one[2] = {0, (int)N};
	set_scenes(ctx, 1, one);
}

void reserve_bodies(pdc_context* ctx, int64_t bodies, int64_t samples) {
	if (bodies > INT32_MAX || samples > INT32_MAX) throw HipError("reservation beyond 2^31", PDC_ERR_CAPACITY);
	const size_t N0 = (size_t)ctx->N, N1 = std::max<size_t>((size_t)bodies, N0);
	hipStream_t s = ctx->stream;
	PDC_HIP(hipStreamSynchronize(s));
	ctx->b_shape.grow_keep(N1, N0, s, true); ctx->b_static.grow_keep(N1, N0, s, true); ctx->b_emit.grow_keep(N1, N0, s, true);
	ctx->b_dims.grow_keep(3 * N1, 3 * N0, s, true); ctx->b_mass.grow_keep(N1, N0, s, true);
	ctx->b_I0.grow_keep(3 * N1, 3 * N0, s, true);
	ctx->b_pos.grow_keep(3 * N1, 3 * N0, s, true); ctx->b_quat.grow_keep(4 * N1, 4 * N0, s, true);
	ctx->b_vel.grow_keep(3 * N1, 3 * N0, s, true); ctx->b_ang.grow_keep(3 * N1, 3 * N0, s, true);
	ctx->b_acc.grow_keep(6 * N1, 6 * N0, s, true);
	ctx->b_R.grow_keep(9 * N1, 0, s, true); ctx->b_aabb.grow_keep(6 * N1, 0, s, true);
	ctx->b_grid.grow_keep(N1, N0, s, true); ctx->b_soff.grow_keep(N1 + 1, N0 + 1, s, true);
	ctx->b_ncell.grow_keep(N1, N0, s, true);
	const size_t n3 = (size_t)ctx->p.grid_n * ctx->p.grid_n * ctx->p.grid_n;
	ctx->grid_pool.grow_keep(std::max<size_t>(N1 * n3, 1), N0 * n3, s, true);
	const size_t total0 = ctx->h_soff.empty() ? 0 : (size_t)ctx->h_soff[N0];
	const size_t total1 = std::max<size_t>((size_t)samples, std::max<size_t>(total0, 1));
	ctx->samples.grow_keep(3 * total1, 3 * total0, s, true);
	ctx->cells.grow_keep(cell_slots((int64_t)total1), cell_slots((int64_t)total0), s, true);
}

int append_bodies(pdc_context* ctx, int n, const int* shape, const double* dims, const double* mass,
                  const double* inertia, const int* is_static, const int* emit) {
	const size_t N0 = (size_t)ctx->N, N1 = N0 + (size_t)n;
	if ((int64_t)N1 > INT32_MAX) throw HipError("body table beyond 2^31 bodies", PDC_ERR_CAPACITY);
	hipStream_t s = ctx->stream;
	PDC_HIP(hipStreamSynchronize(s));
	ctx->b_shape.grow_keep(N1, N0, s); ctx->b_static.grow_keep(N1, N0, s); ctx->b_emit.grow_keep(N1, N0, s);
	ctx->b_dims.grow_keep(3 * N1, 3 * N0, s); ctx->b_mass.grow_keep(N1, N0, s);
	ctx->b_I0.grow_keep(3 * N1, 3 * N0, s);
	ctx->b_pos.grow_keep(3 * N1, 3 * N0, s); ctx->b_quat.grow_keep(4 * N1, 4 * N0, s);
	ctx->b_vel.grow_keep(3 * N1, 3 * N0, s); ctx->b_ang.grow_keep(3 * N1, 3 * N0, s);
	ctx->b_acc.grow_keep(6 * N1, 6 * N0, s);
	ctx->b_R.grow_keep(9 * N1, 0, s); ctx->b_aabb.grow_keep(6 * N1, 0, s);
	ctx->b_grid.grow_keep(N1, N0, s); ctx->b_soff.grow_keep(N1 + 1, N0 + 1, s);
	ctx->b_ncell.grow_keep(N1, N0, s);
	const size_t n3 = (size_t)ctx->p.grid_n * ctx->p.grid_n * ctx->p.grid_n;
	ctx->grid_pool.grow_keep(std::max<size_t>(N1 * n3, 1), N0 * n3, s);
	std::vector<float> smp;
	std::vector<int> cnt, ncell;
	std::vector<SampleCell> cells;
	samples_of(ctx, n, shape, dims, emit, smp, cnt, cells, ncell);
	const int64_t total0 = ctx->h_soff.empty() ? 0 : ctx->h_soff[N0];
	const int64_t total1 = total0 + (int64_t)(smp.size() / 3);
	if (total1 > INT32_MAX) throw HipError("sample table beyond 2^31 samples", PDC_ERR_CAPACITY);
	ctx->samples.grow_keep(std::max<size_t>(3 * (size_t)total1, 3), 3 * (size_t)total0, s);
	ctx->cells.grow_keep(cell_slots(total1), cell_slots(total0), s);
	ctx->h_soff.resize(N1 + 1);
	if (N0 == 0) ctx->h_soff[0] = 0;
	for (size_t b = 0; b < (size_t)n; ++b) ctx->h_soff[N0 + b + 1] = ctx->h_soff[N0 + b] + cnt[b];
	ctx->n_samples = ctx->h_soff[N1];
	if (!smp.empty())
		PDC_HIP(hipMemcpyAsync(ctx->samples.p + 3 * total0, smp.data(), smp.size() * sizeof(float),
		                       hipMemcpyHostToDevice, s));
	PDC_HIP(hipMemcpyAsync(ctx->b_soff.p + N0, ctx->h_soff.data() + N0, ((size_t)n + 1) * sizeof(int),
	                       hipMemcpyHostToDevice, s));
	upload_cells(ctx, (int)N0, n, cells, ncell);
	ctx->N = (int)N1;
	write_bodies(ctx, (int)N0, n, shape, dims, mass, inertia, is_static, emit, N0 == 0);
	PDC_HIP(hipStreamSynchronize(s));
	// the new bodies form a new last scene; the other scenes keep their friction
	std::vector<int> sbeg = ctx->h_sbeg;
	std::vector<double> mu = ctx->h_mu;
	if (N0 == 0) sbeg.assign(1, 0);
	sbeg.push_back((int)N1);
	set_scenes(ctx, (int)sbeg.size() - 1, sbeg.data());
	for (size_t k = 0; k < mu.size() && k + 1 < sbeg.size() && N0 > 0; ++k) ctx->h_mu[k] = mu[k];
	return (int)N0;
}

void replace_bodies(pdc_context* ctx, int first, int n, const int* shape, const double* dims, const double* mass,
                    const double* inertia, const int* is_static, const int* emit) {
	if (first < 0 || n < 1 || first + n > ctx->N) throw std::invalid_argument("body range outside the table");
	std::vector<float> smp;
	std::vector<int> cnt, ncell;
	std::vector<SampleCell> cells;
	samples_of(ctx, n, shape, dims, emit, smp, cnt, cells, ncell);
	const int64_t lo = ctx->h_soff[first], hi = ctx->h_soff[first + n];
	const int64_t m = (int64_t)(smp.size() / 3);
	const bool last_emits = emit ? emit[n - 1] != 0 : true;
	if (m > hi - lo)
		throw HipError("the replacing bodies have " + std::to_string(m) + " samples, more than the " +
		                   std::to_string(hi - lo) + " their range holds",
		               PDC_ERR_CAPACITY);
	if (m < hi - lo && last_emits)
		throw std::invalid_argument("the last replacing body must emit no samples when the samples do not "
		                            "fill the range (it keeps the rest)");
	hipStream_t s = ctx->stream;
	PDC_HIP(hipStreamSynchronize(s));
	for (int b = 0; b + 1 < n; ++b) ctx->h_soff[first + b + 1] = ctx->h_soff[first + b] + cnt[b];
	if (!smp.empty())
		PDC_HIP(hipMemcpyAsync(ctx->samples.p + 3 * lo, smp.data(), smp.size() * sizeof(float),
		                       hipMemcpyHostToDevice, s));
	PDC_HIP(hipMemcpyAsync(ctx->b_soff.p + first, ctx->h_soff.data() + first, ((size_t)n + 1) * sizeof(int),
	                       hipMemcpyHostToDevice, s));
	upload_cells(ctx, first, n, cells, ncell);
	write_bodies(ctx, first, n, shape, dims, mass, inertia, is_static, emit, false);
	PDC_HIP(hipStreamSynchronize(s));
	count_scene_classes(ctx);
}

void count_scene_classes(pdc_context* ctx) {
	const int S = ctx->S;
	ctx->h_sclass.assign(S, 0);
	for (int k = 0; k < S; ++k)
		for (int b = ctx->h_sbeg[k]; b < ctx->h_sbeg[k + 1]; ++b)
			if (!ctx->h_static[b] || ctx->h_emit[b]) ctx->h_sclass[k] += 1;
}

void set_scenes(pdc_context* ctx, int S, const int* sbeg) {
	const int N = ctx->N;
	if (S < 1) throw std::invalid_argument("the scene count must be at least 1");
	if (sbeg[0] != 0 || sbeg[S] != N)
		throw std::invalid_argument("scene offsets must start at 0 and end at the body count");
	for (int k = 0; k < S; ++k)
		if (sbeg[k + 1] < sbeg[k]) throw std::invalid_argument("scene offsets must be non-decreasing");
	hipStream_t s = ctx->stream;
	PDC_HIP(hipStreamSynchronize(s));
	ctx->S = S;
	ctx->h_sbeg.assign(sbeg, sbeg + S + 1);
	std::vector<int> bs(std::max(N, 1), 0);
	int cap = 0;
	for (int k = 0; k < S; ++k) {
		for (int b = sbeg[k]; b < sbeg[k + 1]; ++b) bs[b] = k;
		cap += (sbeg[k + 1] - sbeg[k] + kPcgRowsPerWG - 1) / kPcgRowsPerWG;  // PCG tiles of the scene at most
	}
	ctx->cap_tr = std::max(cap, 1);
	ctx->b_scene.ensure(bs.size());
	PDC_HIP(hipMemcpyAsync(ctx->b_scene.p, bs.data(), bs.size() * sizeof(int), hipMemcpyHostToDevice, s));
	ctx->d_sbeg.ensure(S + 1);
	PDC_HIP(hipMemcpyAsync(ctx->d_sbeg.p, sbeg, (S + 1) * sizeof(int), hipMemcpyHostToDevice, s));
	ctx->h_mu.assign(S, ctx->p.friction_coeff);
	ctx->d_ss.ensure(S);
	ctx->d_sctl.ensure(S);
	ctx->d_act.ensure(S);
	ctx->h_sctl.assign(S, SceneCtl{});
	ctx->d_ps.ensure(S);
	PDC_HIP(hipMemsetAsync(ctx->d_ps.p, 0, S * sizeof(PcgScene), s));
	if ((size_t)S > ctx->h_prog_cap) {
		// no kernel of an earlier step writes it any more: every solve ends with a wait
		PDC_HIP(hipStreamSynchronize(s));
		if (ctx->h_prog) PDC_HIP(hipHostFree(ctx->h_prog));
		ctx->h_prog = nullptr;
		ctx->d_prog = nullptr;
		ctx->h_prog_cap = 0;
		PDC_HIP(hipHostMalloc((void**)&ctx->h_prog, S * sizeof(PcgProgress), hipHostMallocMapped | hipHostMallocCoherent));
		PDC_HIP(hipHostGetDevicePointer((void**)&ctx->d_prog, ctx->h_prog, 0));
		ctx->h_prog_cap = S;
	}
	std::memset(ctx->h_prog, 0, S * sizeof(PcgProgress));
	if ((size_t)S > ctx->h_ss_cap) {
		if (ctx->h_ss) PDC_HIP(hipHostFree(ctx->h_ss));
		ctx->h_ss = nullptr;
		ctx->h_ss_cap = 0;
		PDC_HIP(hipHostMalloc((void**)&ctx->h_ss, S * sizeof(SceneSc), hipHostMallocDefault));
		ctx->h_ss_cap = S;
	}
	ctx->scene_stats.assign(S, pdc_solve_stats{});
	count_scene_classes(ctx);
	set_stepped(ctx, 0, nullptr);
}

void set_stepped(pdc_context* ctx, int n, const int* ids) {
	const int S = ctx->S;
	ctx->h_step.assign(S, ids ? 0 : 1);
	if (ids)
		for (int k = 0; k < n; ++k) {
			if (ids[k] < 0 || ids[k] >= S) throw std::invalid_argument("scene id out of range");
			ctx->h_step[ids[k]] = 1;
		}
	ctx->d_step.ensure(S);
	PDC_HIP(hipMemcpyAsync(ctx->d_step.p, ctx->h_step.data(), S * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
	PDC_HIP(hipStreamSynchronize(ctx->stream));
}

// ---------------------------------------------------------------------------
// per body: rotation and world AABB of the grid box
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_body_frames(int N, const double* __restrict__ quat, const double* __restrict__ pos,
                              const GridDesc* __restrict__ grids, double* __restrict__ Rout,
                              double* __restrict__ aabb) {
	const int b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= N) return;
	double q[4] = {quat[4 * b], quat[4 * b + 1], quat[4 * b + 2], quat[4 * b + 3]};
	double R[9];
	quat_rot(q, R);
	for (int k = 0; k < 9; ++k) Rout[9 * b + k] = R[k];
	const GridDesc g = grids[b];
	for (int r = 0; r < 3; ++r) {
		const double h = fabs(R[r * 3 + 0]) * g.half[0] + fabs(R[r * 3 + 1]) * g.half[1] +
		                 fabs(R[r * 3 + 2]) * g.half[2];
		aabb[6 * b + r] = pos[3 * b + r] - h;
		aabb[6 * b + 3 + r] = pos[3 * b + r] + h;
	}
}

__device__ __forceinline__ bool aabb_overlap(const double* A, const double* B) {
	for (int r = 0; r < 3; ++r)
		if (A[r] > B[3 + r] || B[r] > A[3 + r]) return false;
	return true;
}

// Broadphase over the scene's bodies, one wavefront per body a: the lanes
// test the candidates b > a of a's scene 64 at a time, and a ballot gives the
// count (pass 1) or each hit's place in the (a, b)-ordered output (pass 2).
// Static-static pairs are skipped (DynamicSystem.cpp:326-331).  Only the
// scenes stepped by this call take part.
constexpr int kBpWaves = 4;

__device__ __forceinline__ bool bp_candidate(const double* __restrict__ aabb, const int* __restrict__ stat,
                                             const double (&A)[6], int sa, int b, int end) {
	if (b >= end || (sa && stat[b])) return false;
	const double* B = aabb + 6 * (size_t)b;
	return aabb_overlap(A, B);
}

__global__ __launch_bounds__(64 * kBpWaves) void k_bp_count(int N, const double* __restrict__ aabb,
                                                          const int* __restrict__ stat,
                                                          const int* __restrict__ bscene,
                                                          const int* __restrict__ sbeg,
                                                          const int* __restrict__ stepped, int* __restrict__ cnt) {
	const int a = blockIdx.x * kBpWaves + (int)(threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (a >= N) return;
	const int sc = bscene[a];
	int c = 0;
	if (stepped[sc]) {
		const int end = sbeg[sc + 1];
		double A[6];
		for (int k = 0; k < 6; ++k) A[k] = aabb[6 * (size_t)a + k];
		const int sa = stat[a];
		for (int base = a + 1; base < end; base += 64)
			c += __popcll(__ballot(bp_candidate(aabb, stat, A, sa, base + lane, end)));
	}
	if (lane == 0) cnt[a] = c;
}

__global__ __launch_bounds__(64 * kBpWaves) void k_bp_write(int N, const double* __restrict__ aabb,
                                                          const int* __restrict__ stat,
                                                          const int* __restrict__ bscene,
                                                          const int* __restrict__ sbeg,
                                                          const int* __restrict__ stepped,
                                                          const int* __restrict__ off, int cap,
                                                          int* __restrict__ pa, int* __restrict__ pb) {
	const int a = blockIdx.x * kBpWaves + (int)(threadIdx.x >> 6);
	const int lane = threadIdx.x & 63;
	if (a >= N) return;
	const int sc = bscene[a];
	if (!stepped[sc]) return;
	const int end = sbeg[sc + 1];
	double A[6];
	for (int k = 0; k < 6; ++k) A[k] = aabb[6 * (size_t)a + k];
	const int sa = stat[a];
	const unsigned long long below = (1ull << lane) - 1;
	int o = off[a];
	if (o >= cap) return;  // beyond the pair capacity: the host grows it and detects again
	for (int base = a + 1; base < end; base += 64) {
		const int b = base + lane;
		const bool hit = bp_candidate(aabb, stat, A, sa, b, end);
		const unsigned long long m = __ballot(hit);
		if (hit) {
			const int k = o + __popcll(m & below);
			if (k < cap) {
				pa[k] = a;
				pb[k] = b;
			}
		}
		o += __popcll(m);
	}
}

// ---------------------------------------------------------------------------
// Sample pass of a pair (q's samples against g's grid).  The grid box of g
// seen from q's body frame: centre c, axes A (A[r*3+k] = (Rq^T Rg)_rk, g's
// axis k in q's frame), half extents h, and its q-frame box H.  A cell of q's
// samples is kept unless one of the three axes of q or of g separates its box
// from g's grid box by more than a margin far above the rounding of the exact
// per-sample test, so the cull never drops a sample that test would take.
// ---------------------------------------------------------------------------
struct PassGeom {
	double c[3], A[9], h[3], H[3], margin;
};

__device__ __forceinline__ void pass_geom(const double* Rq, const double* xq, const double* Rg, const double* xg,
                                          const GridDesc& G, PassGeom& P) {
	const double d[3] = {xg[0] - xq[0], xg[1] - xq[1], xg[2] - xq[2]};
	for (int r = 0; r < 3; ++r) {
		P.c[r] = Rq[0 * 3 + r] * d[0] + Rq[1 * 3 + r] * d[1] + Rq[2 * 3 + r] * d[2];
		for (int k = 0; k < 3; ++k)
			P.A[r * 3 + k] = Rq[0 * 3 + r] * Rg[0 * 3 + k] + Rq[1 * 3 + r] * Rg[1 * 3 + k] +
			                 Rq[2 * 3 + r] * Rg[2 * 3 + k];
	}
	for (int k = 0; k < 3; ++k) P.h[k] = G.half[k];
	for (int r = 0; r < 3; ++r)
		P.H[r] = fabs(P.A[r * 3 + 0]) * P.h[0] + fabs(P.A[r * 3 + 1]) * P.h[1] + fabs(P.A[r * 3 + 2]) * P.h[2];
	P.margin = 1e-6 * (1.0 + P.h[0] + P.h[1] + P.h[2] + fabs(P.c[0]) + fabs(P.c[1]) + fabs(P.c[2]));
}

__device__ __forceinline__ bool cell_overlaps(const PassGeom& P, const SampleCell& cell) {
	double m[3], e[3], d[3];
	for (int a = 0; a < 3; ++a) {
		m[a] = 0.5 * ((double)cell.lo[a] + (double)cell.hi[a]);
		e[a] = 0.5 * ((double)cell.hi[a] - (double)cell.lo[a]);
		d[a] = P.c[a] - m[a];
	}
	for (int r = 0; r < 3; ++r)
		if (fabs(d[r]) > e[r] + P.H[r] + P.margin) return false;
	for (int k = 0; k < 3; ++k) {
		const double pr = P.A[0 * 3 + k] * d[0] + P.A[1 * 3 + k] * d[1] + P.A[2 * 3 + k] * d[2];
		const double rad = P.h[k] + fabs(P.A[0 * 3 + k]) * e[0] + fabs(P.A[1 * 3 + k]) * e[1] +
		                   fabs(P.A[2 * 3 + k]) * e[2];
		if (fabs(pr) > rad + P.margin) return false;
	}
	return true;
}

// Grid nodes of g under the box of q's grid (the partner's box) seen in g's
// frame: the nodes a brick staged per pair and side would hold.
__device__ __forceinline__ double brick_nodes(const PassGeom& P, const GridDesc& Gq, const GridDesc& Gg) {
	double n = 1.0;
	for (int k = 0; k < 3; ++k) {
		// centre of q in g's frame is -A^T c; half extent sum_r |A_rk| hq_r
		const double ck = -(P.A[0 * 3 + k] * P.c[0] + P.A[1 * 3 + k] * P.c[1] + P.A[2 * 3 + k] * P.c[2]);
		const double hk = fabs(P.A[0 * 3 + k]) * Gq.half[0] + fabs(P.A[1 * 3 + k]) * Gq.half[1] +
		                  fabs(P.A[2 * 3 + k]) * Gq.half[2];
		const double lo = fmax(ck - hk, Gg.lo[k]), hi = fmin(ck + hk, -Gg.lo[k]);
		if (!(hi >= lo)) return 0.0;
		const double i0 = floor((lo - Gg.lo[k]) / Gg.sp[k]), i1 = ceil((hi - Gg.lo[k]) / Gg.sp[k]);
		n *= fmin(i1, (double)(Gg.n - 1)) - fmax(i0, 0.0) + 1.0;
	}
	return n;
}

struct PairArgs {
	const int* pa;
	const int* pb;
	const int* np;         // candidate pairs (may exceed cap)
	int cap;               // pair capacity
	const int* soff;       // per body sample offsets
	const int* ncell;      // per body cells
	const SampleCell* cells;
	const int* emit;
	const float* samples;
	const double* R;
	const double* pos;
	const GridDesc* grids;
	const float* pool;
	int* stage_cnt;        // per pair: samples the cull keeps (both sides)
	const int* stage_off;  // per pair (P+1)
	int64_t stage_cap;
	double* stage_raw;     // 7 doubles per slot
	double* stage_m;
	int* stage_idx;        // 2 ints per slot
	int* big;              // pairs with more than kSortCap hits
	int* pair_nh;
	int* pair_nc;
	int* pair_st;          // per pair kPairStats counters (summed by k_contact_totals)
	Scalars* sc;
	double tie_band;
};

// Per-pair detection counters: samples of the emitting sides, samples the
// cull keeps, grid nodes under the partners' boxes, samples inside the
// partner's grid box, tie-band samples.  Written per pair and summed by one
// workgroup (same-address atomics from every pair serialise across the XCDs).
enum { kStAll = 0, kStRead = 1, kStNodes = 2, kStInGrid = 3, kStTies = 4, kPairStats = 5 };

__device__ __forceinline__ int pairs_of(const PairArgs& A) { return min(*A.np, A.cap); }

// Culled sample count of one pair, one wavefront per pair: for each emitting
// side the samples of the cells kept for the partner (all of a body without
// cells); the sample and brick totals of the pair for the statistics.
__global__ __launch_bounds__(64 * kBpWaves) void k_pair_cull(PairArgs A) {
	const int lane = threadIdx.x & 63;
	const int P = pairs_of(A);
	for (int p = blockIdx.x * kBpWaves + (int)(threadIdx.x >> 6); p < P; p += gridDim.x * kBpWaves) {
		const int a = A.pa[p], b = A.pb[p];
		double Ra[9], Rb[9], xa[3], xb[3];
		for (int k = 0; k < 9; ++k) { Ra[k] = A.R[9 * a + k]; Rb[k] = A.R[9 * b + k]; }
		for (int k = 0; k < 3; ++k) { xa[k] = A.pos[3 * a + k]; xb[k] = A.pos[3 * b + k]; }
		int kept = 0, all = 0;
		double nodes = 0.0;
		for (int pass = 0; pass < 2; ++pass) {
			const int q = pass == 0 ? a : b, g = pass == 0 ? b : a;
			if (!A.emit[q]) continue;
			const int S = A.soff[q + 1] - A.soff[q];
			all += S;
			PassGeom G;
			if (pass == 0) pass_geom(Ra, xa, Rb, xb, A.grids[g], G);
			else pass_geom(Rb, xb, Ra, xa, A.grids[g], G);
			nodes += brick_nodes(G, A.grids[q], A.grids[g]);
			const int nc = A.ncell[q];
			if (nc == 0) {
				kept += S;
				continue;
			}
			int c = 0;
			if (lane < nc) {
				const SampleCell cl = A.cells[A.soff[q] / kCellStride + lane];
				if (cell_overlaps(G, cl)) c = cl.end - cl.beg;
			}
			for (int off = 32; off > 0; off >>= 1) c += __shfl_xor(c, off, 64);
			kept += c;
		}
		if (lane == 0) {
			A.stage_cnt[p] = kept;
			int* st = &A.pair_st[kPairStats * (size_t)p];
			st[kStAll] = all;
			st[kStRead] = kept;
			st[kStNodes] = (int)fmin(nodes, 2147483647.0);
		}
	}
}

// Tests sample s (FP32, body frame of q) against the grid of g; on a hit
// writes position, normal and depth.
__device__ __forceinline__ bool sample_hit(const PairArgs& A, const double* Rq, const double* xq, const double* Rg,
                                           const double* xg, const GridDesc& G, int s, double sign, double* out7,
                                           int* tie, int* ingrid) {
	const float* l = &A.samples[3 * (size_t)s];
	const double l0 = (double)l[0], l1 = (double)l[1], l2 = (double)l[2];
	double p[3];
	for (int r = 0; r < 3; ++r) p[r] = Rq[r * 3 + 0] * l0 + Rq[r * 3 + 1] * l1 + Rq[r * 3 + 2] * l2 + xq[r];
	const double d[3] = {p[0] - xg[0], p[1] - xg[1], p[2] - xg[2]};
	double q[3];
	for (int c = 0; c < 3; ++c) q[c] = Rg[0 * 3 + c] * d[0] + Rg[1 * 3 + c] * d[1] + Rg[2 * 3 + c] * d[2];
	const float* pool = A.pool + G.offset;
	const int n = G.n;
	auto fetch = [&](int i, int j, int k) -> double { return (double)pool[grid_index(n, i, j, k)]; };
	double sd, g[3];
	if (!grid_query(G, q, fetch, &sd, g)) return false;
	*ingrid += 1;
	if (fabs(sd) <= A.tie_band) *tie += 1;
	if (!(sd < 0.0)) return false;
	double gw[3];
	for (int r = 0; r < 3; ++r) gw[r] = Rg[r * 3 + 0] * g[0] + Rg[r * 3 + 1] * g[1] + Rg[r * 3 + 2] * g[2];
	const double gg = gw[0] * gw[0] + gw[1] * gw[1] + gw[2] * gw[2];
	if (!(gg > 0.0)) return false;
	const double gl = sqrt(gg);
	out7[0] = p[0]; out7[1] = p[1]; out7[2] = p[2];
	out7[3] = sign * (gw[0] / gl);
	out7[4] = sign * (gw[1] / gl);
	out7[5] = sign * (gw[2] / gl);
	out7[6] = -sd;
	return true;
}

__device__ __forceinline__ bool pos_less(double ax, double ay, double az, double bx, double by,
                                         double bz) {
	if (ax != bx) return ax < bx;
	if (ay != by) return ay < by;
	return az < bz;
}

// Sample pass: one 256-thread workgroup per candidate pair.  Per side, wave 0
// culls the cells and lists the kept ones with their sample prefix; the
// workgroup then tests the kept samples 256 at a time, compacting hits in
// (side, cell, sample) order into the pair's staging slots.
__global__ __launch_bounds__(256) void k_pair_hits(PairArgs A) {
	__shared__ int wcnt[4];
	__shared__ int s_nh, s_ties, s_in;
	__shared__ int sel_beg[64], pre[65];
	__shared__ int s_nsel;
	const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
	const int P = pairs_of(A);
	for (int p = blockIdx.x; p < P; p += gridDim.x) {
		const int a = A.pa[p], b = A.pb[p];
		const int base = A.stage_off[p];
		if ((int64_t)A.stage_off[p + 1] > A.stage_cap) {
			// staging too small: the host grows it and detects again
			if (tid == 0) {
				A.pair_nh[p] = 0;
				A.pair_st[kPairStats * (size_t)p + kStInGrid] = 0;
				A.pair_st[kPairStats * (size_t)p + kStTies] = 0;
			}
			continue;
		}
		double Ra[9], Rb[9], xa[3], xb[3];
		for (int k = 0; k < 9; ++k) { Ra[k] = A.R[9 * a + k]; Rb[k] = A.R[9 * b + k]; }
		for (int k = 0; k < 3; ++k) { xa[k] = A.pos[3 * a + k]; xb[k] = A.pos[3 * b + k]; }
		__syncthreads();
		if (tid == 0) { s_nh = 0; s_ties = 0; s_in = 0; }
		int myties = 0, myin = 0;
		for (int pass = 0; pass < 2; ++pass) {
			const int q = pass == 0 ? a : b, g = pass == 0 ? b : a;
			if (!A.emit[q]) continue;
			const GridDesc Gg = A.grids[g];
			const double* Rq = pass == 0 ? Ra : Rb;
			const double* xq = pass == 0 ? xa : xb;
			const double* Rg = pass == 0 ? Rb : Ra;
			const double* xg = pass == 0 ? xb : xa;
			const int s0 = A.soff[q];
			const int nc = A.ncell[q];
			if (wid == 0) {
				if (nc == 0) {
					if (lane == 0) {
						sel_beg[0] = 0;
						pre[0] = 0;
						pre[1] = A.soff[q + 1] - s0;
						s_nsel = 1;
					}
				} else {
					PassGeom G;
					pass_geom(Rq, xq, Rg, xg, Gg, G);
					bool keep = false;
					SampleCell cl{};
					if (lane < nc) {
						cl = A.cells[s0 / kCellStride + lane];
						keep = cell_overlaps(G, cl);
					}
					const unsigned long long m = __ballot(keep);
					const int cnt = keep ? cl.end - cl.beg : 0;
					int x = cnt;
					for (int off = 1; off < 64; off <<= 1) {
						const int y = __shfl_up(x, off, 64);
						if (lane >= off) x += y;
					}
					if (keep) {
						const int r = __popcll(m & ((1ull << lane) - 1));
						sel_beg[r] = cl.beg;
						pre[r + 1] = x;
					}
					if (lane == 0) {
						pre[0] = 0;
						s_nsel = __popcll(m);
					}
				}
			}
			__syncthreads();
			const int nsel = s_nsel;
			const int T = pre[nsel];
			const double sign = pass == 0 ? -1.0 : 1.0;
			for (int c0 = 0; c0 < T; c0 += 256) {
				const int t = c0 + tid;
				double h7[7];
				bool hit = false;
				if (t < T) {
					int lo = 0, hi = nsel - 1;  // the cell j with pre[j] <= t < pre[j + 1]
					while (lo < hi) {
						const int mid = (lo + hi + 1) >> 1;
						if (pre[mid] <= t) lo = mid;
						else hi = mid - 1;
					}
					const int smp = s0 + sel_beg[lo] + (t - pre[lo]);
					hit = sample_hit(A, Rq, xq, Rg, xg, Gg, smp, sign, h7, &myties, &myin);
				}
				const unsigned long long m = __ballot(hit);
				const unsigned long long lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
				const int rank = __popcll(m & lt);
				if (lane == 0) wcnt[wid] = __popcll(m);
				__syncthreads();
				int before = 0;
				for (int w = 0; w < wid; ++w) before += wcnt[w];
				const int total = wcnt[0] + wcnt[1] + wcnt[2] + wcnt[3];
				const int nh0 = s_nh;
				if (hit) {
					double* o = &A.stage_raw[7 * (size_t)(base + nh0 + before + rank)];
					for (int k = 0; k < 7; ++k) o[k] = h7[k];
				}
				__syncthreads();
				if (tid == 0) s_nh = nh0 + total;
				__syncthreads();
			}
			__syncthreads();
		}
		if (myties) atomicAdd(&s_ties, myties);
		if (myin) atomicAdd(&s_in, myin);
		__syncthreads();
		if (tid == 0) {
			const int nh = s_nh;
			A.pair_nh[p] = nh;
			A.pair_st[kPairStats * (size_t)p + kStInGrid] = s_in;
			A.pair_st[kPairStats * (size_t)p + kStTies] = s_ties;
			if (nh > kSortCap) A.big[atomicAdd(&A.sc->n_big, 1)] = p;
		}
	}
}

// Merged record of the sorted hits [i, j) of a pair (`at(k)`: the staging
// slot of the k-th sorted hit): the first position, the normalised sum of the
// members' frame normals, the minimum depth (DynamicSystem.cpp:364-381).
template <class At>
__device__ __forceinline__ void merge_group(const double* __restrict__ raw, int i, int j, At at, double* out) {
	const double* o = &raw[7 * (size_t)at(i)];
	for (int k = 0; k < 7; ++k) out[k] = o[k];
	if (j - i == 1) return;
	double Rr[9];
	frame_rotation(&o[3], Rr);
	double n[3] = {Rr[0], Rr[3], Rr[6]};
	double dep = o[6];
	for (int k = i + 1; k < j; ++k) {
		const double* ok = &raw[7 * (size_t)at(k)];
		frame_rotation(&ok[3], Rr);
		n[0] += Rr[0]; n[1] += Rr[3]; n[2] += Rr[6];
		dep = ok[6] < dep ? ok[6] : dep;
	}
	const double nn = sqrt(dot3(n, n));
	if (nn > 0) { n[0] /= nn; n[1] /= nn; n[2] /= nn; }
	out[3] = n[0]; out[4] = n[1]; out[5] = n[2];
	out[6] = dep;
}

// A new group starts at sorted hit i where |p_i - p_{i-1}|^2 > 1e-8
// (DynamicSystem.cpp:356-363).
__device__ __forceinline__ bool group_break(double x0, double y0, double z0, double x1, double y1, double z1) {
	const double d0 = x1 - x0, d1 = y1 - y0, d2 = z1 - z0;
	return d0 * d0 + d1 * d1 + d2 * d2 > 1e-8;
}

// Merge pass: one 256-thread workgroup per pair with 2..kSortCap hits; sorts
// the hits by position (bitonic in LDS) and chain-merges them.
__global__ __launch_bounds__(256) void k_pair_merge(PairArgs A) {
	__shared__ double kx[kSortCap], ky[kSortCap], kz[kSortCap];
	__shared__ short kidx[kSortCap];
	__shared__ unsigned char brk[kSortCap];
	__shared__ int gidx[kSortCap];
	__shared__ int scan_part[256];
	const int tid = threadIdx.x;
	const int P = pairs_of(A);
	for (int p = blockIdx.x; p < P; p += gridDim.x) {
		const int nh = A.pair_nh[p];
		if (nh > kSortCap) continue;  // the global-memory merge takes it
		if (nh == 0) {
			if (tid == 0) A.pair_nc[p] = 0;
			continue;
		}
		const int base = A.stage_off[p];
		if (nh == 1) {
			if (tid < 7) A.stage_m[7 * (size_t)base + tid] = A.stage_raw[7 * (size_t)base + tid];
			if (tid == 0) A.pair_nc[p] = 1;
			continue;
		}
		__syncthreads();
		int M = 1;
		while (M < nh) M <<= 1;
		for (int i = tid; i < M; i += 256) {
			if (i < nh) {
				const double* o = &A.stage_raw[7 * (size_t)(base + i)];
				kx[i] = o[0]; ky[i] = o[1]; kz[i] = o[2];
			} else {
				kx[i] = INFINITY; ky[i] = INFINITY; kz[i] = INFINITY;
			}
			kidx[i] = (short)i;
		}
		__syncthreads();
		for (int k = 2; k <= M; k <<= 1) {
			for (int j = k >> 1; j > 0; j >>= 1) {
				for (int i = tid; i < M; i += 256) {
					const int l = i ^ j;
					if (l > i) {
						const bool up = (i & k) == 0;
						const bool sw = up ? pos_less(kx[l], ky[l], kz[l], kx[i], ky[i], kz[i])
						                   : pos_less(kx[i], ky[i], kz[i], kx[l], ky[l], kz[l]);
						if (sw) {
							double t;
							t = kx[i]; kx[i] = kx[l]; kx[l] = t;
							t = ky[i]; ky[i] = ky[l]; ky[l] = t;
							t = kz[i]; kz[i] = kz[l]; kz[l] = t;
							const short ti = kidx[i]; kidx[i] = kidx[l]; kidx[l] = ti;
						}
					}
				}
				__syncthreads();
			}
		}
		for (int i = tid; i < nh; i += 256)
			brk[i] = (i == 0 || group_break(kx[i - 1], ky[i - 1], kz[i - 1], kx[i], ky[i], kz[i])) ? 1 : 0;
		__syncthreads();
		{
			int loc = 0;
			const int b0 = tid * 8;
			for (int k = 0; k < 8; ++k)
				if (b0 + k < nh) loc += brk[b0 + k];
			scan_part[tid] = loc;
			__syncthreads();
			if (tid == 0) {
				int run = 0;
				for (int t = 0; t < 256; ++t) { const int v = scan_part[t]; scan_part[t] = run; run += v; }
			}
			__syncthreads();
			int run = scan_part[tid];
			for (int k = 0; k < 8; ++k)
				if (b0 + k < nh) { gidx[b0 + k] = run; run += brk[b0 + k]; }
		}
		__syncthreads();
		const double* raw = A.stage_raw;
		auto at = [&](int k) { return base + (int)kidx[k]; };
		for (int i = tid; i < nh; i += 256) {
			if (!brk[i]) continue;
			int j = i + 1;
			while (j < nh && !brk[j]) ++j;
			double out[7];
			merge_group(raw, i, j, at, out);
			double* w = &A.stage_m[7 * (size_t)(base + gidx[i])];
			for (int k = 0; k < 7; ++k) w[k] = out[k];
		}
		if (tid == 0) A.pair_nc[p] = gidx[nh - 1] + brk[nh - 1];
		__syncthreads();
	}
}

// Merge of a pair with more than kSortCap hits, one 256-thread workgroup per
// such pair: a stable bottom-up merge sort of the hit indices by position in
// global memory (the pair's two index arrays in stage_idx), then the same
// chain merge.  For distinct positions the order equals the bitonic sort's.
__device__ __forceinline__ bool hit_less(const double* __restrict__ raw, int base, int u, int v) {
	const double* a = &raw[7 * (size_t)(base + u)];
	const double* b = &raw[7 * (size_t)(base + v)];
	return pos_less(a[0], a[1], a[2], b[0], b[1], b[2]);
}

__global__ __launch_bounds__(256) void k_pair_merge_big(PairArgs A) {
	__shared__ int part[256];
	__shared__ int carry;
	const int tid = threadIdx.x;
	const int nbig = A.sc->n_big;
	const double* raw = A.stage_raw;
	for (int bi = blockIdx.x; bi < nbig; bi += gridDim.x) {
		const int p = A.big[bi];
		const int nh = A.pair_nh[p];
		const int base = A.stage_off[p];
		int* I0 = A.stage_idx + 2 * (size_t)base;
		int* I1 = I0 + nh;
		for (int i = tid; i < nh; i += 256) I0[i] = i;
		__syncthreads();
		int* src = I0;
		int* dst = I1;
		for (int w = 1; w < nh; w <<= 1) {
			for (int e = tid; e < nh; e += 256) {
				const int L = (e / (2 * w)) * (2 * w);
				const int mid = min(L + w, nh), R = min(L + 2 * w, nh);
				const int v = src[e];
				int lo, hi;
				if (e < mid) {  // left run: place after the right elements strictly less
					lo = mid; hi = R;
					while (lo < hi) {
						const int m = (lo + hi) >> 1;
						if (hit_less(raw, base, src[m], v)) lo = m + 1;
						else hi = m;
					}
					dst[e + (lo - mid)] = v;
				} else {        // right run: place after the left elements not greater
					lo = L; hi = mid;
					while (lo < hi) {
						const int m = (lo + hi) >> 1;
						if (!hit_less(raw, base, v, src[m])) lo = m + 1;
						else hi = m;
					}
					dst[L + (e - mid) + (lo - L)] = v;
				}
			}
			__syncthreads();
			int* t = src; src = dst; dst = t;
		}
		// src: sorted indices; dst (free): group number of each group start
		auto key = [&](int i, double* k3) {
			const double* o = &raw[7 * (size_t)(base + src[i])];
			k3[0] = o[0]; k3[1] = o[1]; k3[2] = o[2];
		};
		auto is_break = [&](int i) {
			if (i == 0) return true;
			double u[3], v[3];
			key(i - 1, u);
			key(i, v);
			return group_break(u[0], u[1], u[2], v[0], v[1], v[2]);
		};
		if (tid == 0) carry = 0;
		__syncthreads();
		for (int c0 = 0; c0 < nh; c0 += 256) {
			const int i = c0 + tid;
			const int f = (i < nh && is_break(i)) ? 1 : 0;
			part[tid] = f;
			__syncthreads();
			for (int off = 1; off < 256; off <<= 1) {
				const int y = tid >= off ? part[tid - off] : 0;
				__syncthreads();
				part[tid] += y;
				__syncthreads();
			}
			if (f) dst[i] = carry + part[tid] - 1;
			__syncthreads();
			if (tid == 0) carry += part[255];
			__syncthreads();
		}
		auto at = [&](int k) { return base + src[k]; };
		for (int i = tid; i < nh; i += 256) {
			if (!is_break(i)) continue;
			int j = i + 1;
			while (j < nh && !is_break(j)) ++j;
			double out[7];
			merge_group(raw, i, j, at, out);
			double* w = &A.stage_m[7 * (size_t)(base + dst[i])];
			for (int k = 0; k < 7; ++k) w[k] = out[k];
		}
		if (tid == 0) A.pair_nc[p] = carry;
		__syncthreads();
	}
}

// Totals of the detection for the host's single read-back.
// One workgroup: the detection totals (pair, staging and contact counts; the
// per-pair counters and raw hits summed).
__global__ __launch_bounds__(1024) void k_contact_totals(PairArgs A, const int* __restrict__ coff) {
	__shared__ unsigned long long part[16][kPairStats + 1];
	const int P = pairs_of(A);
	unsigned long long t[kPairStats + 1] = {};
	for (int p = threadIdx.x; p < P; p += blockDim.x) {
		for (int k = 0; k < kPairStats; ++k) t[k] += (unsigned long long)A.pair_st[kPairStats * (size_t)p + k];
		t[kPairStats] += (unsigned long long)A.pair_nh[p];
	}
	for (int k = 0; k <= kPairStats; ++k)
		for (int off = 32; off > 0; off >>= 1) t[k] += __shfl_xor(t[k], off, 64);
	const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
	if (lane == 0)
		for (int k = 0; k <= kPairStats; ++k) part[wid][k] = t[k];
	__syncthreads();
	if (threadIdx.x == 0) {
		unsigned long long u[kPairStats + 1] = {};
		for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
			for (int k = 0; k <= kPairStats; ++k) u[k] += part[w][k];
		A.sc->n_pairs = *A.np;
		A.sc->n_stage = A.stage_off[P];
		A.sc->n_contacts = coff[P];
		A.sc->all_samples = u[kStAll];
		A.sc->read_samples = u[kStRead];
		A.sc->brick_nodes = u[kStNodes];
		A.sc->ingrid_samples = u[kStInGrid];
		A.sc->ties = (int)u[kStTies];
		A.sc->n_raw = (int)u[kPairStats];
	}
}

// One workgroup per pair: final contact records in pair order.
__global__ __launch_bounds__(256) void k_pair_write(const int* __restrict__ pa, const int* __restrict__ pb,
                             const int* __restrict__ stage_off, const int* __restrict__ pair_nc,
                             const int* __restrict__ coff, const double* __restrict__ stage_m,
                             const double* __restrict__ pos, int C, int* __restrict__ id0,
                             int* __restrict__ id1, double* __restrict__ cpos,
                             double* __restrict__ cdepth, double* __restrict__ cT,
                             double* __restrict__ clf) {
	const int p = blockIdx.x;
	const int n = pair_nc[p];
	if (n == 0) return;
	const int a = pa[p], b = pb[p];
	const double xa[3] = {pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]};
	const double xb[3] = {pos[3 * b], pos[3 * b + 1], pos[3 * b + 2]};
	const int base = stage_off[p];
	const int o0 = coff[p];
	for (int g = threadIdx.x; g < n; g += blockDim.x) {
		const double* s = &stage_m[7 * (size_t)(base + g)];
		const int i = o0 + g;
		double R[9], T[18];
		frame_rotation(&s[3], R);
		id0[i] = a;
		id1[i] = b;
		for (int k = 0; k < 3; ++k) cpos[(size_t)k * C + i] = s[k];
		cdepth[i] = fabs(s[6]);
		frame_T(R, s, xa, 0, T);
		for (int k = 0; k < 18; ++k) cT[36 * (size_t)i + k] = T[k];
		frame_T(R, s, xb, 1, T);
		for (int k = 0; k < 18; ++k) cT[36 * (size_t)i + 18 + k] = T[k];
		for (int k = 0; k < 3; ++k) clf[(size_t)k * C + i] = 0.0;
	}
}

// Contact detection with one read-back: the pair list and the staging slots
// live in capacities kept across steps; a kernel that finds its capacity too
// small leaves its part undone, and the host, seeing the totals, grows the
// capacity and detects again (on a growing scene only).
void detect_contacts(pdc_context* ctx) {
	hipStream_t s = ctx->stream;
	const int N = ctx->N;
	ctx->C = 0;
	ctx->det = DetectionStats{};
	ctx->n_raw = 0;
	ctx->n_ties = 0;
	PDC_HIP(hipMemsetAsync(ctx->d_sc.p, 0, sizeof(Scalars), s));
	if (N < 2) {
		ctx->n_pairs = 0;
		return;
	}
	hipLaunchKernelGGL(k_body_frames, dim3(grid_for(N)), dim3(kBlock), 0, s, N, ctx->b_quat.p,
	                   ctx->b_pos.p, ctx->b_grid.p, ctx->b_R.p, ctx->b_aabb.p);
	PDC_CHECK_LAUNCH();
	ctx->bp_cnt.ensure(N + 1);
	ctx->bp_off.ensure(N + 1);
	const int bp_grid = (N + kBpWaves - 1) / kBpWaves;
	hipLaunchKernelGGL(k_bp_count, dim3(bp_grid), dim3(64 * kBpWaves), 0, s, N, ctx->b_aabb.p,
	                   ctx->b_static.p, ctx->b_scene.p, ctx->d_sbeg.p, ctx->d_step.p, ctx->bp_cnt.p);
	PDC_CHECK_LAUNCH();
	scan_exclusive(ctx, ctx->bp_cnt.p, ctx->bp_off.p, N);
	if (ctx->pr_cap < 1) ctx->pr_cap = std::max(1024, 4 * N);
	if (ctx->stage_cap < 1) ctx->stage_cap = std::max<int64_t>(1 << 16, 64 * (int64_t)N);
	PairArgs A{};
	for (int attempt = 0;; ++attempt) {
		const int cap = ctx->pr_cap;
		ctx->pr_a.ensure(cap); ctx->pr_b.ensure(cap); ctx->pr_tmp.ensure(cap + 1);
		ctx->pr_soff.ensure(cap + 1); ctx->pr_nc.ensure(cap + 1); ctx->pr_coff.ensure(cap + 1);
		ctx->pr_nh.ensure(cap + 1); ctx->pr_big.ensure(cap); ctx->pr_st.ensure(kPairStats * (size_t)cap);
		ctx->stage_raw.ensure(7 * (size_t)ctx->stage_cap);
		ctx->stage_m.ensure(7 * (size_t)ctx->stage_cap);
		ctx->stage_idx.ensure(2 * (size_t)ctx->stage_cap);
		if (attempt > 0) PDC_HIP(hipMemsetAsync(ctx->d_sc.p, 0, sizeof(Scalars), s));
		hipLaunchKernelGGL(k_bp_write, dim3(bp_grid), dim3(64 * kBpWaves), 0, s, N, ctx->b_aabb.p,
		                   ctx->b_static.p, ctx->b_scene.p, ctx->d_sbeg.p, ctx->d_step.p, ctx->bp_off.p, cap,
		                   ctx->pr_a.p, ctx->pr_b.p);
		PDC_CHECK_LAUNCH();
		A.pa = ctx->pr_a.p; A.pb = ctx->pr_b.p; A.np = ctx->bp_off.p + N; A.cap = cap;
		A.soff = ctx->b_soff.p; A.ncell = ctx->b_ncell.p; A.cells = ctx->cells.p; A.emit = ctx->b_emit.p;
		A.samples = ctx->samples.p; A.R = ctx->b_R.p; A.pos = ctx->b_pos.p; A.grids = ctx->b_grid.p;
		A.pool = ctx->grid_pool.p; A.stage_cnt = ctx->pr_tmp.p; A.stage_off = ctx->pr_soff.p;
		A.stage_cap = ctx->stage_cap; A.stage_raw = ctx->stage_raw.p; A.stage_m = ctx->stage_m.p;
		A.stage_idx = ctx->stage_idx.p; A.big = ctx->pr_big.p; A.pair_nh = ctx->pr_nh.p;
		A.pair_nc = ctx->pr_nc.p; A.pair_st = ctx->pr_st.p; A.sc = ctx->d_sc.p; A.tie_band = ctx->tie_band;
		// launch width from the last detection; the kernels stride over the
		// pairs there are, so the width only sets the parallelism
		const int hint = std::min(cap, std::max(256, ctx->n_pairs + ctx->n_pairs / 8 + 64));
		hipEvent_t e0;
		kt_begin(ctx, PDC_KT_CONTACTS, &e0);
		hipLaunchKernelGGL(k_pair_cull, dim3((hint + kBpWaves - 1) / kBpWaves), dim3(64 * kBpWaves), 0, s, A);
		PDC_CHECK_LAUNCH();
		scan_exclusive_dev(ctx, ctx->pr_tmp.p, ctx->pr_soff.p, cap, A.np);
		hipLaunchKernelGGL(k_pair_hits, dim3(hint), dim3(256), 0, s, A);
		PDC_CHECK_LAUNCH();
		kt_end(ctx, PDC_KT_CONTACTS, e0);
		if (e0) ctx->kt.pending_contact_slots = 1;
		hipLaunchKernelGGL(k_pair_merge, dim3(hint), dim3(256), 0, s, A);
		PDC_CHECK_LAUNCH();
		hipLaunchKernelGGL(k_pair_merge_big, dim3(64), dim3(256), 0, s, A);
		PDC_CHECK_LAUNCH();
		scan_exclusive_dev(ctx, ctx->pr_nc.p, ctx->pr_coff.p, cap, A.np);
		hipLaunchKernelGGL(k_contact_totals, dim3(1), dim3(1024), 0, s, A, (const int*)ctx->pr_coff.p);
		PDC_CHECK_LAUNCH();
		sync_scalars(ctx);
		const Scalars& h = *ctx->h_sc;
		bool again = false;
		if (h.n_pairs > ctx->pr_cap) {
			ctx->pr_cap = h.n_pairs + h.n_pairs / 4 + 64;
			again = true;
		} else if ((int64_t)h.n_stage > ctx->stage_cap) {
			ctx->stage_cap = (int64_t)h.n_stage + h.n_stage / 4 + 1024;
			again = true;
		}
		if (!again) break;
		ctx->kt.pending_contact_slots = 0;
		if (attempt >= 3) throw HipError("contact detection capacities did not settle", PDC_ERR_CAPACITY);
	}
	const Scalars& h = *ctx->h_sc;
	const int P = h.n_pairs;
	const int C = h.n_contacts;
	ctx->n_pairs = P;
	ctx->n_raw = h.n_raw;
	ctx->n_ties = h.ties;
	ctx->det.pairs = P;
	ctx->det.raw_hits = h.n_raw;
	ctx->det.ties = h.ties;
	ctx->det.samples_all = (int64_t)h.all_samples;
	ctx->det.samples_in_grid = (int64_t)h.ingrid_samples;
	ctx->det.samples_read = (int64_t)h.read_samples;
	ctx->det.brick_nodes = (int64_t)h.brick_nodes;
	ctx->det.large_pairs = h.n_big;
	if (ctx->kt.pending_contact_slots) {
		// algorithmic bytes of the cull and the sample pass: per pair the two
		// bodies' rotations, positions and grid descriptors (368 B); per kept
		// sample its FP32 body-frame position (12 B); per sample inside the
		// partner's grid box the 8 FP32 nodes it interpolates (32 B); per raw
		// hit its staged record (56 B)
		ctx->kt.bytes += 368.0 * (double)h.n_pairs + 12.0 * (double)h.read_samples +
		                 32.0 * (double)h.ingrid_samples + 56.0 * (double)h.n_raw;
		ctx->kt.pending_contact_slots = 0;
	}
	if (C == 0 || P == 0) return;
	ensure_contact_capacity(ctx, C);
	ctx->C = C;
	hipLaunchKernelGGL(k_pair_write, dim3(P), dim3(64), 0, s, ctx->pr_a.p, ctx->pr_b.p,
	                   ctx->pr_soff.p, ctx->pr_nc.p, ctx->pr_coff.p, ctx->stage_m.p, ctx->b_pos.p,
	                   (int)ctx->c_stride, ctx->c_id0.p, ctx->c_id1.p, ctx->c_pos.p, ctx->c_depth.p,
	                   ctx->c_T.p, ctx->c_lf.p);
	PDC_CHECK_LAUNCH();
}

}  // namespace pdc
