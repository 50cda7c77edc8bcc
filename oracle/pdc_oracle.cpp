/*
 * pdc_oracle.cpp — ORACLE, TEST INFRASTRUCTURE ONLY.
 *
 * A CPU FP64 restatement of the reference's per-timestep contact pipeline.
 * It is the checker for the HIP product path and the timed CPU baseline of the
 * benchmark; the product library never links or calls it.  OpenMP loops use
 * gathers over per-body contact lists and fixed-block reductions, so results
 * do not depend on the thread count.
 *
 * Each function cites the reference code it restates (paths relative to the
 * reference repository):
 *   Lorentz-circle constraint     LorentzCircleConstraint.cpp:5-55 (+ .h:18-19)
 *   contact frame                 ContactGeneratorBase.h:21-49, convertContact :57-68
 *   J operators                   ContactInfo.h:46-110
 *   per-pair merge                DynamicSystem.cpp:341-381
 *   fillForces / renumbering      PrimalDualForceSolver.cpp:79-130
 *   characteristic mass / speed   PrimalDualForceSolver.cpp:23-78
 *   Newton loop                   PrimalDualForceSolver.cpp:167-561
 *   residuals                     PrimalDualForceSolver.cpp:132-165,
 *                                 NonSmoothContactForce.cpp:119-172, OnlyNormalForce.cpp:46-89
 *   tangential weight             NonSmoothContactForce.cpp:74-104
 *   Alart-Curnier vector          NonSmoothContactForce.cpp:53-73, OnlyNormalForce.cpp:202-217
 *   assembly                      PrimalDualForceSolver.cpp:322-377,
 *                                 NonSmoothContactForce.cpp:173-253, OnlyNormalForce.cpp:90-151
 *   increments                    NonSmoothContactForce.cpp:254-280, OnlyNormalForce.cpp:152-171
 *   line search + filter          PrimalDualForceSolver.cpp:562-635,
 *                                 NonSmoothContactForce.cpp:281-319, OnlyNormalForce.cpp:172-201
 *   GD fallback, Hu schedule      PrimalDualForceSolver.cpp:441-483
 *   write-back                    PrimalDualForceSolver.cpp:491-510
 *   integration                   DynamicSystem.cpp:137-161
 *   mass matrices                 RigidObjectBase.cpp:106-119
 *
 * The grid-SDF surface-sample contact generator has no counterpart in the
 * reference (which uses FCL and libigl); it restates this project's own
 * method, whose definition is written out in the functions below.
 *
 * Deviations from the reference, all deliberate:
 *   - the direct branch (N < 2000) is a block-sparse Cholesky with a greedy
 *     minimum-degree ordering in place of Eigen's SimplicialLDLT; the PCG
 *     branch is CG in place of AMGCL SA-AMG CG, preconditioned by the inverse
 *     diagonal blocks (block Jacobi) at pcg_tol >= 1e-9 and, below it, by the
 *     inverse blocks of groups of up to eight strongly coupled bodies
 *     (group_bodies); its start is the previous direction only when that has
 *     a smaller residual than zero;
 *   - contact-to-body scatters are gathers (no FP atomics), so sums have a
 *     fixed order.
 */
#include "orc.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <unordered_map>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

/* Below this PCG tolerance the device runs the standard PCG form, whose
 * preconditioner is the group-block inverse (group_bodies); at or above it
 * the single-reduction form with block Jacobi. */
constexpr double kStandardFormTol = 1e-9;

/* ------------------------------------------------------------------------ */
/* small vector / matrix helpers (row-major 3x3)                             */
/* ------------------------------------------------------------------------ */
inline double dot3(const double* a, const double* b) {
	return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
inline void cross3(const double* a, const double* b, double* o) {
	o[0] = a[1] * b[2] - a[2] * b[1];
	o[1] = a[2] * b[0] - a[0] * b[2];
	o[2] = a[0] * b[1] - a[1] * b[0];
}
/* Rotation of a unit quaternion stored (x, y, z, w). */
void quat_rot(const double* q, double* R) {
	const double x = q[0], y = q[1], z = q[2], w = q[3];
	const double tx = 2 * x, ty = 2 * y, tz = 2 * z;
	const double twx = tx * w, twy = ty * w, twz = tz * w;
	const double txx = tx * x, txy = ty * x, txz = tz * x;
	const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
	R[0] = 1 - (tyy + tzz); R[1] = txy - twz;       R[2] = txz + twy;
	R[3] = txy + twz;       R[4] = 1 - (txx + tzz); R[5] = tyz - twx;
	R[6] = txz - twy;       R[7] = tyz + twx;       R[8] = 1 - (txx + tyy);
}
/* Hamilton product o = a * b, quaternions stored (x, y, z, w). */
void quat_mul(const double* a, const double* b, double* o) {
	double r[4];
	r[3] = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
	r[0] = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
	r[1] = a[3] * b[1] + a[1] * b[3] + a[2] * b[0] - a[0] * b[2];
	r[2] = a[3] * b[2] + a[2] * b[3] + a[0] * b[1] - a[1] * b[0];
	std::memcpy(o, r, sizeof(r));
}
/* R * diag(d) * R^T */
void rot_diag_rot(const double* R, const double* d, double* o) {
	for (int r = 0; r < 3; ++r)
		for (int c = 0; c < 3; ++c)
			o[r * 3 + c] = R[r * 3 + 0] * d[0] * R[c * 3 + 0] +
			               R[r * 3 + 1] * d[1] * R[c * 3 + 1] +
			               R[r * 3 + 2] * d[2] * R[c * 3 + 2];
}

/* Sum of f(i), i < n, in fixed 1024-element blocks: the result does not
 * depend on the number of threads. */
template <class F>
double fixed_sum(size_t n, F f) {
	const size_t B = 1024;
	const size_t nb = (n + B - 1) / B;
	std::vector<double> part(nb, 0.0);
#pragma omp parallel for schedule(static)
	for (long b = 0; b < (long)nb; ++b) {
		double s = 0;
		const size_t e = std::min(n, (size_t)(b + 1) * B);
		for (size_t i = (size_t)b * B; i < e; ++i) s += f(i);
		part[b] = s;
	}
	double s = 0;
	for (double v : part) s += v;
	return s;
}

/* ------------------------------------------------------------------------ */
/* Lorentz-circle friction constraint (LorentzCircleConstraint.cpp:5-55)     */
/* ------------------------------------------------------------------------ */
struct Lorentz {
	double mu;
	static constexpr double P = 1e-16;  /* m_perturbation  (.h:18) */
	static constexpr double SH = 1e-8;  /* m_shift         (.h:19) */
	double constraint(const double* f) const {
		const double r = f[1] * f[1] + f[2] * f[2] + P;
		return std::sqrt(r) - mu * f[0] - std::sqrt(P) - SH;
	}
	void gradient(const double* f, double* g) const {
		const double r = std::sqrt(f[1] * f[1] + f[2] * f[2] + P);
		g[0] = f[1] / r;
		g[1] = f[2] / r;
	}
	/* H = (z z^T + p' I) (|f_t|^2 + p')^(-3/2), z = (f2, -f1), p' = max(P, 1e-6) */
	void hessian(const double* f, double* H) const {
		const double p = std::max(P, 1e-6);
		const double r2 = f[1] * f[1] + f[2] * f[2] + p;
		const double r3 = std::pow(r2, -1.5);
		const double z0 = f[2], z1 = -f[1];
		H[0] = (z0 * z0 + p) * r3;
		H[1] = (z0 * z1) * r3;
		H[2] = (z1 * z0) * r3;
		H[3] = (z1 * z1 + p) * r3;
	}
	double radius(double f0) const {
		const double k = mu * f0 + std::sqrt(P) + SH;
		return std::sqrt(k * k - P);
	}
	/* project(in, start = in, target = 0) (.cpp:36-49) */
	void project(const double* in, double* out) const {
		out[0] = in[0]; out[1] = in[1]; out[2] = in[2];
		if (in[0] < 0) out[0] = 0;
		if (constraint(out) > 0) {
			const double r = std::max(radius(in[0]) + 0.0, 0.0);
			const double nt = std::sqrt(out[1] * out[1] + out[2] * out[2]);
			const double s = r / nt;
			out[1] *= s;
			out[2] *= s;
		}
	}
};

/* ------------------------------------------------------------------------ */
/* contact frame (ContactGeneratorBase.h:21-49)                              */
/* T[j] is the 6x3 column-major matrix transformation[j]:                    */
/*   T[j][col*6 + row], rows 0-2 = rot_j, rows 3-5 = c_j x rot_j.col(col),   */
/*   rot_0 = AngleAxis(atan2(|x^n|, x.n), normalize(x^n)), rot_1 = -rot_0.   */
/* ------------------------------------------------------------------------ */
void normal_to_transform(const double* pos, const double* n, const double* x0,
                         const double* x1, double* T0, double* T1) {
	const double axis[3] = {0.0, -n[2], n[1]}; /* (1,0,0) x n */
	const double cs = n[0];
	const double an = std::sqrt(dot3(axis, axis));
	const double angle = std::atan2(an, cs);
	double a[3] = {axis[0], axis[1], axis[2]};
	const double sq = dot3(axis, axis);
	if (sq > 0) {
		const double inv = std::sqrt(sq);
		a[0] = axis[0] / inv; a[1] = axis[1] / inv; a[2] = axis[2] / inv;
	}
	/* Rodrigues rotation of angle about unit axis a */
	const double s = std::sin(angle), c = std::cos(angle);
	const double sa[3] = {s * a[0], s * a[1], s * a[2]};
	const double c1 = 1 - c;
	const double ca[3] = {c1 * a[0], c1 * a[1], c1 * a[2]};
	double R[9];
	R[0] = ca[0] * a[0] + c;
	R[4] = ca[1] * a[1] + c;
	R[8] = ca[2] * a[2] + c;
	double t = ca[0] * a[1];
	R[1] = t - sa[2]; R[3] = t + sa[2];
	t = ca[0] * a[2];
	R[2] = t + sa[1]; R[6] = t - sa[1];
	t = ca[1] * a[2];
	R[5] = t - sa[0]; R[7] = t + sa[0];
	double* Ts[2] = {T0, T1};
	const double* xs[2] = {x0, x1};
	for (int j = 0; j < 2; ++j) {
		const double cj[3] = {pos[0] - xs[j][0], pos[1] - xs[j][1], pos[2] - xs[j][2]};
		if (j)
			for (double& v : R) v = -v;
		for (int col = 0; col < 3; ++col) {
			const double rc[3] = {R[0 * 3 + col], R[1 * 3 + col], R[2 * 3 + col]};
			double cr[3];
			cross3(cj, rc, cr);
			for (int r = 0; r < 3; ++r) {
				Ts[j][col * 6 + r] = rc[r];
				Ts[j][col * 6 + 3 + r] = cr[r];
			}
		}
	}
}

/* ------------------------------------------------------------------------ */
/* bodies, grid SDF, samples                                                 */
/* ------------------------------------------------------------------------ */
struct Body {
	int type = PDC_SHAPE_BOX;
	double dims[3] = {0, 0, 0};
	double mass = 1, I0[3] = {1, 1, 1};
	bool stat = false, emit = true;
	double x[3] = {0, 0, 0}, q[4] = {0, 0, 0, 1}, v[3] = {0, 0, 0}, w[3] = {0, 0, 0};
	double acc[3] = {0, 0, 0}, aacc[3] = {0, 0, 0};
	/* grid */
	int gn = 2;
	double glo[3] = {0, 0, 0}, gsp[3] = {1, 1, 1}, ghalf[3] = {0, 0, 0};
	std::vector<double> samples; /* 3 per sample, body frame */
	const struct OMesh* mesh = nullptr; /* mesh bodies: the mesh and its node grid */
};

/* ------------------------------------------------------------------------ */
/* Triangle meshes (Mesh.cpp:18-110 set-up, Mesh::signedDistance :125-167,    */
/* MeshMeshContactGenerator.cpp:7-119 vertices against the other mesh).      */
/* The signed distance of a point is its distance to the closest point of    */
/* the mesh, negative behind the angle-weighted pseudonormal of the feature   */
/* (face, edge, vertex) holding that point (Baerentzen and Aanaes, IEEE TVCG  */
/* 11 (2005); libigl's signed_distance_pseudonormal in the reference).        */
/* ------------------------------------------------------------------------ */
struct OMesh {
	std::vector<double> verts;   /* body frame, 3 per vertex: the surface samples */
	std::vector<double> tri;     /* per triangle of nonzero area: a, b, c, face normal,
	                                pseudonormals of edges ab, bc, ca and vertices a, b, c */
	double ext[3] = {0, 0, 0};   /* half extents of the grid box before padding */
	std::vector<float> grid;     /* gn^3 nodes */
	int gn = 0;
	double pad = -1.0;
};

/* Pseudonormals.  Vertices with identical coordinates are merged for the
 * adjacency; zero-area triangles take no part in the distance.  Face: the
 * normalised cross product (b - a) x (c - a).  Edge: the normalised sum of the
 * normals of the faces that share it (uniform weighting, Mesh.cpp:66-68).
 * Vertex: the normalised sum of its faces' normals, each times the face's
 * angle at the vertex (angle weighting, Mesh.cpp:64-65).  Sums in face order. */
void omesh_setup(OMesh& m, int nv, const double* v, int nt, const int* f) {
	m.verts.assign(v, v + 3 * (size_t)nv);
	std::map<std::array<double, 3>, int> seen;
	std::vector<int> id(nv);
	for (int i = 0; i < nv; ++i)
		id[i] = seen.emplace(std::array<double, 3>{v[3 * i], v[3 * i + 1], v[3 * i + 2]}, i).first->second;
	std::vector<std::array<double, 3>> face(nt), vnorm(nv, std::array<double, 3>{0.0, 0.0, 0.0});
	std::vector<bool> live(nt, false);
	std::map<std::pair<int, int>, std::array<double, 3>> enorm;
	for (int t = 0; t < nt; ++t) {
		const double* x[3];
		for (int k = 0; k < 3; ++k) x[k] = v + 3 * f[3 * t + k];
		double u[3], w[3], c[3];
		for (int a = 0; a < 3; ++a) { u[a] = x[1][a] - x[0][a]; w[a] = x[2][a] - x[0][a]; }
		cross3(u, w, c);
		const double len = std::sqrt(dot3(c, c));
		if (!(len > 0.0)) continue;
		live[t] = true;
		for (int a = 0; a < 3; ++a) face[t][a] = c[a] / len;
		for (int k = 0; k < 3; ++k) {
			double e1[3], e2[3];
			for (int a = 0; a < 3; ++a) {
				e1[a] = x[(k + 1) % 3][a] - x[k][a];
				e2[a] = x[(k + 2) % 3][a] - x[k][a];
			}
			const double cs = std::min(1.0, std::max(-1.0, dot3(e1, e2) / (std::sqrt(dot3(e1, e1)) * std::sqrt(dot3(e2, e2)))));
			const double ang = std::acos(cs);
			for (int a = 0; a < 3; ++a) vnorm[id[f[3 * t + k]]][a] += ang * face[t][a];
			const int p = id[f[3 * t + k]], q = id[f[3 * t + (k + 1) % 3]];
			auto& e = enorm.emplace(std::make_pair(std::min(p, q), std::max(p, q)),
			                        std::array<double, 3>{0.0, 0.0, 0.0}).first->second;
			for (int a = 0; a < 3; ++a) e[a] += face[t][a];
		}
	}
	auto normalise = [](std::array<double, 3>& x) {
		const double l = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
		if (l > 0.0)
			for (int a = 0; a < 3; ++a) x[a] /= l;
	};
	for (auto& x : vnorm) normalise(x);
	for (auto& kv : enorm) normalise(kv.second);
	m.tri.clear();
	for (int t = 0; t < nt; ++t) {
		if (!live[t]) continue;
		const int c0 = id[f[3 * t]], c1 = id[f[3 * t + 1]], c2 = id[f[3 * t + 2]];
		for (int k = 0; k < 3; ++k)
			for (int a = 0; a < 3; ++a) m.tri.push_back(v[3 * f[3 * t + k] + a]);
		for (int a = 0; a < 3; ++a) m.tri.push_back(face[t][a]);
		const std::pair<int, int> eds[3] = {{std::min(c0, c1), std::max(c0, c1)},
		                                    {std::min(c1, c2), std::max(c1, c2)},
		                                    {std::min(c2, c0), std::max(c2, c0)}};
		for (const auto& e : eds)
			for (int a = 0; a < 3; ++a) m.tri.push_back(enorm.at(e)[a]);
		for (int c : {c0, c1, c2})
			for (int a = 0; a < 3; ++a) m.tri.push_back(vnorm[c][a]);
	}
	double top = 0.0;
	for (int a = 0; a < 3; ++a) {
		m.ext[a] = 0.0;
		for (int i = 0; i < nv; ++i) m.ext[a] = std::max(m.ext[a], std::fabs(v[3 * i + a]));
		top = std::max(top, m.ext[a]);
	}
	for (int a = 0; a < 3; ++a) m.ext[a] = std::max(m.ext[a], 1e-3 * top);
}

/* Closest point of triangle t (a, b, c at t[0..8]) to p by the Voronoi
 * regions of its vertices, edges and face (Ericson, Real-Time Collision
 * Detection, 2005, sec. 5.1.5); returns the squared distance and the
 * pseudonormal of the region's feature in *n. */
double triangle_closest(const double* t, const double* p, const double** n, double* cp) {
	const double *A = t, *B = t + 3, *Cc = t + 6;
	double ab[3], ac[3], ap[3], bp[3], cpp[3];
	for (int a = 0; a < 3; ++a) {
		ab[a] = B[a] - A[a];
		ac[a] = Cc[a] - A[a];
		ap[a] = p[a] - A[a];
		bp[a] = p[a] - B[a];
		cpp[a] = p[a] - Cc[a];
	}
	const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
	const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
	const double d5 = dot3(ab, cpp), d6 = dot3(ac, cpp);
	const double vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
	if (d1 <= 0.0 && d2 <= 0.0) {                       /* vertex a */
		for (int a = 0; a < 3; ++a) cp[a] = A[a];
		*n = t + 21;
	} else if (d3 >= 0.0 && d4 <= d3) {                /* vertex b */
		for (int a = 0; a < 3; ++a) cp[a] = B[a];
		*n = t + 24;
	} else if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {  /* edge ab */
		const double s = d1 / (d1 - d3);
		for (int a = 0; a < 3; ++a) cp[a] = A[a] + s * ab[a];
		*n = t + 12;
	} else if (d6 >= 0.0 && d5 <= d6) {                /* vertex c */
		for (int a = 0; a < 3; ++a) cp[a] = Cc[a];
		*n = t + 27;
	} else if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {  /* edge ca */
		const double s = d2 / (d2 - d6);
		for (int a = 0; a < 3; ++a) cp[a] = A[a] + s * ac[a];
		*n = t + 18;
	} else if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {  /* edge bc */
		const double s = (d4 - d3) / ((d4 - d3) + (d5 - d6));
		for (int a = 0; a < 3; ++a) cp[a] = B[a] + s * (Cc[a] - B[a]);
		*n = t + 15;
	} else {                                           /* face */
		const double den = 1.0 / (va + vb + vc);
		const double sv = vb * den, sw = vc * den;
		for (int a = 0; a < 3; ++a) cp[a] = A[a] + ab[a] * sv + ac[a] * sw;
		*n = t + 9;
	}
	double d[3];
	for (int a = 0; a < 3; ++a) d[a] = p[a] - cp[a];
	return dot3(d, d);
}

/* Signed distance of p to the mesh: the nearest triangle (the first of equal
 * squared distance) gives the distance and the sign. */
double mesh_sdf(const OMesh& m, const double* p) {
	double best = INFINITY, sgn = 1.0;
	const size_t nt = m.tri.size() / 30;
	for (size_t t = 0; t < nt; ++t) {
		const double* n;
		double cp[3];
		const double d2 = triangle_closest(&m.tri[30 * t], p, &n, cp);
		if (d2 < best) {
			double d[3];
			for (int a = 0; a < 3; ++a) d[a] = p[a] - cp[a];
			best = d2;
			sgn = dot3(d, n) < 0.0 ? -1.0 : 1.0;
		}
	}
	return sgn * std::sqrt(best);
}

/* Distance from y >= 0 to the boundary of the m-dimensional ellipsoid with
 * semi-axes a[0] >= ... >= a[m-1] > 0 (Eberly, "Distance from a Point to an
 * Ellipse, an Ellipsoid, or a Hyperellipsoid", 2013), by recursion on the
 * dimension.  Last component positive: the closest point is
 * x_i = r_i y_i / (s + r_i), r_i = (a_i / a_last)^2, s the root of
 * sum_i (r_i z_i / (s + r_i))^2 = 1 (z_i = y_i / a_i) found by bisection of
 * floating-point midpoints on [z_last - 1, |r z| - 1] (0 as the top inside);
 * axes whose component is zero have x_i = 0 and drop out.  Last component
 * zero: the interior closed form x_i = a_i^2 y_i / (a_i^2 - a_last^2) when it
 * lies inside the ellipsoid, else the problem without the last axis. */
static double ellipsoid_distance_nd(int m, const double* a, const double* y) {
	if (m == 1) return std::fabs(y[0] - a[0]);
	const int L = m - 1;
	if (y[L] > 0.0) {
		double ra[3], ry[3];
		int k = 0;
		for (int i = 0; i < L; ++i)
			if (y[i] > 0.0) { ra[k] = a[i]; ry[k] = y[i]; ++k; }
		if (k < L) {
			ra[k] = a[L]; ry[k] = y[L];
			return ellipsoid_distance_nd(k + 1, ra, ry);
		}
		double z[3] = {0, 0, 0}, r[3] = {0, 0, 0}, num[3] = {0, 0, 0};
		double g = 0.0;
		for (int i = 0; i < m; ++i) { z[i] = y[i] / a[i]; g += z[i] * z[i]; }
		g -= 1.0;
		if (g == 0.0) return 0.0;
		for (int i = 0; i < L; ++i) { const double q = a[i] / a[L]; r[i] = q * q; num[i] = r[i] * z[i]; }
		double lo = z[L] - 1.0, hi = 0.0;
		if (!(g < 0.0)) {
			double nn = 0.0;
			for (int i = 0; i < L; ++i) nn += num[i] * num[i];
			nn += z[L] * z[L];
			hi = std::sqrt(nn) - 1.0;
		}
		double s = lo;
		for (int it = 0; it < 1100; ++it) {
			s = 0.5 * (lo + hi);
			if (s == lo || s == hi) break;
			double f = 0.0;
			for (int i = 0; i < L; ++i) { const double t = num[i] / (s + r[i]); f += t * t; }
			const double tl = z[L] / (s + 1.0);
			f += tl * tl;
			f -= 1.0;
			if (f > 0.0) lo = s;
			else if (f < 0.0) hi = s;
			else break;
		}
		double d2 = 0.0;
		for (int i = 0; i < L; ++i) { const double dx = r[i] * y[i] / (s + r[i]) - y[i]; d2 += dx * dx; }
		const double dl = y[L] / (s + 1.0) - y[L];
		d2 += dl * dl;
		return std::sqrt(d2);
	}
	bool interior = true;
	double e[3];
	for (int i = 0; i < L; ++i) {
		const double num = a[i] * y[i], den = a[i] * a[i] - a[L] * a[L];
		if (!(num < den)) { interior = false; break; }
		e[i] = num / den;
	}
	if (interior) {
		double disc = 1.0;
		for (int i = 0; i < L; ++i) disc -= e[i] * e[i];
		if (disc > 0.0) {
			double d2 = 0.0;
			for (int i = 0; i < L; ++i) { const double dx = a[i] * e[i] - y[i]; d2 += dx * dx; }
			const double xl = a[L] * std::sqrt(disc);
			d2 += xl * xl;
			return std::sqrt(d2);
		}
	}
	return ellipsoid_distance_nd(L, a, y);
}

/* Signed distance to the ellipsoid with semi-axes dims: axes ordered by
 * decreasing length (equal lengths in index order), negative inside. */
double ellipsoid_sdf(const double* dims, double px, double py, double pz) {
	const double p[3] = {std::fabs(px), std::fabs(py), std::fabs(pz)};
	int idx[3] = {0, 1, 2};
	std::stable_sort(idx, idx + 3, [&](int u, int v) { return dims[u] > dims[v]; });
	double a[3], y[3];
	for (int i = 0; i < 3; ++i) { a[i] = dims[idx[i]]; y[i] = p[idx[i]]; }
	const double d = ellipsoid_distance_nd(3, a, y);
	double q = 0.0;
	for (int i = 0; i < 3; ++i) { const double z = y[i] / a[i]; q += z * z; }
	return q < 1.0 ? -d : d;
}

double analytic_sdf(const Body& b, double px, double py, double pz) {
	if (b.type == PDC_SHAPE_SPHERE) {
		return std::sqrt(px * px + py * py + pz * pz) - b.dims[0];
	}
	if (b.type == PDC_SHAPE_ELLIPSOID) return ellipsoid_sdf(b.dims, px, py, pz);
	const double qx = std::fabs(px) - b.dims[0];
	const double qy = std::fabs(py) - b.dims[1];
	const double qz = std::fabs(pz) - b.dims[2];
	const double ox = std::max(qx, 0.0), oy = std::max(qy, 0.0), oz = std::max(qz, 0.0);
	const double outside = std::sqrt(ox * ox + oy * oy + oz * oz);
	const double inside = std::min(std::max(qx, std::max(qy, qz)), 0.0);
	return outside + inside;
}

/* Grid: gn^3 nodes over [-e(1+pad), e(1+pad)] per axis, e = half extent
 * (a mesh: its largest |coordinate| per axis, at least 1e-3 of the largest). */
void setup_grid(Body& b, int n, double pad_frac) {
	b.gn = n;
	double e[3];
	if (b.type == PDC_SHAPE_SPHERE) e[0] = e[1] = e[2] = b.dims[0];
	else if (b.type == PDC_SHAPE_MESH) { e[0] = b.mesh->ext[0]; e[1] = b.mesh->ext[1]; e[2] = b.mesh->ext[2]; }
	else { e[0] = b.dims[0]; e[1] = b.dims[1]; e[2] = b.dims[2]; }
	for (int a = 0; a < 3; ++a) {
		const double h = e[a] * (1.0 + pad_frac);
		b.ghalf[a] = h;
		b.glo[a] = -h;
		b.gsp[a] = (2.0 * h) / (double)(n - 1);
	}
}
inline float grid_node(const Body& b, int i, int j, int k) {
	if (b.mesh) return b.mesh->grid[((size_t)k * b.gn + j) * b.gn + i];
	const double px = b.glo[0] + (double)i * b.gsp[0];
	const double py = b.glo[1] + (double)j * b.gsp[1];
	const double pz = b.glo[2] + (double)k * b.gsp[2];
	return (float)analytic_sdf(b, px, py, pz);
}
/* Trilinear value and gradient of the FP32 node grid at local point q.
 * Returns false when q lies outside the grid box. */
bool grid_query(const Body& b, const double* q, double* sd, double* grad) {
	double t[3];
	int i[3];
	const double top = (double)(b.gn - 1);
	for (int a = 0; a < 3; ++a) {
		const double g = (q[a] - b.glo[a]) / b.gsp[a];
		if (!(g >= 0.0 && g <= top)) return false;
		int ia = (int)g;
		if (ia > b.gn - 2) ia = b.gn - 2;
		i[a] = ia;
		t[a] = g - (double)ia;
	}
	const double v000 = grid_node(b, i[0], i[1], i[2]);
	const double v100 = grid_node(b, i[0] + 1, i[1], i[2]);
	const double v010 = grid_node(b, i[0], i[1] + 1, i[2]);
	const double v110 = grid_node(b, i[0] + 1, i[1] + 1, i[2]);
	const double v001 = grid_node(b, i[0], i[1], i[2] + 1);
	const double v101 = grid_node(b, i[0] + 1, i[1], i[2] + 1);
	const double v011 = grid_node(b, i[0], i[1] + 1, i[2] + 1);
	const double v111 = grid_node(b, i[0] + 1, i[1] + 1, i[2] + 1);
	const double tx = t[0], ty = t[1], tz = t[2];
	const double c00 = v000 + tx * (v100 - v000);
	const double c10 = v010 + tx * (v110 - v010);
	const double c01 = v001 + tx * (v101 - v001);
	const double c11 = v011 + tx * (v111 - v011);
	const double c0 = c00 + ty * (c10 - c00);
	const double c1 = c01 + ty * (c11 - c01);
	*sd = c0 + tz * (c1 - c0);
	const double ux = 1.0 - tx, uy = 1.0 - ty, uz = 1.0 - tz;
	grad[0] = (uy * uz * (v100 - v000) + ty * uz * (v110 - v010) +
	           uy * tz * (v101 - v001) + ty * tz * (v111 - v011)) / b.gsp[0];
	grad[1] = (ux * uz * (v010 - v000) + tx * uz * (v110 - v100) +
	           ux * tz * (v011 - v001) + tx * tz * (v111 - v101)) / b.gsp[1];
	grad[2] = (ux * uy * (v001 - v000) + tx * uy * (v101 - v100) +
	           ux * ty * (v011 - v010) + tx * ty * (v111 - v110)) / b.gsp[2];
	return true;
}

/* Surface samples.  Box: on each face (axis 0..2, side -,+) a cell-centred
 * nu x nw grid, n = max(1, round(edge / spacing)).  Sphere: Fibonacci
 * lattice with max(8, round(4 pi r^2 / spacing^2)) points.  Ellipsoid: that
 * lattice scaled by the semi-axes.  Computed in FP64, kept as FP32 values
 * (setup_samples). */
void surface_samples(Body& b, double spacing) {
	b.samples.clear();
	if (b.type == PDC_SHAPE_MESH) { /* the mesh's vertices */
		b.samples = b.mesh->verts;
		return;
	}
	if (b.type == PDC_SHAPE_SPHERE || b.type == PDC_SHAPE_ELLIPSOID) {
		/* the ellipsoid's lattice is the unit one scaled by the semi-axes;
		 * its count uses the area 4 pi (ab + ac + bc) / 3 (the p = 1 case of
		 * Thomsen's approximation, exact for a sphere) */
		const bool ell = b.type == PDC_SHAPE_ELLIPSOID;
		const double ax = b.dims[0], ay = ell ? b.dims[1] : b.dims[0], az = ell ? b.dims[2] : b.dims[0];
		const double PI = 3.14159265358979323846;
		const double mean_sq = ell ? (ax * ay + ax * az + ay * az) / 3.0 : ax * ax;
		long n = std::lround(4.0 * PI * mean_sq / (spacing * spacing));
		if (n < 8) n = 8;
		const double ga = PI * (3.0 - std::sqrt(5.0));
		for (long i = 0; i < n; ++i) {
			const double y = 1.0 - (2.0 * (double)i + 1.0) / (double)n;
			const double rho = std::sqrt(std::max(0.0, 1.0 - y * y));
			const double phi = (double)i * ga;
			b.samples.push_back(ax * (rho * std::cos(phi)));
			b.samples.push_back(ay * y);
			b.samples.push_back(az * (rho * std::sin(phi)));
		}
		return;
	}
	for (int a = 0; a < 3; ++a) {
		const int u = (a + 1) % 3, w = (a + 2) % 3;
		long nu = std::lround(2.0 * b.dims[u] / spacing);
		long nw = std::lround(2.0 * b.dims[w] / spacing);
		if (nu < 1) nu = 1;
		if (nw < 1) nw = 1;
		const double du = 2.0 * b.dims[u] / (double)nu;
		const double dw = 2.0 * b.dims[w] / (double)nw;
		for (int s = -1; s <= 1; s += 2) {
			for (long iu = 0; iu < nu; ++iu) {
				for (long iw = 0; iw < nw; ++iw) {
					double p[3];
					p[a] = (double)s * b.dims[a];
					p[u] = -b.dims[u] + ((double)iu + 0.5) * du;
					p[w] = -b.dims[w] + ((double)iw + 0.5) * dw;
					b.samples.push_back(p[0]);
					b.samples.push_back(p[1]);
					b.samples.push_back(p[2]);
				}
			}
		}
	}
}

void setup_samples(Body& b, double spacing) {
	surface_samples(b, spacing);
	for (double& v : b.samples) v = (double)(float)v;
}

struct Contact {
	int id[2];
	double pos[3];
	double depth;
	double T[2][18];
	double lf[3];
};

/* ------------------------------------------------------------------------ */
/* block-sparse Cholesky with greedy minimum-degree ordering                 */
/* ------------------------------------------------------------------------ */
struct Blk { double a[36]; };

bool chol6(const double* A, double* L) {
	std::memset(L, 0, 36 * sizeof(double));
	for (int j = 0; j < 6; ++j) {
		double d = A[j * 6 + j];
		for (int k = 0; k < j; ++k) d -= L[j * 6 + k] * L[j * 6 + k];
		if (!(d > 0)) return false;
		const double ljj = std::sqrt(d);
		L[j * 6 + j] = ljj;
		for (int i = j + 1; i < 6; ++i) {
			double s = A[i * 6 + j];
			for (int k = 0; k < j; ++k) s -= L[i * 6 + k] * L[j * 6 + k];
			L[i * 6 + j] = s / ljj;
		}
	}
	return true;
}
/* solve L y = b (in place) */
void fwd6(const double* L, double* b) {
	for (int i = 0; i < 6; ++i) {
		double s = b[i];
		for (int k = 0; k < i; ++k) s -= L[i * 6 + k] * b[k];
		b[i] = s / L[i * 6 + i];
	}
}
/* solve L^T x = b (in place) */
void bwd6(const double* L, double* b) {
	for (int i = 5; i >= 0; --i) {
		double s = b[i];
		for (int k = i + 1; k < 6; ++k) s -= L[k * 6 + i] * b[k];
		b[i] = s / L[i * 6 + i];
	}
}

/* Block matrix in block-CSR, symmetric pattern, 6x6 row-major blocks. */
struct BCSR {
	int n = 0;
	std::vector<int> rowptr, col;
	std::vector<Blk> val;
	int find(int r, int c) const {
		for (int k = rowptr[r]; k < rowptr[r + 1]; ++k)
			if (col[k] == c) return k;
		return -1;
	}
};

struct BlockCholesky {
	int n = 0;
	std::vector<int> perm, iperm;
	std::vector<std::vector<int>> pat; /* below-diagonal positions per column */
	std::vector<Blk> Ld;
	std::vector<std::vector<Blk>> Lc;

	void analyze(const BCSR& A) {
		n = A.n;
		std::vector<std::vector<int>> adj(n);
		for (int r = 0; r < n; ++r)
			for (int k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k)
				if (A.col[k] != r) adj[r].push_back(A.col[k]);
		for (auto& a : adj) {
			std::sort(a.begin(), a.end());
			a.erase(std::unique(a.begin(), a.end()), a.end());
		}
		std::vector<char> done(n, 0);
		perm.clear();
		std::vector<std::vector<int>> nbOrig(n);
		for (int step = 0; step < n; ++step) {
			int best = -1;
			size_t bd = (size_t)-1;
			for (int v = 0; v < n; ++v)
				if (!done[v] && adj[v].size() < bd) { bd = adj[v].size(); best = v; }
			const int v = best;
			done[v] = 1;
			perm.push_back(v);
			nbOrig[v] = adj[v];
			const std::vector<int> nb = adj[v];
			for (int a : nb) {
				auto& aa = adj[a];
				aa.erase(std::remove(aa.begin(), aa.end(), v), aa.end());
				for (int b : nb)
					if (b != a) aa.push_back(b);
				std::sort(aa.begin(), aa.end());
				aa.erase(std::unique(aa.begin(), aa.end()), aa.end());
			}
			adj[v].clear();
		}
		iperm.assign(n, 0);
		for (int k = 0; k < n; ++k) iperm[perm[k]] = k;
		pat.assign(n, {});
		for (int k = 0; k < n; ++k) {
			for (int o : nbOrig[perm[k]]) pat[k].push_back(iperm[o]);
			std::sort(pat[k].begin(), pat[k].end());
		}
	}
	bool factorize(const BCSR& A) {
		/* working lower blocks S(i, j), i >= j, positions */
		std::unordered_map<long long, Blk> S;
		S.reserve(n * 8);
		auto key = [&](int i, int j) { return (long long)i * n + j; };
		for (int r = 0; r < n; ++r) {
			for (int k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k) {
				const int c = A.col[k];
				const int pr = iperm[r], pc = iperm[c];
				if (pr >= pc) S[key(pr, pc)] = A.val[k];
			}
		}
		for (int k = 0; k < n; ++k)
			for (int i : pat[k])
				if (!S.count(key(i, k))) { Blk z; std::memset(z.a, 0, sizeof(z.a)); S[key(i, k)] = z; }
		Ld.assign(n, Blk());
		Lc.assign(n, {});
		for (int k = 0; k < n; ++k) {
			if (!chol6(S[key(k, k)].a, Ld[k].a)) return false;
			const double* L = Ld[k].a;
			Lc[k].resize(pat[k].size());
			for (size_t ii = 0; ii < pat[k].size(); ++ii) {
				/* X = S(i,k) L^-T : each row x of X solves L x^T = s^T */
				const Blk& s = S[key(pat[k][ii], k)];
				Blk& X = Lc[k][ii];
				for (int r = 0; r < 6; ++r) {
					double row[6];
					for (int c = 0; c < 6; ++c) row[c] = s.a[r * 6 + c];
					fwd6(L, row);
					for (int c = 0; c < 6; ++c) X.a[r * 6 + c] = row[c];
				}
			}
			for (size_t ii = 0; ii < pat[k].size(); ++ii) {
				for (size_t jj = 0; jj <= ii; ++jj) {
					const int i = pat[k][ii], j = pat[k][jj];
					/* i >= j since pat is sorted */
					Blk& t = S[key(i, j)];
					const Blk& Li = Lc[k][ii];
					const Blk& Lj = Lc[k][jj];
					for (int r = 0; r < 6; ++r)
						for (int c = 0; c < 6; ++c) {
							double s = 0;
							for (int m = 0; m < 6; ++m) s += Li.a[r * 6 + m] * Lj.a[c * 6 + m];
							t.a[r * 6 + c] -= s;
						}
				}
			}
		}
		return true;
	}
	void solve(const double* b, double* x) const {
		std::vector<double> y(6 * (size_t)n);
		for (int k = 0; k < n; ++k)
			for (int r = 0; r < 6; ++r) y[6 * k + r] = b[6 * perm[k] + r];
		for (int k = 0; k < n; ++k) {
			fwd6(Ld[k].a, &y[6 * k]);
			for (size_t ii = 0; ii < pat[k].size(); ++ii) {
				const int i = pat[k][ii];
				const Blk& Lik = Lc[k][ii];
				for (int r = 0; r < 6; ++r) {
					double s = 0;
					for (int m = 0; m < 6; ++m) s += Lik.a[r * 6 + m] * y[6 * k + m];
					y[6 * i + r] -= s;
				}
			}
		}
		for (int k = n - 1; k >= 0; --k) {
			for (size_t ii = 0; ii < pat[k].size(); ++ii) {
				const int i = pat[k][ii];
				const Blk& Lik = Lc[k][ii];
				for (int m = 0; m < 6; ++m) {
					double s = 0;
					for (int r = 0; r < 6; ++r) s += Lik.a[r * 6 + m] * y[6 * i + r];
					y[6 * k + m] -= s;
				}
			}
			bwd6(Ld[k].a, &y[6 * k]);
		}
		for (int k = 0; k < n; ++k)
			for (int r = 0; r < 6; ++r) x[6 * perm[k] + r] = y[6 * k + r];
	}
};

/* y = A x on block-CSR */
void bcsr_mv(const BCSR& A, const double* x, double* y) {
#pragma omp parallel for schedule(static)
	for (int r = 0; r < A.n; ++r) {
		double acc[6] = {0, 0, 0, 0, 0, 0};
		for (int k = A.rowptr[r]; k < A.rowptr[r + 1]; ++k) {
			const double* a = A.val[k].a;
			const double* xc = x + 6 * A.col[k];
			for (int i = 0; i < 6; ++i)
				for (int j = 0; j < 6; ++j) acc[i] += a[i * 6 + j] * xc[j];
		}
		for (int i = 0; i < 6; ++i) y[6 * r + i] = acc[i];
	}
}

/* Inverse of a dense symmetric matrix (n x n, row-major) through its
 * Cholesky factor; false (and `inv` untouched) when a pivot is not positive. */
bool dense_spd_inverse(int n, const std::vector<double>& G, std::vector<double>& inv) {
	std::vector<double> L((size_t)n * n, 0.0);
	for (int j = 0; j < n; ++j) {
		double d = G[(size_t)j * n + j];
		for (int k = 0; k < j; ++k) d -= L[(size_t)j * n + k] * L[(size_t)j * n + k];
		if (!(d > 0) || !std::isfinite(d)) return false;
		const double ljj = std::sqrt(d);
		L[(size_t)j * n + j] = ljj;
		for (int i = j + 1; i < n; ++i) {
			double t = G[(size_t)i * n + j];
			for (int k = 0; k < j; ++k) t -= L[(size_t)i * n + k] * L[(size_t)j * n + k];
			L[(size_t)i * n + j] = t / ljj;
		}
	}
	inv.assign((size_t)n * n, 0.0);
	std::vector<double> e(n);
	for (int c = 0; c < n; ++c) {
		for (int i = 0; i < n; ++i) {
			double t = i == c ? 1.0 : 0.0;
			for (int k = 0; k < i; ++k) t -= L[(size_t)i * n + k] * e[k];
			e[i] = t / L[(size_t)i * n + i];
		}
		for (int i = n - 1; i >= 0; --i) {
			double t = e[i];
			for (int k = i + 1; k < n; ++k) t -= L[(size_t)k * n + i] * e[k];
			e[i] = t / L[(size_t)i * n + i];
		}
		for (int i = 0; i < n; ++i) inv[(size_t)i * n + c] = e[i];
	}
	return true;
}

/* Block preconditioner of the PCG: for every group of block rows (the rows
 * whose `rep` is the same; every row alone when `rep` is null) the inverse of
 * the principal submatrix of A on the group's rows, or, when that is not
 * positive definite, the inverse of its diagonal. */
struct GroupPrec {
	std::vector<std::vector<int>> groups;
	std::vector<std::vector<double>> inv;
	void build(const BCSR& A, const int* rep) {
		const int n = A.n;
		groups.clear();
		std::vector<int> gi(n, -1);
		for (int r = 0; r < n; ++r) {
			const int k = rep ? rep[r] : r;
			if (gi[k] < 0) { gi[k] = (int)groups.size(); groups.emplace_back(); }
			groups[gi[k]].push_back(r);
		}
		inv.assign(groups.size(), {});
#pragma omp parallel for schedule(dynamic)
		for (int g = 0; g < (int)groups.size(); ++g) {
			const std::vector<int>& mem = groups[g];
			const int d = 6 * (int)mem.size();
			std::vector<double> G((size_t)d * d, 0.0);
			for (size_t u = 0; u < mem.size(); ++u)
				for (size_t v = 0; v < mem.size(); ++v) {
					const int k = A.find(mem[u], mem[v]);
					if (k < 0) continue;
					for (int i = 0; i < 6; ++i)
						for (int j = 0; j < 6; ++j)
							G[(6 * u + i) * (size_t)d + 6 * v + j] = A.val[k].a[i * 6 + j];
				}
			if (!dense_spd_inverse(d, G, inv[g])) {
				inv[g].assign((size_t)d * d, 0.0);
				for (int i = 0; i < d; ++i) inv[g][(size_t)i * d + i] = 1.0 / G[(size_t)i * d + i];
			}
		}
	}
	void apply(const std::vector<double>& in, std::vector<double>& out) const {
#pragma omp parallel for schedule(dynamic)
		for (int g = 0; g < (int)groups.size(); ++g) {
			const std::vector<int>& mem = groups[g];
			const int d = 6 * (int)mem.size();
			for (int i = 0; i < d; ++i) {
				double s = 0;
				for (int j = 0; j < d; ++j) s += inv[g][(size_t)i * d + j] * in[6 * mem[j / 6] + j % 6];
				out[6 * mem[i / 6] + i % 6] = s;
			}
		}
	}
};

/* Groups of strongly coupled bodies for the preconditioner of the standard
 * PCG form (pcg_tol < 1e-9).  This is this project's own preconditioner, in
 * place of the reference's AMGCL smoothed-aggregation AMG
 * (PrimalDualForceSolver.cpp:395-415), whose aggregation it resembles in
 * grouping unknowns by coupling strength; the strength here is read off the
 * contact list alone, so the groups are fixed for a timestep.
 *   Strength of the coupling of two groups G and H of dynamic bodies:
 *   n_GH / sqrt(w_G w_H), n_GH the contacts between bodies of G and bodies of
 *   H, w the contacts of a group's bodies (static partners included).
 *   Three levels merge groups of size 1, 2, 4 into groups of 2, 4, 8.  A
 *   level runs up to three rounds of handshake matching: every unmerged group
 *   of the level's size picks the unmerged neighbour group of that size with
 *   the strongest coupling (ties: the smaller key), and two groups that pick
 *   each other merge.  A group's key is the smallest key of its bodies.
 *   Static bodies stay alone.
 * Inputs per body: key (the global id), stat, w; per pair of bodies with
 * contacts between them: (a, b, count).  Returns per body the index of the
 * body holding its group's key. */
std::vector<int> group_bodies(int n, const std::vector<long long>& key, const std::vector<char>& stat,
                              const std::vector<long long>& w, const std::vector<std::array<int, 3>>& pairs) {
	std::vector<std::vector<std::pair<int, int>>> adj(n);
	for (const auto& pr : pairs)
		if (!stat[pr[0]] && !stat[pr[1]] && pr[0] != pr[1]) {
			adj[pr[0]].push_back({pr[1], pr[2]});
			adj[pr[1]].push_back({pr[0], pr[2]});
		}
	std::vector<int> rep(n), size(n, 1);
	std::vector<long long> wg(w);
	std::vector<std::vector<int>> mem(n);
	for (int r = 0; r < n; ++r) { rep[r] = r; mem[r] = {r}; }
	for (int level = 0; level < 3; ++level) {
		const int s = 1 << level;
		auto eligible = [&](int g) { return rep[g] == g && !stat[g] && size[g] == s; };
		for (int round = 0; round < 3; ++round) {
			std::vector<int> best(n, -1);
			for (int g = 0; g < n; ++g) {
				if (!eligible(g)) continue;
				std::map<int, long long> nGH;
				for (int m : mem[g])
					for (const auto& e : adj[m]) {
						const int h = rep[e.first];
						if (h != g && eligible(h)) nGH[h] += e.second;
					}
				int bh = -1;
				long long bn = 0;
				for (const auto& kv : nGH) {
					const int h = kv.first;
					const long long nh = kv.second;
					if (bh < 0) { bh = h; bn = nh; continue; }
					/* nh^2 / w_h against bn^2 / w_bh, exactly */
					const unsigned __int128 lhs = (unsigned __int128)(nh * nh) * (unsigned long long)wg[bh];
					const unsigned __int128 rhs = (unsigned __int128)(bn * bn) * (unsigned long long)wg[h];
					if (lhs > rhs || (lhs == rhs && key[h] < key[bh])) { bh = h; bn = nh; }
				}
				best[g] = bh;
			}
			bool merged = false;
			for (int g = 0; g < n; ++g) {
				const int h = best[g];
				if (h < 0 || best[h] != g || !(key[g] < key[h])) continue;
				for (int m : mem[h]) rep[m] = g;
				mem[g].insert(mem[g].end(), mem[h].begin(), mem[h].end());
				mem[h].clear();
				size[g] = 2 * s;
				size[h] = 0;
				wg[g] += wg[h];
				merged = true;
			}
			if (!merged) break;
		}
	}
	return rep;
}

/* PCG preconditioned by the inverse blocks of the groups given by `rep` (per
 * block row, the row holding its group's key; null: every row alone, block
 * Jacobi).  Starts from x as given when `warm` (r = b - A x), else from
 * zero.  Iterates while it < maxit and |r| > tol |b| (the convergence test of
 * a textbook CG, as in AMGCL's cg, whose initial guess is the caller's X:
 * PrimalDualForceSolver.cpp:406-414).  Returns the iteration count. */
int bcsr_pcg(const BCSR& A, const double* b, double tol, int maxit, double* x, bool warm,
             const int* rep = nullptr) {
	const int n = A.n;
	const size_t m = 6 * (size_t)n;
	GroupPrec P;
	P.build(A, rep);
	std::vector<double> r(b, b + m), z(m), p(m), q(m);
	if (warm) {
		/* the warm start is kept only when its residual is smaller than that
		 * of the zero start, |b - A x0| < |b|: a start far from the solution
		 * (the increment of a gradient-descent fallback, or of an earlier
		 * Newton iteration) puts the relative tolerance below the rounding
		 * level of its residual */
		bcsr_mv(A, x, q.data());
		for (size_t i = 0; i < m; ++i) r[i] = b[i] - q[i];
		const double r2 = fixed_sum(m, [&](size_t i) { return r[i] * r[i]; });
		const double b2 = fixed_sum(m, [&](size_t i) { return b[i] * b[i]; });
		if (!(r2 < b2)) {
			std::fill(x, x + m, 0.0);
			r.assign(b, b + m);
		}
	} else {
		std::fill(x, x + m, 0.0);
	}
	auto prec = [&](const std::vector<double>& in, std::vector<double>& out) { P.apply(in, out); };
	const double bn = std::sqrt(fixed_sum(m, [&](size_t i) { return b[i] * b[i]; }));
	const double eps = tol * bn;
	prec(r, z);
	p = z;
	double rz = fixed_sum(m, [&](size_t i) { return r[i] * z[i]; });
	int it = 0;
	for (; it < maxit; ++it) {
		const double rn = std::sqrt(fixed_sum(m, [&](size_t i) { return r[i] * r[i]; }));
		if (!(rn > eps)) break;
		if (it > 0) {
			prec(r, z);
			const double rzn = fixed_sum(m, [&](size_t i) { return r[i] * z[i]; });
			const double beta = rzn / rz;
			rz = rzn;
#pragma omp parallel for schedule(static)
			for (long i = 0; i < (long)m; ++i) p[i] = z[i] + beta * p[i];
		}
		bcsr_mv(A, p.data(), q.data());
		const double pq = fixed_sum(m, [&](size_t i) { return p[i] * q[i]; });
		const double alpha = rz / pq;
#pragma omp parallel for schedule(static)
		for (long i = 0; i < (long)m; ++i) {
			x[i] += alpha * p[i];
			r[i] -= alpha * q[i];
		}
	}
	return it;
}

}  // namespace

/* ------------------------------------------------------------------------ */
/* scene                                                                     */
/* ------------------------------------------------------------------------ */
struct orc_scene {
	pdc_params p;
	std::vector<std::unique_ptr<OMesh>> meshes;
	int threads = 0;
	int solver = ORC_SOLVER_REFERENCE;
	std::vector<Body> bodies;
	std::vector<Contact> contacts;
	int raw_count = 0, ties = 0;
	std::vector<double> trace;
	pdc_solve_stats last;
	/* dense capture of the first assembled system (tests) */
	bool capture = false;
	std::vector<double> capA, capRhs;
	std::vector<int> capL2G;
};

namespace {

int max_threads() {
#ifdef _OPENMP
	return omp_get_max_threads();
#else
	return 1;
#endif
}
void set_threads(int t) {
#ifdef _OPENMP
	omp_set_num_threads(t < 1 ? 1 : t);
#else
	(void)t;
#endif
}

/* ---------------- (a) grid-SDF contact generation ---------------- */
void world_aabb(const Body& b, double* lo, double* hi) {
	double R[9];
	quat_rot(b.q, R);
	for (int r = 0; r < 3; ++r) {
		const double h = std::fabs(R[r * 3 + 0]) * b.ghalf[0] + std::fabs(R[r * 3 + 1]) * b.ghalf[1] +
		                 std::fabs(R[r * 3 + 2]) * b.ghalf[2];
		lo[r] = b.x[r] - h;
		hi[r] = b.x[r] + h;
	}
}

struct Raw { double pos[3]; double n[3]; double depth; };

/* Tests the samples of `qb` against the grid of `gb`; sign = -1 when qb is
 * objId[0] (normal = -grad of the other body), +1 when it is objId[1]
 * (MeshMeshContactGenerator.cpp:43-55 convention). */
void sample_pass(const Body& qb, const Body& gb, double sign, double tie_band,
                 std::vector<Raw>& out, int& ties) {
	double Rq[9], Rg[9];
	quat_rot(qb.q, Rq);
	quat_rot(gb.q, Rg);
	const size_t ns = qb.samples.size() / 3;
	for (size_t s = 0; s < ns; ++s) {
		const double* l = &qb.samples[3 * s];
		double p[3];
		for (int r = 0; r < 3; ++r)
			p[r] = Rq[r * 3 + 0] * l[0] + Rq[r * 3 + 1] * l[1] + Rq[r * 3 + 2] * l[2] + qb.x[r];
		const double d[3] = {p[0] - gb.x[0], p[1] - gb.x[1], p[2] - gb.x[2]};
		double q[3];
		for (int c = 0; c < 3; ++c) q[c] = Rg[0 * 3 + c] * d[0] + Rg[1 * 3 + c] * d[1] + Rg[2 * 3 + c] * d[2];
		double sd, g[3];
		if (!grid_query(gb, q, &sd, g)) continue;
		if (std::fabs(sd) <= tie_band) ++ties;
		if (!(sd < 0.0)) continue;
		double gw[3];
		for (int r = 0; r < 3; ++r) gw[r] = Rg[r * 3 + 0] * g[0] + Rg[r * 3 + 1] * g[1] + Rg[r * 3 + 2] * g[2];
		const double gg = gw[0] * gw[0] + gw[1] * gw[1] + gw[2] * gw[2];
		if (!(gg > 0.0)) continue;
		const double gl = std::sqrt(gg);
		Raw rc;
		for (int r = 0; r < 3; ++r) {
			rc.pos[r] = p[r];
			rc.n[r] = sign * (gw[r] / gl);
		}
		rc.depth = -sd;
		out.push_back(rc);
	}
}

bool pos_less(const double* a, const double* b) {
	if (a[0] != b[0]) return a[0] < b[0];
	if (a[1] != b[1]) return a[1] < b[1];
	return a[2] < b[2];
}

/* Per pair: raw sample hits -> frames -> lexicographic sort -> chain merge
 * (DynamicSystem.cpp:341-381). */
void pair_contacts(const std::vector<Body>& bodies, int a, int b, double tie_band,
                   std::vector<Contact>& out, int& raw, int& ties) {
	const Body& A = bodies[a];
	const Body& B = bodies[b];
	std::vector<Raw> rs;
	if (A.emit) sample_pass(A, B, -1.0, tie_band, rs, ties);
	if (B.emit) sample_pass(B, A, +1.0, tie_band, rs, ties);
	raw += (int)rs.size();
	if (rs.empty()) return;
	std::vector<Contact> inter(rs.size());
	for (size_t i = 0; i < rs.size(); ++i) {
		Contact& c = inter[i];
		c.id[0] = a;
		c.id[1] = b;
		std::memcpy(c.pos, rs[i].pos, sizeof(c.pos));
		c.depth = std::fabs(rs[i].depth);
		c.lf[0] = c.lf[1] = c.lf[2] = 0;
		normal_to_transform(c.pos, rs[i].n, A.x, B.x, c.T[0], c.T[1]);
	}
	std::sort(inter.begin(), inter.end(),
	          [](const Contact& u, const Contact& v) { return pos_less(u.pos, v.pos); });
	size_t i = 0;
	while (i < inter.size()) {
		size_t j = i + 1;
		while (j < inter.size()) {
			const double d0 = inter[j].pos[0] - inter[j - 1].pos[0];
			const double d1 = inter[j].pos[1] - inter[j - 1].pos[1];
			const double d2 = inter[j].pos[2] - inter[j - 1].pos[2];
			if (d0 * d0 + d1 * d1 + d2 * d2 > 1e-8) break;
			++j;
		}
		if (j - i == 1) {
			out.push_back(inter[i]);
		} else {
			Contact t = inter[i];
			double n[3] = {t.T[0][0], t.T[0][1], t.T[0][2]};
			for (size_t k = i + 1; k < j; ++k) {
				n[0] += inter[k].T[0][0];
				n[1] += inter[k].T[0][1];
				n[2] += inter[k].T[0][2];
				t.depth = std::min(inter[k].depth, t.depth);
			}
			const double nn = std::sqrt(dot3(n, n));
			if (nn > 0) { n[0] /= nn; n[1] /= nn; n[2] /= nn; }
			normal_to_transform(t.pos, n, A.x, B.x, t.T[0], t.T[1]);
			out.push_back(t);
		}
		i = j;
	}
}

double tie_band_of(const std::vector<Body>& bodies) {
	double m = 1e300;
	for (const Body& b : bodies)
		for (int a = 0; a < 3; ++a) m = std::min(m, b.gsp[a]);
	return 1e-7 * m;
}

void detect(orc_scene* s) {
	auto& bodies = s->bodies;
	const int n = (int)bodies.size();
	std::vector<double> lo(3 * (size_t)n), hi(3 * (size_t)n);
	for (int i = 0; i < n; ++i) world_aabb(bodies[i], &lo[3 * i], &hi[3 * i]);
	std::vector<std::pair<int, int>> pairs;
	for (int a = 0; a < n; ++a)
		for (int b = a + 1; b < n; ++b) {
			if (bodies[a].stat && bodies[b].stat) continue;
			bool ov = true;
			for (int r = 0; r < 3; ++r)
				if (lo[3 * a + r] > hi[3 * b + r] || lo[3 * b + r] > hi[3 * a + r]) { ov = false; break; }
			if (ov) pairs.emplace_back(a, b);
		}
	const double tb = tie_band_of(bodies);
	std::vector<std::vector<Contact>> per(pairs.size());
	std::vector<int> raw(pairs.size(), 0), ties(pairs.size(), 0);
	set_threads(s->threads > 0 ? s->threads : (pairs.size() < 32 ? 1 : max_threads()));
#pragma omp parallel for schedule(dynamic, 4)
	for (long k = 0; k < (long)pairs.size(); ++k)
		pair_contacts(bodies, pairs[k].first, pairs[k].second, tb, per[k], raw[k], ties[k]);
	s->contacts.clear();
	s->raw_count = 0;
	s->ties = 0;
	for (size_t k = 0; k < pairs.size(); ++k) {
		s->contacts.insert(s->contacts.end(), per[k].begin(), per[k].end());
		s->raw_count += raw[k];
		s->ties += ties[k];
	}
}

/* ---------------- (b)-(e) the primal-dual Newton solve ---------------- */
struct Solver {
	const pdc_params& p;
	std::vector<Contact>& cs;     /* local ids */
	std::vector<Body*>& ob;       /* local bodies */
	orc_scene* scene;
	int C, N;
	bool nsc;
	int fd, cd;
	Lorentz lz;
	double mbar = 1, s = 1;
	/* per body */
	std::vector<double> Ms;       /* 36 per body: scaled mass matrix, dense 6x6 */
	std::vector<double> Minv;     /* 36 per body: physical inverse mass matrix */
	std::vector<char> stat;
	std::vector<std::vector<std::pair<int, int>>> inc; /* (contact, side) per body */
	BCSR A;
	/* per contact */
	std::vector<double> K, cvals, grads, hess, invC, KnInv, KtInv, sf;
	double compliance = 0, reduction = 0;

	Solver(const pdc_params& pp, std::vector<Contact>& c, std::vector<Body*>& o, orc_scene* sc)
		: p(pp), cs(c), ob(o), scene(sc) {
		C = (int)cs.size();
		N = (int)ob.size();
		nsc = p.force_model == PDC_NON_SMOOTH_CONTACT;
		fd = nsc ? 3 : 1;
		cd = nsc ? 2 : 1;
		lz.mu = p.friction_coeff;
	}

	/* u = toLocal(vels) = -T0^T v_a - T1^T v_b  (ContactInfo.h:46-51) */
	void to_local(int i, const double* v, double* u) const {
		const Contact& c = cs[i];
		const double* va = v + 6 * c.id[0];
		const double* vb = v + 6 * c.id[1];
		for (int k = 0; k < 3; ++k) {
			double s0 = 0, s1 = 0;
			for (int r = 0; r < 6; ++r) {
				s0 += c.T[0][k * 6 + r] * va[r];
				s1 += c.T[1][k * 6 + r] * vb[r];
			}
			u[k] = -s0 - s1;
		}
	}
	/* T_j * f (the 6-vector hTf = -toGlobal(j, f)) */
	void t_times(int i, int j, const double* f, double* o) const {
		const Contact& c = cs[i];
		for (int r = 0; r < 6; ++r) {
			double a = 0;
			for (int k = 0; k < fd; ++k) a += c.T[j][k * 6 + r] * f[k];
			o[r] = a;
		}
	}
	bool active(int b) const { return !stat[b]; }

	/* group_bodies on this solve's contacts: keys are the global ids, w the
	 * contact count of each body, one pair per two bodies in contact */
	std::vector<int> contact_groups() const {
		std::vector<long long> key(N), w(N, 0);
		for (int b = 0; b < N; ++b) key[b] = (long long)(ob[b] - &scene->bodies[0]);
		std::map<std::pair<int, int>, int> cnt;
		for (const Contact& c : cs) {
			w[c.id[0]]++;
			w[c.id[1]]++;
			cnt[{std::min(c.id[0], c.id[1]), std::max(c.id[0], c.id[1])}]++;
		}
		std::vector<std::array<int, 3>> pairs;
		for (const auto& kv : cnt) pairs.push_back({kv.first.first, kv.first.second, kv.second});
		return group_bodies(N, key, stat, w, pairs);
	}

	/* PrimalDualForceSolver.cpp:23-78 */
	void nondim(double dt, const double* g) {
		double avg = 0;
		int act = 0;
		for (int b = 0; b < N; ++b)
			if (active(b)) {
				double R[9], Iw[9];
				quat_rot(ob[b]->q, R);
				rot_diag_rot(R, ob[b]->I0, Iw);
				avg += 3 * ob[b]->mass + (Iw[0] + Iw[4] + Iw[8]);
				act++;
			}
		avg = act > 0 ? avg / act / 6 : 1;
		std::vector<double> speed(N, 0.0);
		for (int i = 0; i < C; ++i) {
			bool hasStatic = false;
			double rel[3] = {0, 0, 0};
			for (int j = 0; j < 2; ++j) {
				const int id = cs[i].id[j];
				double ve[6] = {ob[id]->v[0], ob[id]->v[1], ob[id]->v[2],
				                ob[id]->w[0], ob[id]->w[1], ob[id]->w[2]};
				if (active(id)) {
					for (int r = 0; r < 3; ++r) ve[r] += g[r] * dt;
				} else {
					hasStatic = true;
				}
				double half[3];
				for (int k = 0; k < 3; ++k) {
					double a = 0;
					for (int r = 0; r < 6; ++r) a += cs[i].T[j][k * 6 + r] * ve[r];
					half[k] = -a;
				}
				for (int k = 0; k < 3; ++k) rel[k] += half[k];
				speed[id] = std::max(speed[id], std::sqrt(dot3(half, half)));
			}
			if (hasStatic)
				for (int j = 0; j < 2; ++j) {
					const int id = cs[i].id[j];
					speed[id] = std::max(speed[id], std::sqrt(dot3(rel, rel)));
				}
		}
		double cs_ = 0, msum = 0;
		for (int b = 0; b < N; ++b)
			if (active(b)) {
				msum += ob[b]->mass;
				cs_ += speed[b] * ob[b]->mass;
			}
		cs_ /= msum;
		const double gn = std::sqrt(dot3(g, g));
		cs_ = std::max(cs_, gn * dt * 0.25);
		cs_ = std::max(cs_, 1e-4);
		mbar = avg;
		s = cs_;
	}

	/* calculateConstraints (NonSmoothContactForce.cpp:105-118, OnlyNormalForce.cpp:40-45) */
	void constraints(const std::vector<double>& f) {
#pragma omp parallel for schedule(static)
		for (int i = 0; i < C; ++i) {
			const double* fi = &f[fd * i];
			if (nsc) {
				cvals[2 * i] = -fi[0];
				cvals[2 * i + 1] = lz.constraint(fi);
				lz.gradient(fi, &grads[2 * i]);
				lz.hessian(fi, &hess[4 * i]);
			} else {
				cvals[i] = -fi[0];
			}
		}
	}

	struct Res {
		std::vector<double> ru, rf, wrf, rl, ac;
	};

	/* calculateResiduals (PrimalDualForceSolver.cpp:132-165) + ACVector */
	void residuals(bool stored, const std::vector<double>& v, const std::vector<double>& vold,
	               const std::vector<double>& f, const std::vector<double>& lam, double muc,
	               Res& R) {
		R.ru.assign(6 * (size_t)N, 0.0);
		R.rf.assign(fd * (size_t)C, 0.0);
		R.wrf.assign(fd * (size_t)C, 0.0);
		R.rl.assign(cd * (size_t)C, 0.0);
		R.ac.assign(fd * (size_t)C, 0.0);
		/* per contact: rf, rl, weight, AC */
#pragma omp parallel for schedule(static)
		for (int i = 0; i < C; ++i) {
			const double* fi = &f[fd * i];
			double u[3];
			to_local(i, v.data(), u);
			if (nsc) {
				double rf[3] = {-u[0], -u[1], -u[2]};
				if (p.spring_based_force) rf[0] -= compliance * mbar * fi[0];
				rf[0] += lam[2 * i];
				double gcur[2];
				if (stored) { gcur[0] = grads[2 * i]; gcur[1] = grads[2 * i + 1]; }
				else lz.gradient(fi, gcur);
				rf[1] -= lam[2 * i + 1] * gcur[0];
				rf[2] -= lam[2 * i + 1] * gcur[1];
				rf[0] -= K[i];
				double c0, c1;
				if (stored) { c0 = cvals[2 * i]; c1 = cvals[2 * i + 1]; }
				else { c0 = -fi[0]; c1 = lz.constraint(fi); }
				R.rl[2 * i] = lam[2 * i] * c0 + muc;
				R.rl[2 * i + 1] = lam[2 * i + 1] * c1 + muc;
				/* tangential weight (NonSmoothContactForce.cpp:74-104) */
				const double tv[2] = {u[1], u[2]};
				const double tsq = tv[0] * tv[0] + tv[1] * tv[1];
				double tn[2] = {tv[0], tv[1]};
				if (tsq > 0) { const double l = std::sqrt(tsq); tn[0] = tv[0] / l; tn[1] = tv[1] / l; }
				double fw = 0;
				for (int j = 0; j < 2; ++j) {
					const int id = cs[i].id[j];
					if (!active(id)) continue;
					double hh[6];
					for (int r = 0; r < 6; ++r) hh[r] = cs[i].T[j][1 * 6 + r] * tn[0] + cs[i].T[j][2 * 6 + r] * tn[1];
					const double* Mi = &Minv[36 * id];
					double q = 0;
					for (int r = 0; r < 6; ++r) {
						double t = 0;
						for (int c = 0; c < 6; ++c) t += Mi[r * 6 + c] * hh[c];
						q += hh[r] * t;
					}
					fw += q * mbar;
				}
				const double rad = lz.radius(fi[0]);
				const double w = std::min(rad / (std::sqrt(tsq) + 1e-8) * fw, 1.0);
				for (int k = 0; k < 3; ++k) R.rf[3 * i + k] = rf[k];
				R.wrf[3 * i] = rf[0];
				R.wrf[3 * i + 1] = rf[1] * w;
				R.wrf[3 * i + 2] = rf[2] * w;
				/* Alart-Curnier (NonSmoothContactForce.cpp:53-73) */
				const double tpn = fi[0] - (u[0] + K[i] + (p.spring_based_force ? 1.0 : 0.0) * compliance * mbar * fi[0]);
				R.ac[3 * i] = std::max(tpn, 0.0) - fi[0];
				const double r = lz.radius(fi[0]);
				const double td[2] = {fi[1] - u[1], fi[2] - u[2]};
				const double tdn = std::sqrt(td[0] * td[0] + td[1] * td[1]);
				double pr[2] = {td[0], td[1]};
				if (!(tdn < r)) { pr[0] = td[0] * r / tdn; pr[1] = td[1] * r / tdn; }
				R.ac[3 * i + 1] = pr[0] - fi[1];
				R.ac[3 * i + 2] = pr[1] - fi[2];
			} else {
				double rf = -u[0];
				if (p.spring_based_force) rf -= compliance * mbar * fi[0];
				rf += lam[i];
				rf -= K[i];
				R.rf[i] = rf;
				R.wrf[i] = rf;
				const double c0 = stored ? cvals[i] : -fi[0];
				R.rl[i] = lam[i] * c0 + muc;
				const double tpn = fi[0] - (u[0] + K[i] + (p.spring_based_force ? 1.0 : 0.0) * compliance * mbar * fi[0]);
				R.ac[i] = std::max(tpn, 0.0) - fi[0];
			}
		}
		/* per body: ru = Ms (v - vold) (static rows 0) + sum_j T_j f over active */
#pragma omp parallel for schedule(static)
		for (int b = 0; b < N; ++b) {
			double* ru = &R.ru[6 * b];
			if (!active(b)) continue;
			const double* M = &Ms[36 * b];
			double d[6];
			for (int r = 0; r < 6; ++r) d[r] = v[6 * b + r] - vold[6 * b + r];
			for (int r = 0; r < 6; ++r) {
				double a = 0;
				for (int c = 0; c < 6; ++c) a += M[r * 6 + c] * d[c];
				ru[r] = a;
			}
			double add[6] = {0, 0, 0, 0, 0, 0};
			for (auto& e : inc[b]) {
				double h[6];
				t_times(e.first, e.second, &f[fd * e.first], h);
				for (int r = 0; r < 6; ++r) add[r] += h[r];
			}
			for (int r = 0; r < 6; ++r) ru[r] += add[r];
		}
	}

	static double sq(const std::vector<double>& a) {
		return fixed_sum(a.size(), [&](size_t i) { return a[i] * a[i]; });
	}

	struct LS {
		std::vector<double> v, f, l;
		double step = 0, sErr = 0, acErr = 0;
		int trials = 0;
	};

	/* filterLineSearch (NonSmoothContactForce.cpp:281-319, OnlyNormalForce.cpp:172-201) */
	void filter(const std::vector<double>& f, const std::vector<double>& lam,
	            std::vector<double>& df, std::vector<double>& dl) {
#pragma omp parallel for schedule(static)
		for (int i = 0; i < C; ++i) {
			if (nsc) {
				const double* cf = &f[3 * i];
				double tf[3] = {cf[0] + df[3 * i], cf[1] + df[3 * i + 1], cf[2] + df[3 * i + 2]};
				if (tf[0] < 0) {
					tf[0] = 0;
					for (int k = 0; k < 3; ++k) df[3 * i + k] = tf[k] - cf[k];
				}
				const double co = lz.constraint(tf);
				if (co > 0) {
					double pf[3];
					lz.project(tf, pf);
					for (int k = 0; k < 3; ++k) df[3 * i + k] = pf[k] - cf[k];
				}
			} else {
				const double tf = f[i] + df[i];
				if (tf < 0) df[i] = 0 - f[i];
			}
		}
#pragma omp parallel for schedule(static)
		for (long i = 0; i < (long)dl.size(); ++i)
			if (dl[i] < 0 && lam[i] + dl[i] < 0) dl[i] = -lam[i];
	}

	/* lineSearch (PrimalDualForceSolver.cpp:562-635) */
	LS line_search(const std::vector<double>& v, const std::vector<double>& f,
	               const std::vector<double>& lam, const std::vector<double>& du,
	               const std::vector<double>& df, const std::vector<double>& dl,
	               const std::vector<double>& vold, double sErr, double sacErr, double muc,
	               int maxLS) {
		LS res;
		res.v = v; res.f = f; res.l = lam;
		res.sErr = sErr; res.acErr = sacErr; res.step = 0;
		double step = 1.0;
		Res R;
		for (int it = 0; it < maxLS; ++it) {
			std::vector<double> duIn(du.size()), dfIn(df.size()), dlIn(dl.size());
			for (size_t k = 0; k < du.size(); ++k) duIn[k] = step * du[k];
			for (size_t k = 0; k < df.size(); ++k) dfIn[k] = step * df[k];
			for (size_t k = 0; k < dl.size(); ++k) dlIn[k] = step * dl[k];
			filter(f, lam, dfIn, dlIn);
			std::vector<double> lI(lam.size()), vI(v.size()), fI(f.size());
			for (size_t k = 0; k < lam.size(); ++k) lI[k] = lam[k] + 0.99 * dlIn[k];
			for (size_t k = 0; k < v.size(); ++k) vI[k] = v[k] + 0.99 * duIn[k];
			for (size_t k = 0; k < f.size(); ++k) fI[k] = f[k] + 0.99 * dfIn[k];
			residuals(false, vI, vold, fI, lI, muc, R);
			res.trials++;
			const double ru2 = sq(R.ru), rl2 = sq(R.rl);
			const double inErr = ru2 + sq(R.wrf) + rl2;
			const double sacIn = ru2 + sq(R.ac) + rl2;
			const double gErr = -fixed_sum(duIn.size(), [&](size_t k) { return duIn[k] * R.ru[k]; }) +
			                    fixed_sum(dfIn.size(), [&](size_t k) { return dfIn[k] * R.rf[k]; }) +
			                    fixed_sum(dlIn.size(), [&](size_t k) { return dlIn[k] / lI[k] * R.rl[k]; });
			const bool acpt = (inErr < sErr) | (sacIn < sacErr);
			if ((acpt && gErr > 0) || it == maxLS - 1) {
				if (acpt) {
					res.v = vI; res.f = fI; res.l = lI;
					res.step = step; res.sErr = inErr; res.acErr = sacIn;
					return res;
				}
				break;
			}
			step *= 0.5;
		}
		return res;
	}

	/* assembly (PrimalDualForceSolver.cpp:322-377; NonSmoothContactForce.cpp:173-253;
	 * OnlyNormalForce.cpp:90-151).  A and rhs are unscaled on return. */
	void assemble(double eps, const std::vector<double>& lam, const Res& R, std::vector<double>& rhs) {
		invC.assign(cd * (size_t)C, 0.0);
		KnInv.assign(C, 0.0);
		KtInv.assign(nsc ? 4 * (size_t)C : 0, 0.0);
		sf.assign(fd * (size_t)C, 0.0);
#pragma omp parallel for schedule(static)
		for (int i = 0; i < C; ++i) {
			if (nsc) {
				const double ic0 = 1. / (cvals[2 * i] - eps * lam[2 * i]);
				const double ic1 = 1. / (cvals[2 * i + 1] - eps * lam[2 * i + 1]);
				invC[2 * i] = ic0; invC[2 * i + 1] = ic1;
				const double* g = &grads[2 * i];
				double lN = -lam[2 * i] * ic0;
				if (p.spring_based_force) lN += compliance * mbar;
				const double lm = -lam[2 * i + 1] * ic1;
				const double L00 = lm * g[0] * g[0], L01 = lm * g[0] * g[1], L11 = lm * g[1] * g[1];
				sf[3 * i] = R.rf[3 * i] + (-R.rl[2 * i] * ic0);
				sf[3 * i + 1] = R.rf[3 * i + 1] + g[0] * R.rl[2 * i + 1] * ic1;
				sf[3 * i + 2] = R.rf[3 * i + 2] + g[1] * R.rl[2 * i + 1] * ic1;
				KnInv[i] = 1. / (eps + lN);
				const double* H = &hess[4 * i];
				const double l1 = lam[2 * i + 1];
				const double k00 = eps + L00 + l1 * H[0];
				const double k01 = 0.0 + L01 + l1 * H[1];
				const double k10 = 0.0 + L01 + l1 * H[2];
				const double k11 = eps + L11 + l1 * H[3];
				const double det = k00 * k11 - k01 * k10;
				KtInv[4 * i + 0] = k11 / det;
				KtInv[4 * i + 1] = -k01 / det;
				KtInv[4 * i + 2] = -k10 / det;
				KtInv[4 * i + 3] = k00 / det;
			} else {
				const double ic = 1. / (cvals[i] - eps);
				invC[i] = ic;
				double reg = -lam[i] * ic;
				if (p.spring_based_force) reg += compliance * mbar;
				sf[i] = R.rf[i] + (-R.rl[i] * ic);
				KnInv[i] = 1. / (eps + reg);
			}
		}
		/* blocks: mass + eps on the diagonal, contact terms gathered per row */
		rhs.assign(6 * (size_t)N, 0.0);
#pragma omp parallel for schedule(dynamic, 16)
		for (int b = 0; b < N; ++b) {
			/* zero row */
			for (int k = A.rowptr[b]; k < A.rowptr[b + 1]; ++k) std::memset(A.val[k].a, 0, sizeof(Blk));
			const int d = A.find(b, b);
			double* D = A.val[d].a;
			if (!active(b)) {
				for (int r = 0; r < 6; ++r) D[r * 6 + r] = 1.0 + eps;
				continue;
			}
			const double* M = &Ms[36 * b];
			for (int r = 0; r < 36; ++r) D[r] = M[r];
			for (int r = 0; r < 6; ++r) D[r * 6 + r] += eps;
			double y[6] = {0, 0, 0, 0, 0, 0};
			for (auto& e : inc[b]) {
				const int i = e.first, j = e.second;
				const Contact& c = cs[i];
				/* kInvSf and y += getTrans(j) * kInvSf = -T_j kInvSf */
				double kf[3] = {sf[fd * i] * KnInv[i], 0, 0};
				if (nsc) {
					const double* Ki = &KtInv[4 * i];
					kf[1] = Ki[0] * sf[3 * i + 1] + Ki[1] * sf[3 * i + 2];
					kf[2] = Ki[2] * sf[3 * i + 1] + Ki[3] * sf[3 * i + 2];
				}
				for (int r = 0; r < 6; ++r) {
					double a = 0;
					for (int k = 0; k < fd; ++k) a += c.T[j][k * 6 + r] * kf[k];
					y[r] -= a;
				}
				for (int k = 0; k < 2; ++k) {
					const int id2 = c.id[k];
					if (!active(id2)) continue;
					const int blk = A.find(b, id2);
					double* Bv = A.val[blk].a;
					/* (T_j e0) KnInv (T_k e0)^T + T_j,t KtInv T_k,t^T */
					for (int r = 0; r < 6; ++r) {
						const double a0 = c.T[j][r] * KnInv[i];
						double at0 = 0, at1 = 0;
						if (nsc) {
							const double t1 = c.T[j][6 + r], t2 = c.T[j][12 + r];
							const double* Ki = &KtInv[4 * i];
							at0 = t1 * Ki[0] + t2 * Ki[2];
							at1 = t1 * Ki[1] + t2 * Ki[3];
						}
						for (int cc = 0; cc < 6; ++cc) {
							double val = a0 * c.T[k][cc];
							if (nsc) val += at0 * c.T[k][6 + cc] + at1 * c.T[k][12 + cc];
							Bv[r * 6 + cc] += val;
						}
					}
				}
			}
			for (int r = 0; r < 6; ++r) rhs[6 * b + r] = y[r] - R.ru[6 * b + r];
		}
		for (int b = 0; b < N; ++b)
			if (!active(b))
				for (int r = 0; r < 6; ++r) rhs[6 * b + r] = 0;
	}

	/* retrieveNonSmoothForceInc (NonSmoothContactForce.cpp:254-280) */
	void increments(const std::vector<double>& lam, const std::vector<double>& du, const Res& R,
	                std::vector<double>& df, std::vector<double>& dl) {
		df.assign(fd * (size_t)C, 0.0);
		dl.assign(cd * (size_t)C, 0.0);
#pragma omp parallel for schedule(static)
		for (int i = 0; i < C; ++i) {
			double u[3];
			to_local(i, du.data(), u);
			if (nsc) {
				double d[3] = {sf[3 * i] - u[0], sf[3 * i + 1] - u[1], sf[3 * i + 2] - u[2]};
				d[0] = KnInv[i] * d[0];
				const double* Ki = &KtInv[4 * i];
				const double t0 = Ki[0] * d[1] + Ki[1] * d[2];
				const double t1 = Ki[2] * d[1] + Ki[3] * d[2];
				d[1] = t0; d[2] = t1;
				for (int k = 0; k < 3; ++k) df[3 * i + k] = d[k];
				dl[2 * i] = -invC[2 * i] * (-lam[2 * i] * d[0] + R.rl[2 * i]);
				const double gd = grads[2 * i] * d[1] + grads[2 * i + 1] * d[2];
				dl[2 * i + 1] = -invC[2 * i + 1] * (lam[2 * i + 1] * gd + R.rl[2 * i + 1]);
			} else {
				const double d = (sf[i] - u[0]) * KnInv[i];
				df[i] = d;
				dl[i] = -invC[i] * (-lam[i] * d + R.rl[i]);
			}
		}
	}

	void run(double dt, const double* g, pdc_solve_stats* st) {
		const size_t nR = 6 * (size_t)N, nF = fd * (size_t)C, nL = cd * (size_t)C;
		const size_t nU = nR + nF + nL;
		set_threads(scene->threads > 0 ? scene->threads : (nU < 1000 ? 1 : max_threads()));
		stat.assign(N, 0);
		for (int b = 0; b < N; ++b) stat[b] = ob[b]->stat ? 1 : 0;
		std::vector<double> v(nR), vold(nR);
		for (int b = 0; b < N; ++b) {
			for (int r = 0; r < 3; ++r) {
				v[6 * b + r] = ob[b]->v[r];
				v[6 * b + 3 + r] = ob[b]->w[r];
			}
			for (int r = 0; r < 6; ++r) vold[6 * b + r] = v[6 * b + r];
			if (active(b))
				for (int r = 0; r < 3; ++r) vold[6 * b + r] += g[r] * dt;
		}
		nondim(dt, g);
		for (auto& x : v) x /= s;
		for (auto& x : vold) x /= s;
		/* mass matrices */
		Ms.assign(36 * (size_t)N, 0.0);
		Minv.assign(36 * (size_t)N, 0.0);
		for (int b = 0; b < N; ++b) {
			double R[9], Iw[9], Ii[9];
			quat_rot(ob[b]->q, R);
			rot_diag_rot(R, ob[b]->I0, Iw);
			const double inv0[3] = {1.0 / ob[b]->I0[0], 1.0 / ob[b]->I0[1], 1.0 / ob[b]->I0[2]};
			rot_diag_rot(R, inv0, Ii);
			double* M = &Ms[36 * b];
			double* Mi = &Minv[36 * b];
			const double im = 1.0 / ob[b]->mass;
			for (int r = 0; r < 3; ++r) {
				Mi[r * 6 + r] = im;
				for (int c = 0; c < 3; ++c) Mi[(3 + r) * 6 + 3 + c] = Ii[r * 3 + c];
			}
			if (!active(b)) {
				for (int r = 0; r < 6; ++r) M[r * 6 + r] = 1.0;
			} else {
				for (int r = 0; r < 3; ++r) {
					M[r * 6 + r] = ob[b]->mass / mbar;
					for (int c = 0; c < 3; ++c) M[(3 + r) * 6 + 3 + c] = Iw[r * 3 + c] / mbar;
				}
			}
		}
		/* incidence and block pattern (linsysReserve, NonSmoothForceBase.h:151-187) */
		inc.assign(N, {});
		for (int i = 0; i < C; ++i)
			for (int j = 0; j < 2; ++j) inc[cs[i].id[j]].emplace_back(i, j);
		std::vector<std::vector<int>> nbr(N);
		for (int b = 0; b < N; ++b) nbr[b].push_back(b);
		for (int i = 0; i < C; ++i) {
			const int a = cs[i].id[0], b = cs[i].id[1];
			if (active(a) && active(b)) {
				nbr[a].push_back(b);
				nbr[b].push_back(a);
			}
		}
		A.n = N;
		A.rowptr.assign(N + 1, 0);
		A.col.clear();
		for (int b = 0; b < N; ++b) {
			auto& l = nbr[b];
			std::sort(l.begin(), l.end());
			l.erase(std::unique(l.begin(), l.end()), l.end());
			for (int c : l) A.col.push_back(c);
			A.rowptr[b + 1] = (int)A.col.size();
		}
		A.val.assign(A.col.size(), Blk());
		/* preprocess (NonSmoothContactForce.cpp:8-48) */
		const double damp = p.spring_damp * p.spring_k;
		const double tk = dt * p.spring_k;
		compliance = 1. / (dt * (tk + damp));
		reduction = tk / (tk + damp);
		K.assign(C, 0.0);
		for (int i = 0; i < C; ++i) {
			if (p.spring_based_force) {
				K[i] += -reduction * cs[i].depth / dt / s;
			} else {
				double u[3];
				to_local(i, v.data(), u);
				K[i] += u[0] * p.restitution;
			}
		}
		/* initForces */
		std::vector<double> f(nF, 0.0), lam(nL, 1e-2);
		for (int i = 0; i < C; ++i) f[fd * i] = 1.0;
		cvals.assign(nL, 0.0);
		grads.assign(nsc ? 2 * (size_t)C : 0, 0.0);
		hess.assign(nsc ? 4 * (size_t)C : 0, 0.0);

		double Hu = 1e-4, exitErr = 0;
		int it = 0, pcgTotal = 0, pcgHits = 0, gdCount = 0, trials = 0;
		const bool direct = scene->solver == ORC_SOLVER_DIRECT ||
		                    (scene->solver == ORC_SOLVER_REFERENCE && N < 2000);
		BlockCholesky chol;
		bool analyzed = false;
		/* the standard PCG form's groups, fixed for the timestep */
		std::vector<int> grp;
		if (!direct && p.pcg_tol < kStandardFormTol) grp = contact_groups();
		scene->trace.clear();
		Res R;
		std::vector<double> rhs, du(nR, 0.0), df, dl, Dsc(nR), z(nR);
		for (it = 0; it < p.iterations; ++it) {
			constraints(f);
			const double gap = -fixed_sum(nL, [&](size_t k) { return lam[k] * cvals[k]; }) / (double)nL;
			const double muc = gap * 0.1;
			residuals(true, v, vold, f, lam, muc, R);
			const double ru2 = sq(R.ru), wrf2 = sq(R.wrf), rl2 = sq(R.rl), ac2 = sq(R.ac);
			const double err = ru2 + wrf2 + rl2;
			const double sErr = err;
			const double sacErr = ru2 + ac2 + rl2;
			const double eps = Hu * std::pow(err, 0.66667 * 0.5);
			const double momenErr = std::sqrt(ru2) / (double)nR;
			exitErr = std::max(momenErr, std::sqrt(wrf2) / (double)nF);
			exitErr = std::max(exitErr, std::sqrt(rl2) / (double)nL);
			exitErr = std::max(exitErr, gap);
			exitErr = std::min(exitErr, std::max(momenErr, std::sqrt(ac2) / (double)C));
			double tr[8] = {err, exitErr, eps, Hu, 0, 0, gap, sacErr};
			if (exitErr < p.tolerance) {
				scene->trace.insert(scene->trace.end(), tr, tr + 8);
				break;
			}
			assemble(eps, lam, R, rhs);
			if (scene->capture && it == 0) {
				const size_t n6 = nR;
				scene->capA.assign(n6 * n6, 0.0);
				for (int b = 0; b < N; ++b)
					for (int k = A.rowptr[b]; k < A.rowptr[b + 1]; ++k)
						for (int r = 0; r < 6; ++r)
							for (int c = 0; c < 6; ++c)
								scene->capA[(6 * b + r) * n6 + 6 * A.col[k] + c] = A.val[k].a[r * 6 + c];
				scene->capRhs = rhs;
			}
			/* Jacobi scaling D = diag(A)^-1/2 */
			for (int b = 0; b < N; ++b) {
				const double* Dg = A.val[A.find(b, b)].a;
				for (int r = 0; r < 6; ++r) Dsc[6 * b + r] = 1 / std::sqrt(Dg[r * 6 + r]);
			}
			for (int b = 0; b < N; ++b)
				for (int k = A.rowptr[b]; k < A.rowptr[b + 1]; ++k) {
					const int c = A.col[k];
					for (int r = 0; r < 6; ++r)
						for (int cc = 0; cc < 6; ++cc) A.val[k].a[r * 6 + cc] *= Dsc[6 * b + r] * Dsc[6 * c + cc];
				}
			for (size_t k = 0; k < nR; ++k) rhs[k] *= Dsc[k];
			int iters = 0;
			if (direct) {
				if (!analyzed) { chol.analyze(A); analyzed = true; }
				if (chol.factorize(A)) chol.solve(rhs.data(), z.data());
				else std::fill(z.begin(), z.end(), 0.0);
			} else {
				/* initial guess: the previous iteration's du (the reference hands
				 * AMGCL X = du, the unscaled increment, or -ru after a gradient-
				 * descent fallback; :407-414, :451); zero in the first iteration,
				 * where the reference's du is freshly allocated */
				if (it > 0) z = du;
				iters = bcsr_pcg(A, rhs.data(), p.pcg_tol, p.pcg_max_iter, z.data(), it > 0,
				                 grp.empty() ? nullptr : grp.data());
				pcgTotal += iters;
				if (iters == p.pcg_max_iter) pcgHits++;
			}
			for (size_t k = 0; k < nR; ++k) du[k] = z[k] * Dsc[k];
			increments(lam, du, R, df, dl);
			LS ls = line_search(v, f, lam, du, df, dl, vold, sErr, sacErr, muc, 10);
			trials += ls.trials;
			bool useGD = false;
			if (ls.step == 0) {
				useGD = true;
				gdCount++;
				for (size_t k = 0; k < nR; ++k) du[k] = -R.ru[k];
				ls = line_search(v, f, lam, du, R.rf, R.rl, vold, sErr, sacErr, muc, 20);
				trials += ls.trials;
			}
			v = ls.v; f = ls.f; lam = ls.l;
			tr[4] = ls.step;
			tr[5] = iters;
			scene->trace.insert(scene->trace.end(), tr, tr + 8);
			if (!useGD && ls.step == 1) {
				/* x1.3 after a capped iterative solve: the reference solves
				 * iteratively only at 2000 or more bodies in contact; below that
				 * its LDLT branch leaves iters at 0 (:379-394, :473-479) */
				if (iters == p.pcg_max_iter && !direct && N >= 2000) Hu *= 1.3;
				else Hu /= 4;
			} else {
				Hu *= 4;
			}
			Hu = std::max(Hu, 1e-6);
		}
		/* write back (PrimalDualForceSolver.cpp:491-510) */
		for (int b = 0; b < N; ++b) {
			for (int r = 0; r < 3; ++r) {
				ob[b]->acc[r] = v[6 * b + r] * s - ob[b]->v[r];
				ob[b]->aacc[r] = v[6 * b + 3 + r] * s - ob[b]->w[r];
			}
		}
		for (int i = 0; i < C; ++i) {
			for (int k = 0; k < 3; ++k) cs[i].lf[k] = 0;
			for (int k = 0; k < fd; ++k) cs[i].lf[k] = f[fd * i + k] * s * mbar;
		}
		if (st) {
			st->n_contacts = C;
			st->n_local_bodies = N;
			st->newton_iterations = it;
			st->pcg_iterations = pcgTotal;
			st->pcg_maxiter_hits = pcgHits;
			st->gd_fallbacks = gdCount;
			st->line_search_trials = trials;
			st->exit_error = exitErr;
			st->char_mass = mbar;
			st->char_speed = s;
			st->n_blocks = (int)A.col.size();
		}
	}
};

/* fillForces (PrimalDualForceSolver.cpp:79-130) */
void fill_forces(orc_scene* sc, std::vector<Contact>& contacts, double dt, const double* g,
                 pdc_solve_stats* st) {
	for (Body& b : sc->bodies) {
		for (int r = 0; r < 3; ++r) { b.acc[r] = g[r] * dt; b.aacc[r] = 0; }
	}
	if (st) std::memset(st, 0, sizeof(*st));
	if (contacts.empty()) return;
	std::unordered_map<int, int> g2l;
	std::vector<int> l2g;
	for (const Contact& c : contacts)
		for (int j = 0; j < 2; ++j)
			if (!g2l.count(c.id[j])) {
				const int lid = (int)g2l.size();
				g2l[c.id[j]] = lid;
				l2g.push_back(c.id[j]);
			}
	std::vector<Body*> ob(l2g.size());
	for (size_t k = 0; k < l2g.size(); ++k) ob[k] = &sc->bodies[l2g[k]];
	std::vector<Contact> local = contacts;
	for (Contact& c : local)
		for (int j = 0; j < 2; ++j) c.id[j] = g2l[c.id[j]];
	Solver solver(sc->p, local, ob, sc);
	solver.run(dt, g, st);
	for (size_t i = 0; i < contacts.size(); ++i)
		std::memcpy(contacts[i].lf, local[i].lf, sizeof(contacts[i].lf));
	sc->capL2G = l2g;
}

/* DynamicSystem.cpp:137-161 */
void integrate(orc_scene* sc, double dt) {
	for (Body& b : sc->bodies) {
		if (!b.stat) {
			for (int r = 0; r < 3; ++r) {
				b.w[r] += b.aacc[r];
				b.v[r] += b.acc[r];
			}
		}
		const double ri[3] = {b.w[0] * dt, b.w[1] * dt, b.w[2] * dt};
		const double rsq = dot3(ri, ri);
		double qi[4] = {0, 0, 0, 1};
		if (rsq > 0) {
			const double ang = std::sqrt(rsq);
			const double ax[3] = {ri[0] / ang, ri[1] / ang, ri[2] / ang};
			const double sh = std::sin(0.5 * ang);
			qi[0] = sh * ax[0]; qi[1] = sh * ax[1]; qi[2] = sh * ax[2];
			qi[3] = std::cos(0.5 * ang);
		}
		for (int r = 0; r < 3; ++r) b.x[r] += b.v[r] * dt;
		quat_mul(qi, b.q, b.q);
	}
}

void load_contacts(int C, const int32_t* ids, const double* T, const double* depth,
                   std::vector<Contact>& out) {
	out.resize(C);
	for (int i = 0; i < C; ++i) {
		Contact& c = out[i];
		c.id[0] = ids[2 * i];
		c.id[1] = ids[2 * i + 1];
		c.depth = depth ? depth[i] : 0.0;
		c.pos[0] = c.pos[1] = c.pos[2] = 0;
		std::memcpy(c.T[0], T + 36 * (size_t)i, 18 * sizeof(double));
		std::memcpy(c.T[1], T + 36 * (size_t)i + 18, 18 * sizeof(double));
		c.lf[0] = c.lf[1] = c.lf[2] = 0;
	}
}

void apply_state(orc_scene* s, const double* q, const double* v, const double* w) {
	for (size_t b = 0; b < s->bodies.size(); ++b) {
		Body& B = s->bodies[b];
		if (q) std::memcpy(B.q, q + 4 * b, 4 * sizeof(double));
		if (v) std::memcpy(B.v, v + 3 * b, 3 * sizeof(double));
		if (w) std::memcpy(B.w, w + 3 * b, 3 * sizeof(double));
	}
}

}  // namespace

/* ------------------------------------------------------------------------ */
/* C entry points                                                            */
/* ------------------------------------------------------------------------ */
extern "C" {

orc_scene* orc_create(const pdc_params* p, int32_t n, const int32_t* shape_type,
                      const double* dims, const double* mass, const double* inertia_diag,
                      const int32_t* is_static, const int32_t* emit_samples) {
	return orc_create_meshes(p, n, shape_type, dims, mass, inertia_diag, is_static, emit_samples, 0, nullptr,
	                         nullptr, nullptr, nullptr);
}

orc_scene* orc_create_meshes(const pdc_params* p, int32_t n, const int32_t* shape_type,
                             const double* dims, const double* mass, const double* inertia_diag,
                             const int32_t* is_static, const int32_t* emit_samples, int32_t n_meshes,
                             const int32_t* mesh_nv, const double* mesh_v, const int32_t* mesh_nt,
                             const int32_t* mesh_tri) {
	if (!p || n < 0 || n_meshes < 0) return nullptr;
	orc_scene* s = new orc_scene();
	s->p = *p;
	std::memset(&s->last, 0, sizeof(s->last));
	size_t ov = 0, ot = 0;
	for (int m = 0; m < n_meshes; ++m) {
		auto M = std::make_unique<OMesh>();
		omesh_setup(*M, mesh_nv[m], mesh_v + 3 * ov, mesh_nt[m], mesh_tri + 3 * ot);
		ov += (size_t)mesh_nv[m];
		ot += (size_t)mesh_nt[m];
		/* node grid */
		Body tmp;
		tmp.type = PDC_SHAPE_MESH;
		tmp.mesh = M.get();
		setup_grid(tmp, p->grid_n, p->grid_pad_frac);
		const int gn = p->grid_n;
		M->gn = gn;
		M->pad = p->grid_pad_frac;
		M->grid.assign((size_t)gn * gn * gn, 0.0f);
		const int64_t n3 = (int64_t)gn * gn * gn;
#pragma omp parallel for schedule(static)
		for (int64_t t = 0; t < n3; ++t) {
			const int i = (int)(t % gn), j = (int)((t / gn) % gn), k = (int)(t / ((int64_t)gn * gn));
			const double q[3] = {tmp.glo[0] + (double)i * tmp.gsp[0], tmp.glo[1] + (double)j * tmp.gsp[1],
			                     tmp.glo[2] + (double)k * tmp.gsp[2]};
			M->grid[(size_t)t] = (float)mesh_sdf(*M, q);
		}
		s->meshes.push_back(std::move(M));
	}
	s->bodies.resize(n);
	for (int i = 0; i < n; ++i) {
		Body& b = s->bodies[i];
		b.type = shape_type[i];
		for (int a = 0; a < 3; ++a) b.dims[a] = dims[3 * i + a];
		if (b.type == PDC_SHAPE_MESH) {
			const int id = (int)dims[3 * i];
			if (id < 0 || id >= n_meshes) {
				delete s;
				return nullptr;
			}
			b.mesh = s->meshes[id].get();
		}
		b.mass = mass[i];
		for (int a = 0; a < 3; ++a) b.I0[a] = inertia_diag[3 * i + a];
		b.stat = is_static[i] != 0;
		b.emit = emit_samples ? emit_samples[i] != 0 : true;
		setup_grid(b, p->grid_n, p->grid_pad_frac);
		setup_samples(b, p->sample_spacing);
	}
	return s;
}
void orc_destroy(orc_scene* s) { delete s; }
void orc_set_params(orc_scene* s, const pdc_params* p) { s->p = *p; }
void orc_set_threads(orc_scene* s, int32_t t) { s->threads = t; }
void orc_set_linear_solver(orc_scene* s, int32_t w) { s->solver = w; }
void orc_set_state(orc_scene* s, const double* pos, const double* quat, const double* vel,
                   const double* angvel) {
	for (size_t b = 0; b < s->bodies.size(); ++b) {
		Body& B = s->bodies[b];
		if (pos) std::memcpy(B.x, pos + 3 * b, 3 * sizeof(double));
	}
	apply_state(s, quat, vel, angvel);
}
void orc_get_state(const orc_scene* s, double* pos, double* quat, double* vel, double* angvel) {
	for (size_t b = 0; b < s->bodies.size(); ++b) {
		const Body& B = s->bodies[b];
		if (pos) std::memcpy(pos + 3 * b, B.x, 3 * sizeof(double));
		if (quat) std::memcpy(quat + 4 * b, B.q, 4 * sizeof(double));
		if (vel) std::memcpy(vel + 3 * b, B.v, 3 * sizeof(double));
		if (angvel) std::memcpy(angvel + 3 * b, B.w, 3 * sizeof(double));
	}
}

int32_t orc_detect_contacts(orc_scene* s, int32_t* n_ties) {
	detect(s);
	if (n_ties) *n_ties = s->ties;
	return (int32_t)s->contacts.size();
}
int32_t orc_raw_contact_count(const orc_scene* s) { return s->raw_count; }

int32_t orc_get_contacts(const orc_scene* s, int32_t cap, int32_t* ids, double* pos,
                         double* depth, double* transformation, double* local_force) {
	const int C = (int)s->contacts.size();
	if (cap < C) return -1;
	for (int i = 0; i < C; ++i) {
		const Contact& c = s->contacts[i];
		if (ids) { ids[2 * i] = c.id[0]; ids[2 * i + 1] = c.id[1]; }
		if (pos) std::memcpy(pos + 3 * i, c.pos, 3 * sizeof(double));
		if (depth) depth[i] = c.depth;
		if (transformation) {
			std::memcpy(transformation + 36 * (size_t)i, c.T[0], 18 * sizeof(double));
			std::memcpy(transformation + 36 * (size_t)i + 18, c.T[1], 18 * sizeof(double));
		}
		if (local_force) std::memcpy(local_force + 3 * i, c.lf, 3 * sizeof(double));
	}
	return C;
}

int32_t orc_fill_forces(orc_scene* s, int32_t C, const int32_t* ids, const double* T,
                        const double* depth, const double* quat, const double* vel,
                        const double* angvel, double* out_acc, double* out_lf,
                        pdc_solve_stats* st) {
	std::vector<Contact> cs;
	load_contacts(C, ids, T, depth, cs);
	for (int i = 0; i < C; ++i)
		for (int j = 0; j < 2; ++j)
			if (cs[i].id[j] < 0 || cs[i].id[j] >= (int)s->bodies.size()) return PDC_ERR_INVALID_ARGUMENT;
	apply_state(s, quat, vel, angvel);
	fill_forces(s, cs, s->p.timestep, s->p.gravity, st);
	for (size_t b = 0; b < s->bodies.size(); ++b) {
		if (out_acc) {
			std::memcpy(out_acc + 6 * b, s->bodies[b].acc, 3 * sizeof(double));
			std::memcpy(out_acc + 6 * b + 3, s->bodies[b].aacc, 3 * sizeof(double));
		}
	}
	if (out_lf)
		for (int i = 0; i < C; ++i) std::memcpy(out_lf + 3 * i, cs[i].lf, 3 * sizeof(double));
	return PDC_OK;
}

int32_t orc_step(orc_scene* s, int32_t n_steps, pdc_solve_stats* st) {
	for (int k = 0; k < n_steps; ++k) {
		detect(s);
		pdc_solve_stats local;
		fill_forces(s, s->contacts, s->p.timestep, s->p.gravity, &local);
		local.ties = s->ties;
		s->last = local;
		integrate(s, s->p.timestep);
	}
	if (st) *st = s->last;
	return PDC_OK;
}

int32_t orc_trace(const orc_scene* s, int32_t cap, double* out) {
	const int n = (int)(s->trace.size() / 8);
	const int m = std::min(n, cap);
	std::memcpy(out, s->trace.data(), 8 * (size_t)m * sizeof(double));
	return m;
}

void orc_lorentz(double mu, const double f[3], double* c, double g[2], double H[4], double* radius) {
	Lorentz lz;
	lz.mu = mu;
	if (c) *c = lz.constraint(f);
	if (g) lz.gradient(f, g);
	if (H) lz.hessian(f, H);
	if (radius) *radius = lz.radius(f[0]);
}
void orc_lorentz_project(double mu, const double in[3], double out[3]) {
	Lorentz lz;
	lz.mu = mu;
	lz.project(in, out);
}
void orc_normal_to_transform(const double pos[3], const double n[3], const double x0[3],
                             const double x1[3], double T[36]) {
	normal_to_transform(pos, n, x0, x1, T, T + 18);
}
int32_t orc_sdf_query(const orc_scene* s, int32_t b, const double q[3], double* sd, double grad[3]) {
	return grid_query(s->bodies[b], q, sd, grad) ? 1 : 0;
}
float orc_sdf_node(const orc_scene* s, int32_t b, int32_t i, int32_t j, int32_t k) {
	return grid_node(s->bodies[b], i, j, k);
}
void orc_mesh_sdf(int32_t nv, const double* v, int32_t nt, const int32_t* tri, int32_t n, const double* pts,
                  double* out) {
	OMesh m;
	omesh_setup(m, nv, v, nt, tri);
	for (int32_t i = 0; i < n; ++i) out[i] = mesh_sdf(m, pts + 3 * i);
}
int32_t orc_samples(const orc_scene* s, int32_t b, int32_t cap, double* out) {
	const auto& v = s->bodies[b].samples;
	const int n = (int)(v.size() / 3);
	if (out) std::memcpy(out, v.data(), 3 * (size_t)std::min(n, cap) * sizeof(double));
	return n;
}
void orc_ellipsoid_sdf(const double dims[3], int32_t n, const double* pts, double* out) {
	for (int32_t i = 0; i < n; ++i) out[i] = ellipsoid_sdf(dims, pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
}

int32_t orc_first_system(orc_scene* s, int32_t C, const int32_t* ids, const double* T,
                         const double* depth, const double* quat, const double* vel,
                         const double* angvel, int32_t cap, double* Aout, double* rhs,
                         int32_t* l2g) {
	pdc_params saved = s->p;
	s->capture = true;
	s->capA.clear();
	s->p.iterations = 1;
	std::vector<Contact> cs;
	load_contacts(C, ids, T, depth, cs);
	apply_state(s, quat, vel, angvel);
	fill_forces(s, cs, s->p.timestep, s->p.gravity, nullptr);
	s->capture = false;
	s->p = saved;
	const int n6 = (int)s->capRhs.size();
	if (s->capA.empty() || cap < n6) return -1;
	std::memcpy(Aout, s->capA.data(), (size_t)n6 * n6 * sizeof(double));
	std::memcpy(rhs, s->capRhs.data(), (size_t)n6 * sizeof(double));
	if (l2g)
		for (size_t k = 0; k < s->capL2G.size(); ++k) l2g[k] = s->capL2G[k];
	return n6;
}

static void dense_to_bcsr(int n, const double* A, BCSR& B) {
	const int nb = n / 6;
	B.n = nb;
	B.rowptr.assign(nb + 1, 0);
	B.col.clear();
	B.val.clear();
	for (int r = 0; r < nb; ++r) {
		for (int c = 0; c < nb; ++c) {
			bool nz = false;
			for (int i = 0; i < 6 && !nz; ++i)
				for (int j = 0; j < 6; ++j)
					if (A[(6 * r + i) * (size_t)n + 6 * c + j] != 0.0) { nz = true; break; }
			if (!nz && r != c) continue;
			Blk b;
			for (int i = 0; i < 6; ++i)
				for (int j = 0; j < 6; ++j) b.a[i * 6 + j] = A[(6 * r + i) * (size_t)n + 6 * c + j];
			B.col.push_back(c);
			B.val.push_back(b);
		}
		B.rowptr[r + 1] = (int)B.col.size();
	}
}

int32_t orc_block_cholesky_solve(int32_t n, const double* A, const double* b, double* x) {
	if (n % 6) return -1;
	BCSR B;
	dense_to_bcsr(n, A, B);
	BlockCholesky ch;
	ch.analyze(B);
	if (!ch.factorize(B)) return 1;
	ch.solve(b, x);
	return 0;
}
int32_t orc_pcg_solve(int32_t n, const double* A, const double* b, double tol, int32_t maxit,
                      double* x, int32_t warm, const int32_t* rep) {
	if (n % 6) return -1;
	BCSR B;
	dense_to_bcsr(n, A, B);
	if (rep)
		for (int r = 0; r < n / 6; ++r)
			if (rep[r] < 0 || rep[r] >= n / 6) return -1;
	return bcsr_pcg(B, b, tol, maxit, x, warm != 0, rep);
}

int32_t orc_group_bodies(int32_t n, const int64_t* key, const int32_t* stat, const int64_t* w, int32_t np,
                         const int32_t* pa, const int32_t* pb, const int32_t* pn, int32_t* rep) {
	if (n < 0 || np < 0) return -1;
	std::vector<long long> k(key, key + n), ww(w, w + n);
	std::vector<char> st(n);
	for (int r = 0; r < n; ++r) st[r] = stat[r] ? 1 : 0;
	std::vector<std::array<int, 3>> pairs;
	for (int q = 0; q < np; ++q) {
		if (pa[q] < 0 || pa[q] >= n || pb[q] < 0 || pb[q] >= n || pn[q] < 1) return -1;
		pairs.push_back({pa[q], pb[q], pn[q]});
	}
	const std::vector<int> g = group_bodies(n, k, st, ww, pairs);
	for (int r = 0; r < n; ++r) rep[r] = g[r];
	return 0;
}

}  // extern "C"
