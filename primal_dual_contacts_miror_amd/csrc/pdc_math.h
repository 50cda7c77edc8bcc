// pdc_math.h — device/host math shared by the HIP kernels.
//
// The grid-SDF functions define this project's contact generator and must give
// bit-identical FP64 results on host and device: the library is compiled with
// -ffp-contract=off and every expression below fixes its evaluation order.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#define PDC_HD __host__ __device__ __forceinline__

namespace pdc {

PDC_HD double dot3(const double* a, const double* b) {
	return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
PDC_HD void cross3(const double* a, const double* b, double* o) {
	o[0] = a[1] * b[2] - a[2] * b[1];
	o[1] = a[2] * b[0] - a[0] * b[2];
	o[2] = a[0] * b[1] - a[1] * b[0];
}

// Rotation (row-major) of a unit quaternion stored (x, y, z, w); the standard
// formula used by Eigen's Quaternion::toRotationMatrix.
PDC_HD void quat_rot(const double* q, double* R) {
	const double x = q[0], y = q[1], z = q[2], w = q[3];
	const double tx = 2 * x, ty = 2 * y, tz = 2 * z;
	const double twx = tx * w, twy = ty * w, twz = tz * w;
	const double txx = tx * x, txy = ty * x, txz = tz * x;
	const double tyy = ty * y, tyz = tz * y, tzz = tz * z;
	R[0] = 1 - (tyy + tzz); R[1] = txy - twz;       R[2] = txz + twy;
	R[3] = txy + twz;       R[4] = 1 - (txx + tzz); R[5] = tyz - twx;
	R[6] = txz - twy;       R[7] = tyz + twx;       R[8] = 1 - (txx + tyy);
}

// Hamilton product o = a * b, quaternions stored (x, y, z, w).
PDC_HD void quat_mul(const double* a, const double* b, double* o) {
	double r[4];
	r[3] = a[3] * b[3] - a[0] * b[0] - a[1] * b[1] - a[2] * b[2];
	r[0] = a[3] * b[0] + a[0] * b[3] + a[1] * b[2] - a[2] * b[1];
	r[1] = a[3] * b[1] + a[1] * b[3] + a[2] * b[0] - a[0] * b[2];
	r[2] = a[3] * b[2] + a[2] * b[3] + a[0] * b[1] - a[1] * b[0];
	o[0] = r[0]; o[1] = r[1]; o[2] = r[2]; o[3] = r[3];
}

// o = R diag(d) R^T, stored as the 6 unique entries (00 01 02 11 12 22).
PDC_HD void rot_diag_rot_sym(const double* R, const double* d, double* o) {
	int k = 0;
	for (int r = 0; r < 3; ++r)
		for (int c = r; c < 3; ++c)
			o[k++] = R[r * 3 + 0] * d[0] * R[c * 3 + 0] + R[r * 3 + 1] * d[1] * R[c * 3 + 1] +
			         R[r * 3 + 2] * d[2] * R[c * 3 + 2];
}
PDC_HD void sym3_mv(const double* s, const double* x, double* y) {
	y[0] = s[0] * x[0] + s[1] * x[1] + s[2] * x[2];
	y[1] = s[1] * x[0] + s[3] * x[1] + s[4] * x[2];
	y[2] = s[2] * x[0] + s[4] * x[1] + s[5] * x[2];
}

// ---------------------------------------------------------------------------
// Lorentz-circle constraint (reference LorentzCircleConstraint.cpp:5-55):
//   c  = sqrt(|f_t|^2 + P) - mu f_n - sqrt(P) - SH,  P = 1e-16, SH = 1e-8
//   g  = f_t / sqrt(|f_t|^2 + P)
//   H  = (z z^T + p I) (|f_t|^2 + p)^(-3/2), z = (f_2, -f_1), p = max(P, 1e-6)
//   radius(f_n) = sqrt(k^2 - P), k = mu f_n + sqrt(P) + SH
// ---------------------------------------------------------------------------
struct Lorentz {
	static constexpr double P = 1e-16;
	static constexpr double SQP = 1e-8;  // sqrt(P), exactly representable as computed
	static constexpr double SH = 1e-8;
	PDC_HD static double constraint(double mu, double f0, double f1, double f2) {
		const double r = f1 * f1 + f2 * f2 + P;
		return sqrt(r) - mu * f0 - sqrt(P) - SH;
	}
	PDC_HD static void gradient(double f1, double f2, double* g) {
		const double r = sqrt(f1 * f1 + f2 * f2 + P);
		g[0] = f1 / r;
		g[1] = f2 / r;
	}
	// symmetric 2x2: H00, H01, H11
	PDC_HD static void hessian(double f1, double f2, double* H) {
		const double p = 1e-6;
		const double r2 = f1 * f1 + f2 * f2 + p;
		const double r3 = 1.0 / (r2 * sqrt(r2));
		const double z0 = f2, z1 = -f1;
		H[0] = (z0 * z0 + p) * r3;
		H[1] = (z0 * z1) * r3;
		H[2] = (z1 * z1 + p) * r3;
	}
	PDC_HD static double radius(double mu, double f0) {
		const double k = mu * f0 + sqrt(P) + SH;
		return sqrt(k * k - P);
	}
	// project(in, start = in, target = 0) (LorentzCircleConstraint.cpp:36-49)
	PDC_HD static void project(double mu, const double* in, double* out) {
		out[0] = in[0]; out[1] = in[1]; out[2] = in[2];
		if (in[0] < 0) out[0] = 0;
		if (constraint(mu, out[0], out[1], out[2]) > 0) {
			double r = radius(mu, in[0]);
			r = r > 0.0 ? r : 0.0;
			const double nt = sqrt(out[1] * out[1] + out[2] * out[2]);
			const double s = r / nt;
			out[1] *= s;
			out[2] *= s;
		}
	}
};

// ---------------------------------------------------------------------------
// Contact frame (reference ContactGeneratorBase.h:21-49): rot maps x^ to n by
// the angle-axis rotation (atan2(|x^ x n|, x^.n), normalize(x^ x n)); T_j =
// [rot_j; c_j x rot_j(:,k)], c_j = pos - x_j, rot_1 = -rot_0.  Output: the
// rotation (row-major) of side 0.  n = -x^ gives rot = -I.
// ---------------------------------------------------------------------------
PDC_HD void frame_rotation(const double* n, double* R) {
	const double axis[3] = {0.0, -n[2], n[1]};
	const double sq = dot3(axis, axis);
	const double an = sqrt(sq);
	const double angle = atan2(an, n[0]);
	double a[3] = {axis[0], axis[1], axis[2]};
	if (sq > 0) {
		const double inv = sqrt(sq);
		a[0] = axis[0] / inv; a[1] = axis[1] / inv; a[2] = axis[2] / inv;
	}
	const double s = sin(angle), c = cos(angle);
	const double sa[3] = {s * a[0], s * a[1], s * a[2]};
	const double c1 = 1 - c;
	const double ca[3] = {c1 * a[0], c1 * a[1], c1 * a[2]};
	R[0] = ca[0] * a[0] + c;
	R[4] = ca[1] * a[1] + c;
	R[8] = ca[2] * a[2] + c;
	double t = ca[0] * a[1];
	R[1] = t - sa[2]; R[3] = t + sa[2];
	t = ca[0] * a[2];
	R[2] = t + sa[1]; R[6] = t - sa[1];
	t = ca[1] * a[2];
	R[5] = t - sa[0]; R[7] = t + sa[0];
}
// T (6x3 column-major, T[col*6 + row]) of side j from the side-0 rotation.
PDC_HD void frame_T(const double* R0, const double* pos, const double* xj, int j, double* T) {
	const double cj[3] = {pos[0] - xj[0], pos[1] - xj[1], pos[2] - xj[2]};
	const double sg = j ? -1.0 : 1.0;
	for (int col = 0; col < 3; ++col) {
		const double rc[3] = {sg * R0[0 * 3 + col], sg * R0[1 * 3 + col], sg * R0[2 * 3 + col]};
		double cr[3];
		cross3(cj, rc, cr);
		T[col * 6 + 0] = rc[0]; T[col * 6 + 1] = rc[1]; T[col * 6 + 2] = rc[2];
		T[col * 6 + 3] = cr[0]; T[col * 6 + 4] = cr[1]; T[col * 6 + 5] = cr[2];
	}
}

// ---------------------------------------------------------------------------
// Euclidean distance from a point to an ellipsoid (D. Eberly, "Distance from
// a Point to an Ellipse, an Ellipsoid, or a Hyperellipsoid", Geometric Tools,
// 2013).  With semi-axes a0 >= a1 >= a2 and the point y in the first octant,
// the closest point is x_i = r_i y_i / (s + r_i), r_i = (a_i / a2)^2, where s
// is the root of F(s) = sum_i (r_i z_i / (s + r_i))^2 - 1, z_i = y_i / a_i;
// F decreases on (-1, inf) when y2 > 0, and the root lies in
// [z2 - 1, |(r0 z0, r1 z1, z2)| - 1] (or [z2 - 1, 0] inside).  Bisection over
// floating-point midpoints ends when the midpoint equals an end; the cap is
// above the number of halvings a double interval admits.  A zero component
// reduces the problem to an ellipse, or (y2 = 0, point inside near the plane
// of the two larger axes) to the closed form of the interior case.
// ---------------------------------------------------------------------------
constexpr int kEllipsoidBisect = 1100;

PDC_HD double ellipse_root(double n0, double r0, double z1, double g) {
	double s0 = z1 - 1.0;
	double s1 = g < 0.0 ? 0.0 : sqrt(n0 * n0 + z1 * z1) - 1.0;
	double s = s0;
	for (int it = 0; it < kEllipsoidBisect; ++it) {
		s = 0.5 * (s0 + s1);
		if (s == s0 || s == s1) break;
		const double t0 = n0 / (s + r0);
		const double t1 = z1 / (s + 1.0);
		const double f = t0 * t0 + t1 * t1 - 1.0;
		if (f > 0.0) s0 = s;
		else if (f < 0.0) s1 = s;
		else break;
	}
	return s;
}

PDC_HD double ellipsoid_root(double n0, double r0, double n1, double r1, double z2, double g) {
	double s0 = z2 - 1.0;
	double s1 = g < 0.0 ? 0.0 : sqrt(n0 * n0 + n1 * n1 + z2 * z2) - 1.0;
	double s = s0;
	for (int it = 0; it < kEllipsoidBisect; ++it) {
		s = 0.5 * (s0 + s1);
		if (s == s0 || s == s1) break;
		const double t0 = n0 / (s + r0);
		const double t1 = n1 / (s + r1);
		const double t2 = z2 / (s + 1.0);
		const double f = t0 * t0 + t1 * t1 + t2 * t2 - 1.0;
		if (f > 0.0) s0 = s;
		else if (f < 0.0) s1 = s;
		else break;
	}
	return s;
}

// Ellipse with semi-axes a0 >= a1 > 0, point (y0, y1) >= 0.
PDC_HD double ellipse_distance(double a0, double a1, double y0, double y1) {
	if (y1 > 0.0) {
		if (y0 > 0.0) {
			const double z0 = y0 / a0, z1 = y1 / a1;
			const double g = z0 * z0 + z1 * z1 - 1.0;
			double x0 = y0, x1 = y1;
			if (g != 0.0) {
				const double q0 = a0 / a1;
				const double r0 = q0 * q0;
				const double s = ellipse_root(r0 * z0, r0, z1, g);
				x0 = r0 * y0 / (s + r0);
				x1 = y1 / (s + 1.0);
			}
			const double d0 = x0 - y0, d1 = x1 - y1;
			return sqrt(d0 * d0 + d1 * d1);
		}
		return fabs(y1 - a1);
	}
	const double num = a0 * y0, den = a0 * a0 - a1 * a1;
	if (num < den) {
		const double xd = num / den;
		const double x0 = a0 * xd, x1 = a1 * sqrt(1.0 - xd * xd);
		const double d0 = x0 - y0;
		return sqrt(d0 * d0 + x1 * x1);
	}
	return fabs(y0 - a0);
}

// Ellipsoid with semi-axes a0 >= a1 >= a2 > 0, point y >= 0.
PDC_HD double ellipsoid_distance(double a0, double a1, double a2, double y0, double y1, double y2) {
	if (y2 > 0.0) {
		if (y1 > 0.0) {
			if (y0 > 0.0) {
				const double z0 = y0 / a0, z1 = y1 / a1, z2 = y2 / a2;
				const double g = z0 * z0 + z1 * z1 + z2 * z2 - 1.0;
				double x0 = y0, x1 = y1, x2 = y2;
				if (g != 0.0) {
					const double q0 = a0 / a2, q1 = a1 / a2;
					const double r0 = q0 * q0, r1 = q1 * q1;
					const double s = ellipsoid_root(r0 * z0, r0, r1 * z1, r1, z2, g);
					x0 = r0 * y0 / (s + r0);
					x1 = r1 * y1 / (s + r1);
					x2 = y2 / (s + 1.0);
				}
				const double d0 = x0 - y0, d1 = x1 - y1, d2 = x2 - y2;
				return sqrt(d0 * d0 + d1 * d1 + d2 * d2);
			}
			return ellipse_distance(a1, a2, y1, y2);
		}
		if (y0 > 0.0) return ellipse_distance(a0, a2, y0, y2);
		return fabs(y2 - a2);
	}
	const double den0 = a0 * a0 - a2 * a2, den1 = a1 * a1 - a2 * a2;
	const double num0 = a0 * y0, num1 = a1 * y1;
	if (num0 < den0 && num1 < den1) {
		const double e0 = num0 / den0, e1 = num1 / den1;
		const double disc = 1.0 - e0 * e0 - e1 * e1;
		if (disc > 0.0) {
			const double x0 = a0 * e0, x1 = a1 * e1, x2 = a2 * sqrt(disc);
			const double d0 = x0 - y0, d1 = x1 - y1;
			return sqrt(d0 * d0 + d1 * d1 + x2 * x2);
		}
	}
	return ellipse_distance(a0, a1, y0, y1);
}

// Signed distance to the ellipsoid with semi-axes dims (any order): the axes
// sorted descending (ties keep their order), negative where
// (p0/a0)^2 + (p1/a1)^2 + (p2/a2)^2 < 1.
PDC_HD double ellipsoid_sdf(const double* dims, double px, double py, double pz) {
	double a0 = dims[0], a1 = dims[1], a2 = dims[2];
	double y0 = fabs(px), y1 = fabs(py), y2 = fabs(pz);
	double t;
	if (a0 < a1) { t = a0; a0 = a1; a1 = t; t = y0; y0 = y1; y1 = t; }
	if (a1 < a2) { t = a1; a1 = a2; a2 = t; t = y1; y1 = y2; y2 = t; }
	if (a0 < a1) { t = a0; a0 = a1; a1 = t; t = y0; y0 = y1; y1 = t; }
	const double d = ellipsoid_distance(a0, a1, a2, y0, y1, y2);
	const double z0 = y0 / a0, z1 = y1 / a1, z2 = y2 / a2;
	return z0 * z0 + z1 * z1 + z2 * z2 < 1.0 ? -d : d;
}

// ---------------------------------------------------------------------------
// Grid SDF of a primitive.  Grid: n^3 FP32 nodes over [-h, h] per axis,
// h = half extent * (1 + pad), spacing 2h/(n-1); node (i,j,k) at
// lo + (i,j,k)*spacing holds (float) of the analytic FP64 distance.
// Types: 0 box (half extents), 1 sphere (radius in dims[0]), 2 ellipsoid
// (semi-axes).
// ---------------------------------------------------------------------------
PDC_HD double analytic_sdf(int type, const double* dims, double px, double py, double pz) {
	if (type == 1) {
		return sqrt(px * px + py * py + pz * pz) - dims[0];
	}
	if (type == 2) return ellipsoid_sdf(dims, px, py, pz);
	const double qx = fabs(px) - dims[0];
	const double qy = fabs(py) - dims[1];
	const double qz = fabs(pz) - dims[2];
	const double ox = qx > 0.0 ? qx : 0.0;
	const double oy = qy > 0.0 ? qy : 0.0;
	const double oz = qz > 0.0 ? qz : 0.0;
	const double outside = sqrt(ox * ox + oy * oy + oz * oz);
	const double m = qx > (qy > qz ? qy : qz) ? qx : (qy > qz ? qy : qz);
	const double inside = m < 0.0 ? m : 0.0;
	return outside + inside;
}

// ---------------------------------------------------------------------------
// Triangle meshes (Mesh::signedDistance, src/CollisionObject/Mesh.cpp:125-167,
// which calls libigl's signed_distance_pseudonormal).  A triangle is a record
// of kTriRec doubles: vertices a, b, c (0-8), the unit face normal (9-11), the
// pseudonormals of edges ab, bc, ca (12-20) and of vertices a, b, c (21-29).
// The signed distance of a point is its distance to the closest point of the
// mesh, negative when the point lies behind the pseudonormal of the feature
// (face, edge or vertex) holding that closest point (Baerentzen and Aanaes,
// "Signed distance computation using the angle weighted pseudonormal", IEEE
// TVCG 11 (2005)).  Among triangles at the same squared distance the first in
// triangle order decides.
// ---------------------------------------------------------------------------
constexpr int kTriRec = 30;

// Closest point of triangle t to p by the Voronoi regions of its vertices,
// edges and face (Ericson, Real-Time Collision Detection, 2005, sec. 5.1.5).
// Returns the squared distance; *feat: 0 face, 1-3 vertex a, b, c, 4 edge ab,
// 5 edge bc, 6 edge ca.
PDC_HD double tri_closest(const double* t, const double* p, double* cp, int* feat) {
	const double* a = t;
	const double* b = t + 3;
	const double* c = t + 6;
	const double ab[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
	const double ac[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
	const double ap[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
	const double d1 = dot3(ab, ap), d2 = dot3(ac, ap);
	int f;
	if (d1 <= 0.0 && d2 <= 0.0) {
		cp[0] = a[0]; cp[1] = a[1]; cp[2] = a[2];
		f = 1;
	} else {
		const double bp[3] = {p[0] - b[0], p[1] - b[1], p[2] - b[2]};
		const double d3 = dot3(ab, bp), d4 = dot3(ac, bp);
		const double cq[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
		const double d5 = dot3(ab, cq), d6 = dot3(ac, cq);
		const double vc = d1 * d4 - d3 * d2;
		const double vb = d5 * d2 - d1 * d6;
		const double va = d3 * d6 - d5 * d4;
		if (d3 >= 0.0 && d4 <= d3) {
			cp[0] = b[0]; cp[1] = b[1]; cp[2] = b[2];
			f = 2;
		} else if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
			const double v = d1 / (d1 - d3);
			for (int k = 0; k < 3; ++k) cp[k] = a[k] + v * ab[k];
			f = 4;
		} else if (d6 >= 0.0 && d5 <= d6) {
			cp[0] = c[0]; cp[1] = c[1]; cp[2] = c[2];
			f = 3;
		} else if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
			const double w = d2 / (d2 - d6);
			for (int k = 0; k < 3; ++k) cp[k] = a[k] + w * ac[k];
			f = 6;
		} else if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0) {
			const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
			for (int k = 0; k < 3; ++k) cp[k] = b[k] + w * (c[k] - b[k]);
			f = 5;
		} else {
			const double den = 1.0 / (va + vb + vc);
			const double v = vb * den, w = vc * den;
			for (int k = 0; k < 3; ++k) cp[k] = a[k] + ab[k] * v + ac[k] * w;
			f = 0;
		}
	}
	*feat = f;
	const double d[3] = {p[0] - cp[0], p[1] - cp[1], p[2] - cp[2]};
	return dot3(d, d);
}

// One triangle of a mesh distance query: keeps the smallest squared distance
// (*best) and the sign of its feature's pseudonormal side (*sgn).
PDC_HD void mesh_sdf_step(const double* t, const double* p, double* best, double* sgn) {
	double cp[3];
	int f;
	const double d2 = tri_closest(t, p, cp, &f);
	if (d2 < *best) {
		const double* n = f == 0 ? t + 9 : f <= 3 ? t + 21 + 3 * (f - 1) : t + 12 + 3 * (f - 4);
		const double d[3] = {p[0] - cp[0], p[1] - cp[1], p[2] - cp[2]};
		*best = d2;
		*sgn = dot3(d, n) < 0.0 ? -1.0 : 1.0;
	}
}

struct GridDesc {
	double lo[3];
	double sp[3];
	double half[3];
	int n;
	int64_t offset;  // into the float node pool
};

// Trilinear value and gradient at local point q; false when outside the grid box.
template <class Fetch>
PDC_HD bool grid_query(const GridDesc& g, const double* q, Fetch fetch, double* sd, double* grad) {
	double t[3];
	int i[3];
	const double top = (double)(g.n - 1);
	for (int a = 0; a < 3; ++a) {
		const double gg = (q[a] - g.lo[a]) / g.sp[a];
		if (!(gg >= 0.0 && gg <= top)) return false;
		int ia = (int)gg;
		if (ia > g.n - 2) ia = g.n - 2;
		i[a] = ia;
		t[a] = gg - (double)ia;
	}
	const double v000 = fetch(i[0], i[1], i[2]);
	const double v100 = fetch(i[0] + 1, i[1], i[2]);
	const double v010 = fetch(i[0], i[1] + 1, i[2]);
	const double v110 = fetch(i[0] + 1, i[1] + 1, i[2]);
	const double v001 = fetch(i[0], i[1], i[2] + 1);
	const double v101 = fetch(i[0] + 1, i[1], i[2] + 1);
	const double v011 = fetch(i[0], i[1] + 1, i[2] + 1);
	const double v111 = fetch(i[0] + 1, i[1] + 1, i[2] + 1);
	const double tx = t[0], ty = t[1], tz = t[2];
	const double c00 = v000 + tx * (v100 - v000);
	const double c10 = v010 + tx * (v110 - v010);
	const double c01 = v001 + tx * (v101 - v001);
	const double c11 = v011 + tx * (v111 - v011);
	const double c0 = c00 + ty * (c10 - c00);
	const double c1 = c01 + ty * (c11 - c01);
	*sd = c0 + tz * (c1 - c0);
	const double ux = 1.0 - tx, uy = 1.0 - ty, uz = 1.0 - tz;
	grad[0] = (uy * uz * (v100 - v000) + ty * uz * (v110 - v010) + uy * tz * (v101 - v001) +
	           ty * tz * (v111 - v011)) / g.sp[0];
	grad[1] = (ux * uz * (v010 - v000) + tx * uz * (v110 - v100) + ux * tz * (v011 - v001) +
	           tx * tz * (v111 - v101)) / g.sp[1];
	grad[2] = (ux * uy * (v001 - v000) + tx * uy * (v101 - v100) + ux * ty * (v011 - v010) +
	           tx * ty * (v111 - v110)) / g.sp[2];
	return true;
}

// Node layout inside one grid: x fastest, then y, then z.
PDC_HD int64_t grid_index(int n, int i, int j, int k) {
	return ((int64_t)k * n + j) * n + i;
}

}  // namespace pdc
