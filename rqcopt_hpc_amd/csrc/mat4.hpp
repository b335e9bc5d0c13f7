// 4x4 complex matrix algebra and the tangent-space maps of U(4), usable on
// host and device.  Matrices are row-major zc[16].  Each routine follows the
// semantics of the reference function named beside it (src/matrix.c); the
// order of floating-point operations matches it where that is cheap.
#pragma once

#include "zc.hpp"

namespace rqc {

RQC_HD void m4_zero(zc* a) { for (int e = 0; e < 16; e++) { a[e] = {0, 0}; } }

RQC_HD void m4_identity(zc* a) { for (int e = 0; e < 16; e++) { a[e] = {(e % 5 == 0) ? 1.0 : 0.0, 0}; } }

// c = a b (reference src/matrix.c:356-369: c[i][k] accumulated over j in order)
RQC_HD void m4_mul(const zc* a, const zc* b, zc* c)
{
	for (int i = 0; i < 4; i++) {
		for (int k = 0; k < 4; k++) {
			zc s = {0, 0};
			for (int j = 0; j < 4; j++) { s += a[4*i + j] * b[4*j + k]; }
			c[4*i + k] = s;
		}
	}
}

RQC_HD void m4_transpose(const zc* a, zc* at)
{
	for (int i = 0; i < 4; i++) { for (int j = 0; j < 4; j++) { at[4*i + j] = a[4*j + i]; } }
}

RQC_HD void m4_adjoint(const zc* a, zc* ah)
{
	for (int i = 0; i < 4; i++) { for (int j = 0; j < 4; j++) { ah[4*i + j] = conj(a[4*j + i]); } }
}

RQC_HD int m4_upper(int t, int which)
{
	// (0,1) (0,2) (0,3) (1,2) (1,3) (2,3)
	const int r[6] = {0, 0, 0, 1, 1, 2};
	const int c[6] = {1, 2, 3, 2, 3, 3};
	return which == 0 ? r[t] : c[t];
}

// Hermitian part (reference src/matrix.c:150-174)
RQC_HD void m4_symm(const zc* w, zc* z)
{
	for (int k = 0; k < 4; k++) { z[5*k] = {w[5*k].x, 0}; }
	for (int t = 0; t < 6; t++) {
		const int r = m4_upper(t, 0), c = m4_upper(t, 1);
		z[4*r + c] = 0.5 * (w[4*r + c] + conj(w[4*c + r]));
		z[4*c + r] = conj(z[4*r + c]);
	}
}

// skew-Hermitian part (reference src/matrix.c:180-211)
RQC_HD void m4_antisymm(const zc* w, zc* z)
{
	for (int k = 0; k < 4; k++) { z[5*k] = {0, w[5*k].y}; }
	for (int t = 0; t < 6; t++) {
		const int r = m4_upper(t, 0), c = m4_upper(t, 1);
		z[4*r + c] = 0.5 * (w[4*r + c] - conj(w[4*c + r]));
		z[4*c + r] = -conj(z[4*r + c]);
	}
}

// real 16-vector -> skew-Hermitian matrix (reference src/matrix.c:218-244)
RQC_HD void m4_real_to_antisymm(const double* r, zc* w)
{
	for (int k = 0; k < 4; k++) { w[5*k] = {0, r[5*k]}; }
	for (int t = 0; t < 6; t++) {
		const int a = m4_upper(t, 0), b = m4_upper(t, 1);
		const double p = r[4*a + b], q = r[4*b + a];
		w[4*a + b] = {0.5 * (p - q), 0.5 * (p + q)};
		w[4*b + a] = -conj(w[4*a + b]);
	}
}

// reference src/matrix.c:279-288
RQC_HD void m4_antisymm_to_real(const zc* w, double* r)
{
	for (int e = 0; e < 16; e++) { r[e] = w[e].x + w[e].y; }
}

// z = v A(r) (reference src/matrix.c:309-314)
RQC_HD void m4_real_to_tangent(const double* r, const zc* v, zc* z)
{
	zc a[16];
	m4_real_to_antisymm(r, a);
	m4_mul(v, a, z);
}

// r = antisymm_to_real(antisymm(v^dagger z)) (reference src/matrix.c:322-329)
RQC_HD void m4_tangent_to_real(const zc* v, const zc* z, double* r)
{
	zc w[16], t[16];
	m4_adjoint(v, w);
	m4_mul(w, z, t);
	m4_antisymm(t, w);
	m4_antisymm_to_real(w, r);
}

// p = z - u symm(u^dagger z) (reference src/matrix.c:336-347)
RQC_HD void m4_project_tangent(const zc* u, const zc* z, zc* p)
{
	zc v[16], w[16];
	m4_adjoint(u, v);
	m4_mul(v, z, w);
	m4_symm(w, v);
	m4_mul(u, v, w);
	for (int e = 0; e < 16; e++) { p[e] = z[e] - w[e]; }
}

// (abc + cba)/2 (reference src/matrix.c:376-388)
RQC_HD void m4_symmetric_triple(const zc* a, const zc* b, const zc* c, zc* ret)
{
	zc u[16], v[16];
	m4_mul(a, b, u);
	m4_mul(u, c, ret);
	m4_mul(c, b, u);
	m4_mul(u, a, v);
	for (int e = 0; e < 16; e++) { ret[e] = 0.5 * (ret[e] + v[e]); }
}

// Gauss-Jordan inverse with partial pivoting; -1 if singular (reference src/matrix.c:410-475)
RQC_HD int m4_inverse(const zc* a, zc* ainv)
{
	zc m[16];
	for (int e = 0; e < 16; e++) { m[e] = a[e]; }
	m4_identity(ainv);
	for (int k = 0; k < 4; k++) {
		int piv = k;
		double p = cabs(m[5*k]);
		for (int i = k + 1; i < 4; i++) {
			const double q = cabs(m[4*i + k]);
			if (q > p) { piv = i; p = q; }
		}
		if (p == 0) { return -1; }
		if (piv != k) {
			for (int c = 0; c < 4; c++) {
				zc t = m[4*k + c]; m[4*k + c] = m[4*piv + c]; m[4*piv + c] = t;
				t = ainv[4*k + c]; ainv[4*k + c] = ainv[4*piv + c]; ainv[4*piv + c] = t;
			}
		}
		for (int i = 0; i < 4; i++) {
			if (i == k) { continue; }
			const zc s = cdiv(m[4*i + k], m[5*k]);
			for (int c = k + 1; c < 4; c++) { m[4*i + c] -= s * m[4*k + c]; }
			m[4*i + k] = {0, 0};
			for (int c = 0; c < 4; c++) { ainv[4*i + c] -= s * ainv[4*k + c]; }
		}
		for (int c = k + 1; c < 4; c++) { m[4*k + c] = cdiv(m[4*k + c], m[5*k]); }
		for (int c = 0; c < 4; c++) { ainv[4*k + c] = cdiv(ainv[4*k + c], m[5*k]); }
		m[5*k] = {1, 0};
	}
	return 0;
}

// unitary polar factor by Newton iteration u <- (u + u^{-dagger})/2 (Higham 1986;
// reference src/matrix.c:488-510: at most 14 steps, stop after step 4 once the
// update moves entries by less than 1e-14)
RQC_HD void m4_polar(const zc* a, zc* u)
{
	for (int e = 0; e < 16; e++) { u[e] = a[e]; }
	for (int k = 0; k < 14; k++) {
		zc v[16], w[16];
		m4_inverse(u, v);
		m4_adjoint(v, w);
		for (int e = 0; e < 16; e++) { u[e] = 0.5 * (u[e] + w[e]); }
		if (k >= 4) {
			double d = 0;
			for (int e = 0; e < 16; e++) { d = fmax(d, cabs(u[e] - w[e])); }
			if (d < 1e-14) { break; }
		}
	}
}

// v = polar(u (I + A(eta))) (reference src/matrix.c:517-531)
RQC_HD void m4_retract(const zc* u, const double* eta, zc* v)
{
	zc z[16], w[16];
	m4_real_to_antisymm(eta, z);
	for (int k = 0; k < 4; k++) { z[5*k].x += 1; }
	m4_mul(u, z, w);
	m4_polar(w, v);
}

// Riemannian gradient and Hessian-vector product of one gate as real 16-vectors
// (reference src/target.c:650-671): dv, hz are the Euclidean (holomorphic)
// gradient and directional second derivative, v the gate, z the direction.
RQC_HD void m4_project_grad_hvp(const zc* v, const zc* z, const zc* dv, const zc* hz, double* grad, double* hvp)
{
	zc g[16], h[16], gh[16], pd[16];
	for (int e = 0; e < 16; e++) { g[e] = conj(dv[e]); }
	m4_tangent_to_real(v, g, grad);
	if (hvp != nullptr) {
		for (int e = 0; e < 16; e++) { h[e] = conj(hz[e]); }
		m4_adjoint(g, gh);
		m4_symmetric_triple(z, gh, v, pd);
		for (int e = 0; e < 16; e++) { h[e] -= pd[e]; }
		m4_tangent_to_real(v, h, hvp);
	}
}

}  // namespace rqc
