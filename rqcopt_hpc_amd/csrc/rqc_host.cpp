// Host entry points of librqcopt_hip: 4x4 algebra (matrix.h), seeded random
// numbers (rng.h), the RTR-HVP driver (trust_region.h) and its brickwall glue
// (brickwall_opt.h).  These are O(nlayers) host computations that the
// reference also runs on the host; every target / gradient / HVP they request
// runs on the GPU through target.h.  The 4x4 algebra is the same
// __host__ __device__ code the GPU finalize kernel uses for the projection.

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mat4.hpp"
#include "rqcopt/matrix.h"
#include "rqcopt/rng.h"
#include "rqcopt/trust_region.h"
#include "rqcopt/brickwall_opt.h"
#include "rqcopt/circuit_opt.h"
#include "rqcopt/rqc_device.h"
#include "rqcopt/target.h"
#include "rqcopt/rng.h"



static inline zc* Z(struct mat4x4* a) { return reinterpret_cast<zc*>(a->data); }
static inline const zc* Z(const struct mat4x4* a) { return reinterpret_cast<const zc*>(a->data); }

// ============================================================== matrix.h
extern "C" void zero_matrix(struct mat4x4* a) { rqc::m4_zero(Z(a)); }
extern "C" void identity_matrix(struct mat4x4* a) { rqc::m4_identity(Z(a)); }

extern "C" void scale_matrix(struct mat4x4* RQC_RESTRICT a, const double x)
{
	for (int e = 0; e < 16; e++) { Z(a)[e] = x * Z(a)[e]; }
}

extern "C" void conjugate_matrix(struct mat4x4* a)
{
	for (int e = 0; e < 16; e++) { Z(a)[e] = conj(Z(a)[e]); }
}

extern "C" void add_matrix(struct mat4x4* RQC_RESTRICT a, const struct mat4x4* RQC_RESTRICT b)
{
	for (int e = 0; e < 16; e++) { Z(a)[e] += Z(b)[e]; }
}

extern "C" void sub_matrix(struct mat4x4* RQC_RESTRICT a, const struct mat4x4* RQC_RESTRICT b)
{
	for (int e = 0; e < 16; e++) { Z(a)[e] -= Z(b)[e]; }
}

extern "C" void add_matrices(const struct mat4x4* RQC_RESTRICT a, const struct mat4x4* RQC_RESTRICT b, struct mat4x4* RQC_RESTRICT c)
{
	for (int e = 0; e < 16; e++) { Z(c)[e] = Z(a)[e] + Z(b)[e]; }
}

extern "C" void sub_matrices(const struct mat4x4* RQC_RESTRICT a, const struct mat4x4* RQC_RESTRICT b, struct mat4x4* RQC_RESTRICT c)
{
	for (int e = 0; e < 16; e++) { Z(c)[e] = Z(a)[e] - Z(b)[e]; }
}

extern "C" void transpose(const struct mat4x4* RQC_RESTRICT a, struct mat4x4* RQC_RESTRICT at) { rqc::m4_transpose(Z(a), Z(at)); }
extern "C" void adjoint(const struct mat4x4* RQC_RESTRICT a, struct mat4x4* RQC_RESTRICT ah) { rqc::m4_adjoint(Z(a), Z(ah)); }
extern "C" void symm(const struct mat4x4* RQC_RESTRICT w, struct mat4x4* RQC_RESTRICT z) { rqc::m4_symm(Z(w), Z(z)); }
extern "C" void antisymm(const struct mat4x4* RQC_RESTRICT w, struct mat4x4* RQC_RESTRICT z) { rqc::m4_antisymm(Z(w), Z(z)); }
extern "C" void real_to_antisymm(const double* r, struct mat4x4* w) { rqc::m4_real_to_antisymm(r, Z(w)); }
extern "C" void antisymm_to_real(const struct mat4x4* w, double* r) { rqc::m4_antisymm_to_real(Z(w), r); }

extern "C" void real_to_tangent(const double* r, const struct mat4x4* RQC_RESTRICT v, struct mat4x4* RQC_RESTRICT z)
{
	rqc::m4_real_to_tangent(r, Z(v), Z(z));
}

extern "C" void tangent_to_real(const struct mat4x4* RQC_RESTRICT v, const struct mat4x4* RQC_RESTRICT z, double* r)
{
	rqc::m4_tangent_to_real(Z(v), Z(z), r);
}

extern "C" void project_tangent(const struct mat4x4* RQC_RESTRICT u, const struct mat4x4* RQC_RESTRICT z, struct mat4x4* RQC_RESTRICT p)
{
	rqc::m4_project_tangent(Z(u), Z(z), Z(p));
}

extern "C" void multiply_matrices(const struct mat4x4* RQC_RESTRICT a, const struct mat4x4* RQC_RESTRICT b, struct mat4x4* RQC_RESTRICT c)
{
	rqc::m4_mul(Z(a), Z(b), Z(c));
}

extern "C" void symmetric_triple_matrix_product(const struct mat4x4* RQC_RESTRICT a, const struct mat4x4* RQC_RESTRICT b,
	const struct mat4x4* RQC_RESTRICT c, struct mat4x4* RQC_RESTRICT ret)
{
	rqc::m4_symmetric_triple(Z(a), Z(b), Z(c), Z(ret));
}

extern "C" int inverse_matrix(const struct mat4x4* RQC_RESTRICT a, struct mat4x4* RQC_RESTRICT ainv)
{
	return rqc::m4_inverse(Z(a), Z(ainv));
}

extern "C" void polar_factor(const struct mat4x4* RQC_RESTRICT a, struct mat4x4* RQC_RESTRICT u) { rqc::m4_polar(Z(a), Z(u)); }

extern "C" void retract(const struct mat4x4* RQC_RESTRICT u, const double* RQC_RESTRICT eta, struct mat4x4* RQC_RESTRICT v)
{
	rqc::m4_retract(Z(u), eta, Z(v));
}

// ============================================================== rng.h
// PCG32 with XSH-RR output (O'Neill 2014); two streams give 64 bits
extern "C" uint32_t pcg32_random_r(struct pcg32_random* rng)
{
	const uint64_t old = rng->state;
	rng->state = old * 6364136223846793005ULL + rng->inc;
	const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
	const uint32_t rot = (uint32_t)(old >> 59u);
	return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

extern "C" void pcg32_srandom_r(const uint64_t initstate, const uint64_t initseq, struct pcg32_random* rng)
{
	rng->state = 0;
	rng->inc = (initseq << 1u) | 1u;
	pcg32_random_r(rng);
	rng->state += initstate;
	pcg32_random_r(rng);
}

extern "C" void pcg32x2_srandom_r(uint64_t seed1, uint64_t seed2, uint64_t seq1, uint64_t seq2, struct pcg32x2_random* rng)
{
	const uint64_t mask = ~0ULL >> 1;
	if ((seq1 & mask) == (seq2 & mask)) { seq2 = ~seq2; }  // the two streams must differ
	pcg32_srandom_r(seed1, seq1, &rng->gen[0]);
	pcg32_srandom_r(seed2, seq2, &rng->gen[1]);
}

extern "C" uint64_t pcg32x2_random_r(struct pcg32x2_random* rng)
{
	const uint64_t hi = pcg32_random_r(&rng->gen[0]);
	const uint64_t lo = pcg32_random_r(&rng->gen[1]);
	return (hi << 32) | lo;
}

// seeding through a temporary generator (reference src/rng.c:13-25; the high
// word of each 64-bit seed is drawn first)
extern "C" void seed_rng_state(const uint64_t seed, struct rng_state* state)
{
	struct pcg32_random t;
	pcg32_srandom_r(seed, 6364136223846793005ULL * seed + 1442695040888963407ULL, &t);
	uint64_t v[4];
	for (int k = 0; k < 4; k++) {
		const uint64_t hi = pcg32_random_r(&t);
		const uint64_t lo = pcg32_random_r(&t);
		v[k] = (hi << 32) | lo;
	}
	pcg32x2_srandom_r(v[0], v[1], v[2], v[3], &state->pcgstate);
}

extern "C" uint32_t rand_uint32(struct rng_state* state) { return pcg32_random_r(&state->pcgstate.gen[0]); }
extern "C" uint64_t rand_uint64(struct rng_state* state) { return pcg32x2_random_r(&state->pcgstate); }
extern "C" double randu(struct rng_state* state) { return std::ldexp((double)rand_uint64(state), -64); }

// Box-Muller (1958), using log(u1) with u1 = 0 mapped to 1
extern "C" double randn(struct rng_state* state)
{
	double u1 = randu(state);
	const double u2 = randu(state);
	if (u1 == 0) { u1 = 1; }
	return std::sqrt(-2 * std::log(u1)) * std::sin(6.2831853071795864769 * u2);
}

extern "C" numeric crandn(struct rng_state* state)
{
	const double x = randn(state);
	const double y = randn(state);
	const double s = std::sqrt(2.0);
	return numeric(x / s, y / s);
}

// rejection below (2^32 - bound) mod bound, so that r mod bound is unbiased
// (reference src/pcg_basic.c:58-89, the PCG authors' bounded draw)
extern "C" uint32_t pcg32_boundedrand_r(const uint32_t bound, struct pcg32_random* rng)
{
	const uint32_t threshold = (uint32_t)(-bound) % bound;
	for (;;) {
		const uint32_t r = pcg32_random_r(rng);
		if (r >= threshold) { return r % bound; }
	}
}

extern "C" uint64_t pcg32x2_boundedrand_r(const uint64_t bound, struct pcg32x2_random* rng)
{
	const uint64_t threshold = (uint64_t)(-bound) % bound;
	for (;;) {
		const uint64_t r = pcg32x2_random_r(rng);
		if (r >= threshold) { return r % bound; }
	}
}

extern "C" uint64_t rand_interval(const uint64_t bound, struct rng_state* state)
{
	return pcg32x2_boundedrand_r(bound, &state->pcgstate);
}

// num_samples distinct values of [0, bound) by Floyd's algorithm (reference
// src/rng.c:58-79): draw t in [0, j) for j = bound - num_samples + 1 ..
// bound, and take j - 1 instead when t was already drawn
extern "C" void rand_choice(const uint64_t bound, const uint64_t num_samples, struct rng_state* state, uint64_t* ret)
{
	for (uint64_t i = 0; i < num_samples && i < bound; i++) {
		const uint64_t j = bound - num_samples + 1 + i;
		const uint64_t t = rand_interval(j, state);
		bool drawn = false;
		for (uint64_t k = 0; k < i && !drawn; k++) { drawn = (ret[k] == t); }
		ret[i] = drawn ? j - 1 : t;
	}
}

extern "C" float randuf(struct rng_state* state) { return std::ldexp((float)rand_uint32(state), -32); }

// single-precision Box-Muller, u1 = 0 mapped to 1
extern "C" float randnf(struct rng_state* state)
{
	float u1 = randuf(state);
	const float u2 = randuf(state);
	if (u1 == 0) { u1 = 1; }
	return std::sqrt(-2 * std::log(u1)) * std::sin((float)6.2831853071795864769 * u2);
}

extern "C" std::complex<float> crandnf(struct rng_state* state)
{
	const float x = randnf(state);
	const float y = randnf(state);
	const float s = std::sqrt(2.f);
	return std::complex<float>(x / s, y / s);
}

// ============================================================== trust_region.h
static double ddot(int n, const double* x, const double* y)
{
	double s = 0;
	for (int i = 0; i < n; i++) { s += x[i] * y[i]; }
	return s;
}

// roots of x^2 + 2 p x + q = 0, ascending (reference src/trust_region.c:14-40)
static void solve_quadratic(double p, double q, double x[2])
{
	if (p == 0) {
		const double a = std::sqrt(-q);
		x[0] = -a; x[1] = a;
		return;
	}
	const double a = -(p + std::copysign(1.0, p) * std::sqrt(p * p - q));
	const double b = q / a;
	if (a <= b) { x[0] = a; x[1] = b; } else { x[0] = b; x[1] = a; }
}

// t > 0 with ||b + t d|| = radius (reference src/trust_region.c:49-65)
static double move_to_boundary(const double* b, const double* d, int n, double radius)
{
	const double dsq = ddot(n, d, d);
	if (dsq == 0) {
		fprintf(stderr, "input vector 'd' is zero\n");
		return 0;
	}
	const double p = ddot(n, b, d) / dsq;
	const double q = (ddot(n, b, b) - radius * radius) / dsq;
	double t[2];
	solve_quadratic(p, q, t);
	if (t[1] < 0) { fprintf(stderr, "encountered t < 0: t = %g\n", t[1]); }
	return t[1];
}

// Steihaug-Toint truncated CG (reference src/trust_region.c:83-141 and :159-217);
// hv(d, out) applies the Hessian, as a matrix or through HVP evaluations
template <typename HessApply>
static bool truncated_cg(const double* grad, int n, double radius, const struct truncated_cg_params* params, double* z,
	HessApply hv)
{
	std::vector<double> r(grad, grad + n), d(n), hd(n);
	double rsq = ddot(n, r.data(), r.data());
	const double stoptol = std::fmax(params->abstol, params->reltol * std::sqrt(rsq));
	for (int i = 0; i < n; i++) { z[i] = 0; d[i] = -r[i]; }
	for (int j = 0; j < params->maxiter; j++) {
		hv(d.data(), hd.data());
		const double dhd = ddot(n, d.data(), hd.data());
		const double t = move_to_boundary(z, d.data(), n, radius);
		const double alpha = rsq / dhd;
		if (dhd <= 0 || alpha > t) {
			for (int i = 0; i < n; i++) { z[i] += t * d[i]; }
			return true;
		}
		for (int i = 0; i < n; i++) { r[i] += alpha * hd[i]; }
		for (int i = 0; i < n; i++) { z[i] += alpha * d[i]; }
		const double rsq_next = ddot(n, r.data(), r.data());
		if (std::sqrt(rsq_next) <= stoptol) { break; }
		const double beta = rsq_next / rsq;
		for (int i = 0; i < n; i++) { d[i] = beta * d[i] - r[i]; }
		rsq = rsq_next;
	}
	return false;
}

// y = A x for a row-major n x n matrix
static void matvec(int n, const double* A, const double* x, double* y)
{
	for (int a = 0; a < n; a++) { y[a] = ddot(n, &A[(size_t)a * n], x); }
}

extern "C" bool truncated_cg_hmat(const double* grad, const double* hess, int n, double radius, const struct truncated_cg_params* params, double* z)
{
	return truncated_cg(grad, n, radius, params, z, [&](const double* d, double* hd) { matvec(n, hess, d, hd); });
}

extern "C" bool truncated_cg_hvp(const double* RQC_RESTRICT x, const double* RQC_RESTRICT grad, const hessian_vector_product_func f_hvp, void* fdata,
	const int n, const double radius, const struct truncated_cg_params* params, double* RQC_RESTRICT z)
{
	return truncated_cg(grad, n, radius, params, z,
		[&](const double* d, double* hd) { f_hvp(x, fdata, const_cast<double*>(d), hd); });
}

// Riemannian trust region (Absil et al. 2008, Alg. 10; reference src/trust_region.c:230-293
// and :306-368).  derivs(x, grad) evaluates f and the gradient at x and readies the
// Hessian at x for hv(d, out).
template <typename Derivs, typename HessApply>
static void trust_region(const target_func f, void* fdata, retract_func retract_fn, void* rdata, int n, const double* x_init,
	int m, const struct rtr_params* params, int niter, double* f_iter, double* x_final, Derivs derivs, HessApply hv)
{
	if (!(0 <= params->rho_trust && params->rho_trust < 0.25)) {
		fprintf(stderr, "rho_trust must lie in [0, 0.25)\n");
		abort();
	}
	std::vector<double> xa(x_init, x_init + m), xb(m), grad(n), eta(n), heta(n);
	double* x = xa.data();
	double* xn = xb.data();
	double radius = params->radius_init;
	if (params->g_func != nullptr) { params->g_iter[0] = params->g_func(x, params->g_data); }
	for (int k = 0; k < niter; k++) {
		const double fx = derivs(x, grad.data());
		const bool on_boundary = truncated_cg(grad.data(), n, radius, &params->tcg_params, eta.data(),
			[&](const double* d, double* hd) { hv(x, d, hd); });
		retract_fn(x, eta.data(), rdata, xn);
		f_iter[k] = fx;
		// model decrease, Eq. (7.7) of Absil et al.
		hv(x, eta.data(), heta.data());
		const double rho = (f(xn, fdata) - fx) / (ddot(n, grad.data(), eta.data()) + 0.5 * ddot(n, eta.data(), heta.data()));
		if (rho < 0.25) {
			radius *= 0.25;
		} else if (rho > 0.75 && on_boundary) {
			radius = std::fmin(2 * radius, params->maxradius);
		}
		if (rho > params->rho_trust) { std::swap(x, xn); }
		if (params->g_func != nullptr) { params->g_iter[k + 1] = params->g_func(x, params->g_data); }
	}
	std::memcpy(x_final, x, m * sizeof(double));
	f_iter[niter] = f(x_final, fdata);
}

extern "C" void riemannian_trust_region_optimize_hmat(target_func f, target_gradient_hessian_func f_deriv_hess, void* fdata,
	retract_func retract_fn, void* rdata,
	const int n, const double* x_init, const int m, const struct rtr_params* params, const int niter, double* f_iter, double* x_final)
{
	std::vector<double> hess((size_t)n * n);
	trust_region(f, fdata, retract_fn, rdata, n, x_init, m, params, niter, f_iter, x_final,
		[&](const double* x, double* grad) { return f_deriv_hess(x, fdata, grad, hess.data()); },
		[&](const double*, const double* d, double* hd) { matvec(n, hess.data(), d, hd); });
}

extern "C" void riemannian_trust_region_optimize_hvp(const target_func f, const target_gradient_func f_deriv, const hessian_vector_product_func f_hvp, void* fdata,
	retract_func retract_fn, void* rdata,
	const int n, const double* x_init, const int m, const struct rtr_params* params, const int niter, double* f_iter, double* x_final)
{
	trust_region(f, fdata, retract_fn, rdata, n, x_init, m, params, niter, f_iter, x_final,
		[&](const double* x, double* grad) { return f_deriv(x, fdata, grad); },
		[&](const double* x, const double* d, double* hd) { f_hvp(x, fdata, const_cast<double*>(d), hd); });
}

// ============================================================== target.h: Riemannian Hessian matrix
// Projection of the complex Hessian onto the tangent spaces (reference
// src/target.c:937-994): for the upper blocks (i <= j), column k of block (i, j)
// is the tangent basis vector k at V_j pushed through hess, conjugated (the
// derivative convention), projected at V_i on the diagonal blocks together with
// the curvature term 1/2 (Z dv^dagger V + V dv^dagger Z) of the gradient, and
// written in the tangent coordinates at V_i; the lower half by symmetry.
// hess2 (NULL for the unitary targets) is the Hessian of the conjugated gates of
// the block-encoding targets, added unconjugated (reference src/target.c:1665-1680).
static void project_hessian_matrix(int nlayers, const struct mat4x4* vlist, const std::vector<zc>& dvc, const zc* hess,
	const zc* hess2, double* H)
{

	const int m = 16 * nlayers;
	for (int i = 0; i < nlayers; i++) {
		for (int j = i; j < nlayers; j++) {
			for (int k = 0; k < 16; k++) {
				double r[16] = {0};
				r[k] = 1;
				zc Zm[16], G[16];
				rqc::m4_real_to_tangent(r, Z(&vlist[j]), Zm);
				for (int x = 0; x < 16; x++) {
					zc acc = {0, 0};
					for (int y = 0; y < 16; y++) { cmac(acc, hess[(size_t)(i * 16 + x) * m + j * 16 + y], Zm[y]); }
					G[x] = zc{acc.x, -acc.y};
					if (hess2 != nullptr) {
						zc acc2 = {0, 0};
						for (int y = 0; y < 16; y++) { cmac(acc2, hess2[(size_t)(i * 16 + x) * m + j * 16 + y], Zm[y]); }
						G[x] = zc{G[x].x + acc2.x, G[x].y + acc2.y};
					}
				}
				if (i == j) {
					zc P[16], gh[16], T[16];
					rqc::m4_project_tangent(Z(&vlist[i]), G, P);
					rqc::m4_adjoint(&dvc[16 * (size_t)i], gh);
					rqc::m4_symmetric_triple(Zm, gh, Z(&vlist[i]), T);
					for (int e = 0; e < 16; e++) { G[e] = zc{P[e].x - T[e].x, P[e].y - T[e].y}; }
				}
				rqc::m4_tangent_to_real(Z(&vlist[i]), G, r);
				for (int x = 0; x < 16; x++) { H[(size_t)(i * 16 + x) * m + j * 16 + k] = r[x]; }
			}
		}
	}
	for (int a = 0; a < m; a++) {
		for (int b = 0; b < a; b++) { H[(size_t)a * m + b] = H[(size_t)b * m + a]; }
	}
}

// gradient (conjugated, derivative convention) in tangent coordinates and the projected Hessian
// matrix (reference src/target.c:1003-1040 and :1050-1088)
static void gradient_vector_hessian_matrix(const struct mat4x4 vlist[], int nlayers, const std::vector<struct mat4x4>& dv,
	const std::vector<numeric>& hess, double* grad_vec, double* H, const std::vector<numeric>* hess2 = nullptr)
{
	std::vector<zc> dvc(16 * (size_t)nlayers);
	for (int i = 0; i < nlayers; i++) {
		for (int e = 0; e < 16; e++) { dvc[16 * (size_t)i + e] = zc{dv[i].data[e].real(), -dv[i].data[e].imag()}; }
		rqc::m4_tangent_to_real(Z(&vlist[i]), &dvc[16 * (size_t)i], &grad_vec[16 * i]);
	}
	project_hessian_matrix(nlayers, vlist, dvc, reinterpret_cast<const zc*>(hess.data()),
		hess2 != nullptr ? reinterpret_cast<const zc*>(hess2->data()) : nullptr, H);
}

extern "C" int brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[], const long nsamples,
	struct rng_state* rng, numeric* fval, double* grad_vec, double* H)
{
	const int m = 16 * nlayers;
	std::vector<struct mat4x4> dv((size_t)std::max(nlayers, 1));
	std::vector<numeric> hess((size_t)m * m);
	const int ret = brickwall_unitary_target_gradient_hessian_sampling(ufunc, udata, vlist, nlayers, nqubits, perms, nsamples, rng,
		fval, dv.data(), hess.data());
	if (ret < 0) { return ret; }
	gradient_vector_hessian_matrix(vlist, nlayers, dv, hess, grad_vec, H);
	return 0;
}

extern "C" int brickwall_unitary_target_gradient_vector_hessian_matrix(linear_func ufunc, void* udata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[],
	numeric* fval, double* grad_vec, double* H)
{

	const int m = 16 * nlayers;
	std::vector<struct mat4x4> dv((size_t)std::max(nlayers, 1));
	std::vector<numeric> hess((size_t)m * m);
	const int ret = brickwall_unitary_target_gradient_hessian(ufunc, udata, vlist, nlayers, nqubits, perms, fval, dv.data(), hess.data());
	if (ret < 0) { return ret; }
	gradient_vector_hessian_matrix(vlist, nlayers, dv, hess, grad_vec, H);
	return 0;
}

// reference src/target.c:1631-1720
extern "C" int brickwall_blockenc_target_gradient_vector_hessian_matrix(linear_func hfunc, void* hdata,
	const struct mat4x4 vlist[], const int nlayers, const int nqubits, const int* perms[], double* fval, double* grad_vec,
	double* H)
{
	const int m = 16 * nlayers;
	std::vector<struct mat4x4> dv((size_t)std::max(nlayers, 1));
	std::vector<numeric> hess1((size_t)m * m), hess2((size_t)m * m);
	const int ret = brickwall_blockenc_target_gradient_hessian(hfunc, hdata, vlist, nlayers, nqubits, perms, fval, dv.data(),
		hess1.data(), hess2.data());
	if (ret < 0) { return ret; }
	gradient_vector_hessian_matrix(vlist, nlayers, dv, hess1, grad_vec, H, &hess2);
	return 0;
}

// ============================================================== brickwall_opt.h
namespace {

// The last (x, f, projected gradient) of a mixed-precision target evaluation.  A mixed target
// evaluation runs the gradient sweep (its f is corrected for the FP32 rounding of the gates
// with the gradient), so the f callback keeps the gradient, and the gradient callback that the
// trust region makes at the same x after an accepted step returns it instead of sweeping again.
// The values are those a new evaluation would return (evaluations are deterministic).  FP64
// runs keep the reference's calls: a target evaluation there has no gradient.
struct GradMemo
{
	std::vector<double> x, grad;
	double f = 0;
	bool valid = false;

	bool hit(const double* xq, size_t m, double* grad_out, double& f_out) const
	{
		if (!valid || x.size() != m || std::memcmp(x.data(), xq, m * sizeof(double)) != 0) { return false; }
		std::memcpy(grad_out, grad.data(), grad.size() * sizeof(double));
		f_out = f;
		return true;
	}
	void keep(const double* xq, size_t m, const double* g, size_t n, double fv)
	{
		x.assign(xq, xq + m);
		grad.assign(g, g + n);
		f = fv;
		valid = true;
	}
};

struct BrickwallData
{
	linear_func ufunc;
	void* udata;
	const int** perms;
	int nlayers;
	int nqubits;
	GradMemo memo;
};

bool mixed_precision() { return rqc_get_precision() == RQC_PRECISION_MIXED; }

// wrappers of reference src/brickwall_opt.c:22-79
double bw_f_deriv(const double* RQC_RESTRICT x, void* fdata, double* RQC_RESTRICT grad)
{
	auto* d = static_cast<BrickwallData*>(fdata);
	const size_t m = 32 * (size_t)d->nlayers, n = 16 * (size_t)d->nlayers;
	double fm;
	if (mixed_precision() && d->memo.hit(x, m, grad, fm)) { return fm; }
	numeric f;
	if (brickwall_unitary_target_and_projected_gradient(d->ufunc, d->udata, reinterpret_cast<const struct mat4x4*>(x), d->nlayers,
			d->nqubits, d->perms, &f, grad) < 0) {
		fprintf(stderr, "target function and derivative evaluation failed internally\n");
		return -1;
	}
	if (mixed_precision()) { d->memo.keep(x, m, grad, n, f.real()); }
	return f.real();
}

double bw_f(const double* x, void* fdata)
{
	auto* d = static_cast<BrickwallData*>(fdata);
	if (mixed_precision()) {
		std::vector<double> grad(16 * (size_t)d->nlayers);
		return bw_f_deriv(x, fdata, grad.data());
	}
	numeric f;
	if (brickwall_unitary_target(d->ufunc, d->udata, reinterpret_cast<const struct mat4x4*>(x), d->nlayers, d->nqubits, d->perms, &f) < 0) {
		fprintf(stderr, "target function evaluation failed internally\n");
		return -1;
	}
	return f.real();
}

void bw_f_hvp(const double* RQC_RESTRICT x, void* fdata, double* RQC_RESTRICT vec, double* RQC_RESTRICT hvp)
{
	auto* d = static_cast<BrickwallData*>(fdata);
	const auto* vlist = reinterpret_cast<const struct mat4x4*>(x);
	std::vector<struct mat4x4> vdirs(d->nlayers);
	for (int i = 0; i < d->nlayers; i++) { real_to_tangent(&vec[16 * i], &vlist[i], &vdirs[i]); }
	std::vector<double> grad(16 * (size_t)d->nlayers);
	numeric f;
	if (brickwall_unitary_target_projected_hessian_vector_product(d->ufunc, d->udata, vlist, vdirs.data(), d->nlayers, d->perms,
			d->nqubits, &f, grad.data(), hvp) < 0) {
		fprintf(stderr, "Hessian-vector product evaluation failed internally\n");
	}
}

void bw_retract(const double* RQC_RESTRICT x, const double* RQC_RESTRICT eta, void* rdata, double* RQC_RESTRICT xs)
{
	const int nlayers = *static_cast<int*>(rdata);
	const auto* v = reinterpret_cast<const struct mat4x4*>(x);
	auto* w = reinterpret_cast<struct mat4x4*>(xs);
	for (int i = 0; i < nlayers; i++) { retract(&v[i], &eta[16 * i], &w[i]); }
}

}  // namespace

// wrapper of reference src/brickwall_opt.c f_deriv_hess: f, projected gradient and Hessian matrix
static double bw_f_deriv_hess(const double* RQC_RESTRICT x, void* fdata, double* RQC_RESTRICT grad, double* RQC_RESTRICT hess)
{
	auto* d = static_cast<BrickwallData*>(fdata);
	numeric f;
	if (brickwall_unitary_target_gradient_vector_hessian_matrix(d->ufunc, d->udata, reinterpret_cast<const struct mat4x4*>(x),
			d->nlayers, d->nqubits, d->perms, &f, grad, hess) < 0) {
		fprintf(stderr, "target function and derivative evaluation failed internally\n");
		return -1;
	}
	return f.real();
}

extern "C" void optimize_brickwall_circuit_hvp(linear_func ufunc, void* udata,
	const struct mat4x4 vlist_start[], const int nlayers, const int nqubits, const int* perms[],
	struct rtr_params* params, const int niter, double* f_iter, struct mat4x4 vlist_opt[])
{
	BrickwallData fdata{ufunc, udata, perms, nlayers, nqubits, {}};
	params->g_func = nullptr;
	params->g_data = nullptr;
	params->g_iter = nullptr;
	int rdata = nlayers;
	riemannian_trust_region_optimize_hvp(bw_f, bw_f_deriv, bw_f_hvp, &fdata, bw_retract, &rdata,
		nlayers * num_tangent_params, reinterpret_cast<const double*>(vlist_start), nlayers * 32, params, niter, f_iter,
		reinterpret_cast<double*>(vlist_opt));
}

extern "C" void optimize_brickwall_circuit_hmat(linear_func ufunc, void* udata,
	const struct mat4x4 vlist_start[], const int nlayers, const int nqubits, const int* perms[],
	struct rtr_params* params, const int niter, double* f_iter, struct mat4x4 vlist_opt[])
{
	BrickwallData fdata{ufunc, udata, perms, nlayers, nqubits, {}};
	params->g_func = nullptr;
	params->g_data = nullptr;
	params->g_iter = nullptr;
	int rdata = nlayers;
	riemannian_trust_region_optimize_hmat(bw_f, bw_f_deriv_hess, &fdata, bw_retract, &rdata,
		nlayers * num_tangent_params, reinterpret_cast<const double*>(vlist_start), nlayers * 32, params, niter, f_iter,
		reinterpret_cast<double*>(vlist_opt));
}

// ============================================================== circuit_opt.h
namespace {

struct CircuitData
{
	linear_func ufunc;
	void* udata;
	const int* wires;
	int ngates;
	int nqubits;
	GradMemo memo;
};

// wrappers of reference src/circuit_opt.c:22-79 (mixed precision: as bw_f / bw_f_deriv)
double qc_f_deriv(const double* RQC_RESTRICT x, void* fdata, double* RQC_RESTRICT grad)
{
	auto* d = static_cast<CircuitData*>(fdata);
	const size_t m = 32 * (size_t)d->ngates, n = 16 * (size_t)d->ngates;
	double fm;
	if (mixed_precision() && d->memo.hit(x, m, grad, fm)) { return fm; }
	numeric f;
	if (circuit_unitary_target_and_projected_gradient(d->ufunc, d->udata, reinterpret_cast<const struct mat4x4*>(x), d->ngates,
			d->wires, d->nqubits, &f, grad) < 0) {
		fprintf(stderr, "target function and derivative evaluation failed internally\n");
		return -1;
	}
	if (mixed_precision()) { d->memo.keep(x, m, grad, n, f.real()); }
	return f.real();
}

double qc_f(const double* x, void* fdata)
{
	auto* d = static_cast<CircuitData*>(fdata);
	if (mixed_precision()) {
		std::vector<double> grad(16 * (size_t)d->ngates);
		return qc_f_deriv(x, fdata, grad.data());
	}
	numeric f;
	if (circuit_unitary_target(d->ufunc, d->udata, reinterpret_cast<const struct mat4x4*>(x), d->ngates, d->wires, d->nqubits, &f) < 0) {
		fprintf(stderr, "target function evaluation failed internally\n");
		return -1;
	}
	return f.real();
}

void qc_f_hvp(const double* RQC_RESTRICT x, void* fdata, double* RQC_RESTRICT vec, double* RQC_RESTRICT hvp)
{
	auto* d = static_cast<CircuitData*>(fdata);
	const auto* gates = reinterpret_cast<const struct mat4x4*>(x);
	std::vector<struct mat4x4> dirs(d->ngates);
	for (int i = 0; i < d->ngates; i++) { real_to_tangent(&vec[16 * i], &gates[i], &dirs[i]); }
	std::vector<double> grad(16 * (size_t)d->ngates);
	numeric f;
	if (circuit_unitary_target_projected_hessian_vector_product(d->ufunc, d->udata, gates, dirs.data(), d->ngates, d->wires,
			d->nqubits, &f, grad.data(), hvp) < 0) {
		fprintf(stderr, "Hessian-vector product evaluation failed internally\n");
	}
}

}  // namespace

extern "C" void optimize_quantum_circuit(linear_func ufunc, void* udata, const struct mat4x4 gates_start[], const int ngates,
	const int nqubits, const int wires[], struct rtr_params* params, const int niter, double* f_iter, struct mat4x4 gates_opt[])
{
	CircuitData fdata{ufunc, udata, wires, ngates, nqubits, {}};
	params->g_func = nullptr;
	params->g_data = nullptr;
	params->g_iter = nullptr;
	int rdata = ngates;
	riemannian_trust_region_optimize_hvp(qc_f, qc_f_deriv, qc_f_hvp, &fdata, bw_retract, &rdata, ngates * num_tangent_params,
		reinterpret_cast<const double*>(gates_start), ngates * 32, params, niter, f_iter, reinterpret_cast<double*>(gates_opt));
}

// ============================================================== brickwall_opt.h: sampled targets
namespace {

struct BrickwallSamplingData
{
	linear_func ufunc;
	void* udata;
	const int** perms;
	struct rng_state* rng;
	long nsamples;
	int nlayers;
	int nqubits;
};

// wrappers of reference src/brickwall_opt.c:176-215; every call draws fresh samples from rng
double bws_f(const double* x, void* fdata)
{
	auto* d = static_cast<BrickwallSamplingData*>(fdata);
	numeric f;
	if (brickwall_unitary_target_sampling(d->ufunc, d->udata, reinterpret_cast<const struct mat4x4*>(x), d->nlayers, d->nqubits,
			d->perms, d->nsamples, d->rng, &f) < 0) {
		fprintf(stderr, "target function evaluation failed internally\n");
		return -1;
	}
	return f.real();
}

double bws_f_deriv_hess(const double* RQC_RESTRICT x, void* fdata, double* RQC_RESTRICT grad, double* RQC_RESTRICT hess)
{
	auto* d = static_cast<BrickwallSamplingData*>(fdata);
	numeric f;
	if (brickwall_unitary_target_gradient_vector_hessian_matrix_sampling(d->ufunc, d->udata, reinterpret_cast<const struct mat4x4*>(x),
			d->nlayers, d->nqubits, d->perms, d->nsamples, d->rng, &f, grad, hess) < 0) {
		fprintf(stderr, "target function and derivative evaluation failed internally\n");
		return -1;
	}
	return f.real();
}

}  // namespace

extern "C" void optimize_brickwall_circuit_hmat_sampling(linear_func ufunc, void* udata,
	const struct mat4x4 vlist_start[], const int nlayers, const int nqubits, const int* perms[],
	const long nsamples, struct rng_state* rng,
	struct rtr_params* params, const int niter, double* f_iter, struct mat4x4 vlist_opt[])
{
	BrickwallSamplingData fdata{ufunc, udata, perms, rng, nsamples, nlayers, nqubits};
	params->g_func = nullptr;
	params->g_data = nullptr;
	params->g_iter = nullptr;
	int rdata = nlayers;
	riemannian_trust_region_optimize_hmat(bws_f, bws_f_deriv_hess, &fdata, bw_retract, &rdata,
		nlayers * num_tangent_params, reinterpret_cast<const double*>(vlist_start), nlayers * 32, params, niter, f_iter,
		reinterpret_cast<double*>(vlist_opt));
}
