/*
 * Riemannian trust-region driver (Absil, Mahony, Sepulchre 2008, Alg. 10)
 * with Steihaug-Toint truncated CG (Alg. 11), in Hessian-matrix and
 * Hessian-vector-product form.  Replaces reference src/trust_region.h:10-85.  The driver is
 * host code; every f / f_deriv / f_hvp it calls runs on the GPU when it is
 * one of this library's targets.
 */
#pragma once

#include <stdbool.h>
#include "config.h"

/* reference src/trust_region.h:10-16 */
struct truncated_cg_params
{
	int maxiter;
	double abstol;
	double reltol;
};

/* reference src/trust_region.h:23-28 */
static inline void set_truncated_cg_default_params(int n, struct truncated_cg_params* params)
{
	params->maxiter = 2 * n;
	params->abstol  = 1e-8;
	params->reltol  = 1e-6;
}

typedef void (*hessian_vector_product_func)(const double* RQC_RESTRICT x, void* fdata, double* RQC_RESTRICT vec, double* RQC_RESTRICT hvp);
typedef double (*target_func)(const double* x, void* fdata);
typedef double (*target_gradient_func)(const double* RQC_RESTRICT x, void* fdata, double* RQC_RESTRICT grad);
typedef double (*target_gradient_hessian_func)(const double* RQC_RESTRICT x, void* fdata, double* RQC_RESTRICT grad, double* RQC_RESTRICT hess);
typedef void (*retract_func)(const double* RQC_RESTRICT x, const double* RQC_RESTRICT eta, void* rdata, double* RQC_RESTRICT xs);

/* reference src/trust_region.h:52-61 */
struct rtr_params
{
	struct truncated_cg_params tcg_params;
	double rho_trust;
	double radius_init;
	double maxradius;
	target_func g_func;
	void*       g_data;
	double*     g_iter;
};

/* reference src/trust_region.h:68-76 */
static inline void set_rtr_default_params(int n, struct rtr_params* params)
{
	set_truncated_cg_default_params(n, &params->tcg_params);
	params->rho_trust   = 0.125;
	params->radius_init = 0.01;
	params->maxradius   = 0.1;
	params->g_func      = NULL;
	params->g_data      = NULL;
	params->g_iter      = NULL;
}

RQC_EXTERN_C_BEGIN

/* reference src/trust_region.c:83-141: hess is the row-major n x n Hessian */
bool truncated_cg_hmat(const double* grad, const double* hess, int n, double radius, const struct truncated_cg_params* params, double* z);

/* reference src/trust_region.c:159-217 */
bool truncated_cg_hvp(const double* RQC_RESTRICT x, const double* RQC_RESTRICT grad, const hessian_vector_product_func f_hvp, void* fdata,
	const int n, const double radius, const struct truncated_cg_params* params, double* RQC_RESTRICT z);

/* reference src/trust_region.c:230-293 */
void riemannian_trust_region_optimize_hmat(target_func f, target_gradient_hessian_func f_deriv_hess, void* fdata, retract_func retract, void* rdata,
	const int n, const double* x_init, const int m, const struct rtr_params* params, const int niter, double* f_iter, double* x_final);

/* reference src/trust_region.c:306-368 */
void riemannian_trust_region_optimize_hvp(const target_func f, const target_gradient_func f_deriv, const hessian_vector_product_func f_hvp, void* fdata,
	retract_func retract, void* rdata,
	const int n, const double* x_init, const int m, const struct rtr_params* params, const int niter, double* f_iter, double* x_final);

RQC_EXTERN_C_END
