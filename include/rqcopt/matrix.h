/*
 * 4x4 complex matrices and the tangent-space maps of U(4).
 * Replaces reference src/matrix.h:10-70 (complex build).
 */
#pragma once

#include "config.h"

/* 4x4 matrix in row-major order (reference src/matrix.h:10-13) */
struct mat4x4
{
	numeric data[16];
};

/* number of real tangent-space parameters of U(4) (reference src/matrix.h:42-43) */
static const int num_tangent_params = 16;

RQC_EXTERN_C_BEGIN

/* reference src/matrix.c:11-113 */
void zero_matrix(struct mat4x4* a);
void identity_matrix(struct mat4x4* a);
void scale_matrix(struct mat4x4* RQC_RESTRICT a, const double x);
void conjugate_matrix(struct mat4x4* a);
void add_matrix(struct mat4x4* RQC_RESTRICT a, const struct mat4x4* RQC_RESTRICT b);
void sub_matrix(struct mat4x4* RQC_RESTRICT a, const struct mat4x4* RQC_RESTRICT b);
void add_matrices(const struct mat4x4* RQC_RESTRICT a, const struct mat4x4* RQC_RESTRICT b, struct mat4x4* RQC_RESTRICT c);
void sub_matrices(const struct mat4x4* RQC_RESTRICT a, const struct mat4x4* RQC_RESTRICT b, struct mat4x4* RQC_RESTRICT c);

/* reference src/matrix.c:119-144 */
void transpose(const struct mat4x4* RQC_RESTRICT a, struct mat4x4* RQC_RESTRICT at);
void adjoint(const struct mat4x4* RQC_RESTRICT a, struct mat4x4* RQC_RESTRICT ah);

/* reference src/matrix.c:150-211 */
void symm(const struct mat4x4* RQC_RESTRICT w, struct mat4x4* RQC_RESTRICT z);
void antisymm(const struct mat4x4* RQC_RESTRICT w, struct mat4x4* RQC_RESTRICT z);

/* reference src/matrix.c:218-329 */
void real_to_antisymm(const double* r, struct mat4x4* w);
void antisymm_to_real(const struct mat4x4* w, double* r);
void real_to_tangent(const double* r, const struct mat4x4* RQC_RESTRICT v, struct mat4x4* RQC_RESTRICT z);
void tangent_to_real(const struct mat4x4* RQC_RESTRICT v, const struct mat4x4* RQC_RESTRICT z, double* r);

/* reference src/matrix.c:336-347 */
void project_tangent(const struct mat4x4* RQC_RESTRICT u, const struct mat4x4* RQC_RESTRICT z, struct mat4x4* RQC_RESTRICT p);

/* reference src/matrix.c:356-388 */
void multiply_matrices(const struct mat4x4* RQC_RESTRICT a, const struct mat4x4* RQC_RESTRICT b, struct mat4x4* RQC_RESTRICT c);
void symmetric_triple_matrix_product(const struct mat4x4* RQC_RESTRICT a, const struct mat4x4* RQC_RESTRICT b,
	const struct mat4x4* RQC_RESTRICT c, struct mat4x4* RQC_RESTRICT ret);

/* reference src/matrix.c:410-531 */
int inverse_matrix(const struct mat4x4* RQC_RESTRICT a, struct mat4x4* RQC_RESTRICT ainv);
void polar_factor(const struct mat4x4* RQC_RESTRICT a, struct mat4x4* RQC_RESTRICT u);
void retract(const struct mat4x4* RQC_RESTRICT u, const double* RQC_RESTRICT eta, struct mat4x4* RQC_RESTRICT v);

RQC_EXTERN_C_END
