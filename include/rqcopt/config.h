/*
 * rqcopt_hpc_amd — scalar types and allocation helpers of the C ABI.
 *
 * Replaces the reference's src/config.h:6-62 for the complex build
 * (COMPLEX_CIRCUIT): `numeric` is a double-precision complex number stored as
 * (re, im), `intqs` addresses statevector entries.  The headers are valid C99
 * (numeric = double complex) and C++ (numeric = std::complex<double>); both
 * have the same size, alignment and calling convention on x86-64 SysV.
 */
#pragma once

#include <stdlib.h>

#ifdef __cplusplus
#include <complex>
typedef std::complex<double> numeric;
#define RQC_RESTRICT __restrict
#define RQC_EXTERN_C_BEGIN extern "C" {
#define RQC_EXTERN_C_END }
#else
#include <complex.h>
typedef double complex numeric;
#define RQC_RESTRICT restrict
#define RQC_EXTERN_C_BEGIN
#define RQC_EXTERN_C_END
#endif

#ifndef COMPLEX_CIRCUIT
#define COMPLEX_CIRCUIT
#endif

/* alignment of host buffers allocated through this ABI (reference src/config.h:32) */
#define MEM_DATA_ALIGN 64

/* integer type for addressing statevector entries (reference src/config.h:62) */
typedef long int intqs;

/* 64-byte aligned host allocation; size is rounded up to a multiple of 64
   (same contract as the reference's src/config.h:36-45) */
static inline void* aligned_malloc(size_t size)
{
	return aligned_alloc(MEM_DATA_ALIGN, (size + MEM_DATA_ALIGN - 1) & ~((size_t)MEM_DATA_ALIGN - 1));
}

static inline void aligned_free(void* memblock)
{
	free(memblock);
}
