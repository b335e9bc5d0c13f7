/*
 * Seeded pseudo-random numbers for input generation (host side).
 * Replaces reference src/rng.h:14-69 and src/pcg_basic.h:31-60.
 * Generator: two PCG32 (XSH-RR) streams combined to 64 bits; uniform doubles
 * are u64 * 2^-64; normals by Box-Muller; complex normals (x + iy)/sqrt(2).
 */
#pragma once

#include <stdint.h>
#include "config.h"

/* PCG 32-bit generator state (reference src/pcg_basic.h:31-35) */
struct pcg32_random
{
	uint64_t state;
	uint64_t inc;
};

/* two 32-bit generators (reference src/pcg_basic.h:46-49) */
struct pcg32x2_random
{
	struct pcg32_random gen[2];
};

/* reference src/rng.h:14-17 */
struct rng_state
{
	struct pcg32x2_random pcgstate;
};

RQC_EXTERN_C_BEGIN

/* reference src/pcg_basic.h:38-60 */
void pcg32_srandom_r(const uint64_t initstate, const uint64_t initseq, struct pcg32_random* rng);
uint32_t pcg32_random_r(struct pcg32_random* rng);
uint32_t pcg32_boundedrand_r(const uint32_t bound, struct pcg32_random* rng);
void pcg32x2_srandom_r(uint64_t seed1, uint64_t seed2, uint64_t seq1, uint64_t seq2, struct pcg32x2_random* rng);
uint64_t pcg32x2_random_r(struct pcg32x2_random* rng);
uint64_t pcg32x2_boundedrand_r(const uint64_t bound, struct pcg32x2_random* rng);

/* reference src/rng.h:20-69 */
void seed_rng_state(const uint64_t seed, struct rng_state* state);
uint32_t rand_uint32(struct rng_state* state);
uint64_t rand_uint64(struct rng_state* state);
uint64_t rand_interval(const uint64_t bound, struct rng_state* state);
void rand_choice(const uint64_t bound, const uint64_t num_samples, struct rng_state* state, uint64_t* ret);
double randu(struct rng_state* state);
float randuf(struct rng_state* state);
double randn(struct rng_state* state);
float randnf(struct rng_state* state);
#if defined(__cplusplus) && defined(__clang__)
/* std::complex<double> has the layout and return convention of C99 double complex */
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Wreturn-type-c-linkage"
#endif
numeric crandn(struct rng_state* state);
#ifdef __cplusplus
std::complex<float> crandnf(struct rng_state* state);
#else
float complex crandnf(struct rng_state* state);
#endif
#if defined(__cplusplus) && defined(__clang__)
#pragma clang diagnostic pop
#endif

RQC_EXTERN_C_END
