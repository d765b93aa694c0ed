/*
 * nraps_core.h — arithmetic shared by the HIP kernels and the host.
 *
 * Holds the three things whose bits must agree between the GPU and any CPU
 * replay of a history: the PCG32 generator of src/rand.rs, the mapping from
 * a 32-bit draw to a uniform variate, and the natural logarithm used for
 * flight lengths.  The file is valid C99 and HIP C++; in HIP every function
 * is usable from host and device code.
 *
 * Everything that includes this header must be compiled with
 * -ffp-contract=off so that no multiply-add is fused on one side only.
 */
#ifndef NRAPS_CORE_H
#define NRAPS_CORE_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define NRAPS_HD __host__ __device__ static inline __attribute__((always_inline))
#else
#define NRAPS_HD static inline
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* ---- PCG32 (PCG-XSH-RR 64/32), src/rand.rs:36-107 ----------------------- */
#define NRAPS_PCG_MULT 6364136223846793005ULL

typedef struct nraps_pcg32 {
  uint64_t state;
  uint64_t inc; /* always odd: selects the stream */
} nraps_pcg32;

/* One generator step and its 32-bit output (src/rand.rs:74-85). */
NRAPS_HD uint32_t nraps_pcg32_next(nraps_pcg32 *r) {
  uint64_t old = r->state;
  r->state = old * NRAPS_PCG_MULT + r->inc;
  uint32_t xorshifted = (uint32_t)(((old >> 18) ^ old) >> 27);
  uint32_t rot = (uint32_t)(old >> 59);
  return (xorshifted >> rot) | (xorshifted << ((32u - rot) & 31u));
}

/* Prn32::new with an explicit seed in place of the wall clock
   (src/rand.rs:49-71): state 0, inc = 2*stream+1, step, state += seed, step. */
NRAPS_HD void nraps_pcg32_seed(nraps_pcg32 *r, uint64_t seed, uint64_t stream) {
  r->state = 0u;
  r->inc = (stream << 1) | 1u;
  (void)nraps_pcg32_next(r);
  r->state = seed + r->inc; /* state == inc after the first step */
  (void)nraps_pcg32_next(r);
}

/* Id of particle p in generation `cycle` of problem `problem`.  Every
   history owns one generator, so its draws do not depend on scheduling,
   thread count or batch composition.  Requires problem < 2^19,
   cycle < 2^16, p < 2^28. */
NRAPS_HD uint64_t nraps_stream_id(uint64_t problem, uint64_t cycle, uint64_t p) {
  return (problem << 44) | (cycle << 28) | p;
}

/* SplitMix64 output function (Steele, Lea & Flood, "Fast splittable
   pseudorandom number generators", OOPSLA 2014, with the multipliers of
   D. Stafford's Mix13): a bijection of 64-bit words. */
NRAPS_HD uint64_t nraps_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
#define NRAPS_GOLDEN64 0x9E3779B97F4A7C15ULL

/* Generator of one history: PCG32 initialised as Prn32::new, with both the
   initial state and the stream scrambled from the history id.  Ids that
   differ only in high bits (the same particle in another generation) would
   otherwise give LCG states that agree in their low bits at every step. */
NRAPS_HD void nraps_history_rng(nraps_pcg32 *r, uint64_t seed, uint64_t id) {
  nraps_pcg32_seed(r, seed ^ nraps_mix64(id), nraps_mix64(id + NRAPS_GOLDEN64));
}
/* Per-(problem, cycle) service stream (fission-bank resampling offset). */
#define NRAPS_SERVICE_PARTICLE ((1u << 28) - 1u)
#define NRAPS_MAX_HISTORIES ((int64_t)NRAPS_SERVICE_PARTICLE - 1)

/* ---- uniform variates --------------------------------------------------- */
/* FP32 variate exactly as Prn32::next_f32 (src/rand.rs:95-100): the u32 is
   rounded to float and scaled by 2^-32, so the range is [0, 1]. */
NRAPS_HD float nraps_u01_f32(uint32_t u) { return (float)u * 0x1p-32f; }
/* FP64 variate: exact, in [0, 1). */
NRAPS_HD double nraps_u01_f64(uint32_t u) { return (double)u * 0x1p-32; }
/* Flight variate: as above but a zero draw becomes 2^-32 so -ln(U) stays
   finite (the reference would fly an infinite distance). */
NRAPS_HD float nraps_uflight_f32(uint32_t u) { return (float)(u ? u : 1u) * 0x1p-32f; }
NRAPS_HD double nraps_uflight_f64(uint32_t u) { return (double)(u ? u : 1u) * 0x1p-32; }
/* Index in [0, n) from one draw by multiply-shift (replaces gen_range). */
NRAPS_HD uint32_t nraps_uindex(uint32_t u, uint32_t n) {
  return (uint32_t)(((uint64_t)u * (uint64_t)n) >> 32);
}

/* ---- natural logarithm --------------------------------------------------
 * ln(x) for positive, finite, normal x, written here so host and device give
 * identical bits (a vendor log differs between the two in the last place).
 *
 * Method: x = 2^e * m with m in [sqrt(2)/2, sqrt(2)];  f = m - 1 is exact
 * (Sterbenz).  With s = f / (2 + f),  ln(1+f) = 2 atanh(s) = 2s + 2s*T(z),
 * z = s^2, T(z) = sum_{k>=1} z^k / (2k+1).  Since 2s = f - s*f,
 *     ln(1+f) = f - s*(f - 2*T(z)).
 * |s| <= 3 - 2*sqrt(2) = 0.1716, so z <= 0.02944 and ten terms of T leave a
 * truncation error below 2^-55 relative.  The coefficients are the exact
 * series coefficients 1/(2k+1).  ln 2 is split into a 42-bit head (e*hi is
 * exact for |e| < 2^11) and a tail, both rounded from the 60-digit value
 * 0.693147180559945309417232121458176568075500134360255254120680.
 * Error < 1.25 ulp against the exact logarithm (measured max 1.196 ulp over
 * 30 000 samples against a 50-digit log; bound asserted in the tests). */
#define NRAPS_LN2_HI 0x1.62e42fefa38p-1
#define NRAPS_LN2_LO 0x1.ef35793c7673p-45
#define NRAPS_SQRT2 0x1.6a09e667f3bcdp+0

NRAPS_HD double nraps_log(double x) {
  uint64_t b;
  __builtin_memcpy(&b, &x, sizeof b);
  int e = (int)((b >> 52) & 0x7ffu) - 1023;
  b = (b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL;
  double m;
  __builtin_memcpy(&m, &b, sizeof m);
  if (m > NRAPS_SQRT2) {
    m = m * 0.5;
    e = e + 1;
  }
  double f = m - 1.0;
  double s = f / (2.0 + f);
  double z = s * s;
  double t = 1.0 / 21.0;
  t = 1.0 / 19.0 + z * t;
  t = 1.0 / 17.0 + z * t;
  t = 1.0 / 15.0 + z * t;
  t = 1.0 / 13.0 + z * t;
  t = 1.0 / 11.0 + z * t;
  t = 1.0 / 9.0 + z * t;
  t = 1.0 / 7.0 + z * t;
  t = 1.0 / 5.0 + z * t;
  t = 1.0 / 3.0 + z * t;
  t = z * t;
  double lm = f - s * (f - 2.0 * t);
  double de = (double)e;
  return de * NRAPS_LN2_HI + (lm + de * NRAPS_LN2_LO);
}

#ifdef __cplusplus
}
#endif
#endif /* NRAPS_CORE_H */
