// philox.h -- Philox4x64-10 counter-based generator (Salmon, Moraes, Dror,
// Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC'11), usable from
// host and device code.  One call yields four 64-bit words; a word maps to
// a uniform in the open interval (0,1) with 52 random bits.
//
// Counter layout used by the spark step:
//   ctr = {step, channel, stream, purpose}, key = {seed, SRS_PHILOX_KEY1}
// purpose 0: gating draws at `step` (word 0 -> new residual, word 1 -> LCC
//            transition choice); 1: initial residuals; 2: LCC-trigger close
//            time; 3: random initial RyR.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define SRS_HD __host__ __device__ inline
#else
#define SRS_HD inline
#endif

#define SRS_PHILOX_KEY1 0x5352535f414d4431ULL

enum { SRS_PURPOSE_GATING = 0, SRS_PURPOSE_INIT_R = 1, SRS_PURPOSE_LCC_CLOSE = 2, SRS_PURPOSE_RANDOM_RYR = 3 };

SRS_HD void srs_mulhilo64(uint64_t a, uint64_t b, uint64_t* hi, uint64_t* lo) {
#if defined(__HIP_DEVICE_COMPILE__)
  *lo = a * b;
  *hi = __umul64hi(a, b);
#else
  unsigned __int128 p = (unsigned __int128)a * b;
  *lo = (uint64_t)p;
  *hi = (uint64_t)(p >> 64);
#endif
}

SRS_HD void srs_philox4x64_10(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3, uint64_t k0, uint64_t k1,
                              uint64_t out[4]) {
  const uint64_t M0 = 0xD2E7470EE14C6C93ULL, M1 = 0xCA5A826395121157ULL;
  const uint64_t W0 = 0x9E3779B97F4A7C15ULL, W1 = 0xBB67AE8584CAA73BULL;
#if defined(__HIPCC__)
#pragma unroll
#endif
  for (int r = 0; r < 10; ++r) {
    uint64_t hi0, lo0, hi1, lo1;
    srs_mulhilo64(M0, c0, &hi0, &lo0);
    srs_mulhilo64(M1, c2, &hi1, &lo1);
    const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
    c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    k0 += W0; k1 += W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

SRS_HD double srs_open01(uint64_t x) { return ((double)(x >> 12) + 0.5) * (1.0 / 4503599627370496.0); }
