// Host/device arithmetic of the random-forest hot path: the MD5 keyed
// sampler, the candidate parameter / threshold formula, the deterministic
// log2 and the FP32 information gain.  Everything here is compiled for the
// host (g++/clang) and for gfx950 (hipcc) with -ffp-contract=off, so both
// sides round after every operation in the order written.
//
// Reference: src/padenti/kernels/hist_update.cl (md5Rand :278-369, parameter
// and threshold sampling :62-184), src/padenti/prng.hpp:107-195 (host md5Rand),
// src/padenti/kernels/learn_best_feature.cl:20-111 (gain),
// src/padenti/cl_tree_trainer_impl_learn_best_featthr.hpp:183-222 (host
// re-derivation of the winning parameters).
#pragma once
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define PD_HD __host__ __device__ __forceinline__
#else
#define PD_HD inline
#endif

#pragma clang fp contract(off)

namespace pd {

// ---------------------------------------------------------------------------
// md5Rand(seed) = MD5 digest of the 16 little-endian seed bytes, i.e. the
// single block (s0, s1, s2, s3, 0x80, 0 x 9, 128, 0).  Fully unrolled with
// compile-time message indices so the zero words cost nothing.
// ---------------------------------------------------------------------------
PD_HD uint32_t pd_rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }

struct U4 { uint32_t x, y, z, w; };

#define PD_MD5_STEP(fn, a, b, c, d, m, s, k) \
  a = b + pd_rotl(a + fn(b, c, d) + (m) + (uint32_t)(k), s)
#define PD_F(b, c, d) (((b) & (c)) | (~(b) & (d)))
#define PD_G(b, c, d) (((b) & (d)) | ((c) & ~(d)))
#define PD_H(b, c, d) ((b) ^ (c) ^ (d))
#define PD_I(b, c, d) ((c) ^ ((b) | ~(d)))

PD_HD U4 md5rand(U4 s)
{
  const uint32_t m0 = s.x, m1 = s.y, m2 = s.z, m3 = s.w, m4 = 0x80u, m14 = 128u;
  uint32_t a = 0x67452301u, b = 0xefcdab89u, c = 0x98badcfeu, d = 0x10325476u;
  // round 1: message index i
  PD_MD5_STEP(PD_F, a, b, c, d, m0, 7, 0xd76aa478u);
  PD_MD5_STEP(PD_F, d, a, b, c, m1, 12, 0xe8c7b756u);
  PD_MD5_STEP(PD_F, c, d, a, b, m2, 17, 0x242070dbu);
  PD_MD5_STEP(PD_F, b, c, d, a, m3, 22, 0xc1bdceeeu);
  PD_MD5_STEP(PD_F, a, b, c, d, m4, 7, 0xf57c0fafu);
  PD_MD5_STEP(PD_F, d, a, b, c, 0u, 12, 0x4787c62au);
  PD_MD5_STEP(PD_F, c, d, a, b, 0u, 17, 0xa8304613u);
  PD_MD5_STEP(PD_F, b, c, d, a, 0u, 22, 0xfd469501u);
  PD_MD5_STEP(PD_F, a, b, c, d, 0u, 7, 0x698098d8u);
  PD_MD5_STEP(PD_F, d, a, b, c, 0u, 12, 0x8b44f7afu);
  PD_MD5_STEP(PD_F, c, d, a, b, 0u, 17, 0xffff5bb1u);
  PD_MD5_STEP(PD_F, b, c, d, a, 0u, 22, 0x895cd7beu);
  PD_MD5_STEP(PD_F, a, b, c, d, 0u, 7, 0x6b901122u);
  PD_MD5_STEP(PD_F, d, a, b, c, 0u, 12, 0xfd987193u);
  PD_MD5_STEP(PD_F, c, d, a, b, m14, 17, 0xa679438eu);
  PD_MD5_STEP(PD_F, b, c, d, a, 0u, 22, 0x49b40821u);
  // round 2: message index (5i+1) mod 16
  PD_MD5_STEP(PD_G, a, b, c, d, m1, 5, 0xf61e2562u);
  PD_MD5_STEP(PD_G, d, a, b, c, 0u, 9, 0xc040b340u);
  PD_MD5_STEP(PD_G, c, d, a, b, 0u, 14, 0x265e5a51u);
  PD_MD5_STEP(PD_G, b, c, d, a, m0, 20, 0xe9b6c7aau);
  PD_MD5_STEP(PD_G, a, b, c, d, 0u, 5, 0xd62f105du);
  PD_MD5_STEP(PD_G, d, a, b, c, 0u, 9, 0x02441453u);
  PD_MD5_STEP(PD_G, c, d, a, b, 0u, 14, 0xd8a1e681u);
  PD_MD5_STEP(PD_G, b, c, d, a, m4, 20, 0xe7d3fbc8u);
  PD_MD5_STEP(PD_G, a, b, c, d, 0u, 5, 0x21e1cde6u);
  PD_MD5_STEP(PD_G, d, a, b, c, m14, 9, 0xc33707d6u);
  PD_MD5_STEP(PD_G, c, d, a, b, m3, 14, 0xf4d50d87u);
  PD_MD5_STEP(PD_G, b, c, d, a, 0u, 20, 0x455a14edu);
  PD_MD5_STEP(PD_G, a, b, c, d, 0u, 5, 0xa9e3e905u);
  PD_MD5_STEP(PD_G, d, a, b, c, m2, 9, 0xfcefa3f8u);
  PD_MD5_STEP(PD_G, c, d, a, b, 0u, 14, 0x676f02d9u);
  PD_MD5_STEP(PD_G, b, c, d, a, 0u, 20, 0x8d2a4c8au);
  // round 3: message index (3i+5) mod 16
  PD_MD5_STEP(PD_H, a, b, c, d, 0u, 4, 0xfffa3942u);
  PD_MD5_STEP(PD_H, d, a, b, c, 0u, 11, 0x8771f681u);
  PD_MD5_STEP(PD_H, c, d, a, b, 0u, 16, 0x6d9d6122u);
  PD_MD5_STEP(PD_H, b, c, d, a, m14, 23, 0xfde5380cu);
  PD_MD5_STEP(PD_H, a, b, c, d, m1, 4, 0xa4beea44u);
  PD_MD5_STEP(PD_H, d, a, b, c, m4, 11, 0x4bdecfa9u);
  PD_MD5_STEP(PD_H, c, d, a, b, 0u, 16, 0xf6bb4b60u);
  PD_MD5_STEP(PD_H, b, c, d, a, 0u, 23, 0xbebfbc70u);
  PD_MD5_STEP(PD_H, a, b, c, d, 0u, 4, 0x289b7ec6u);
  PD_MD5_STEP(PD_H, d, a, b, c, m0, 11, 0xeaa127fau);
  PD_MD5_STEP(PD_H, c, d, a, b, m3, 16, 0xd4ef3085u);
  PD_MD5_STEP(PD_H, b, c, d, a, 0u, 23, 0x04881d05u);
  PD_MD5_STEP(PD_H, a, b, c, d, 0u, 4, 0xd9d4d039u);
  PD_MD5_STEP(PD_H, d, a, b, c, 0u, 11, 0xe6db99e5u);
  PD_MD5_STEP(PD_H, c, d, a, b, 0u, 16, 0x1fa27cf8u);
  PD_MD5_STEP(PD_H, b, c, d, a, m2, 23, 0xc4ac5665u);
  // round 4: message index 7i mod 16
  PD_MD5_STEP(PD_I, a, b, c, d, m0, 6, 0xf4292244u);
  PD_MD5_STEP(PD_I, d, a, b, c, 0u, 10, 0x432aff97u);
  PD_MD5_STEP(PD_I, c, d, a, b, m14, 15, 0xab9423a7u);
  PD_MD5_STEP(PD_I, b, c, d, a, 0u, 21, 0xfc93a039u);
  PD_MD5_STEP(PD_I, a, b, c, d, 0u, 6, 0x655b59c3u);
  PD_MD5_STEP(PD_I, d, a, b, c, m3, 10, 0x8f0ccc92u);
  PD_MD5_STEP(PD_I, c, d, a, b, 0u, 15, 0xffeff47du);
  PD_MD5_STEP(PD_I, b, c, d, a, m1, 21, 0x85845dd1u);
  PD_MD5_STEP(PD_I, a, b, c, d, 0u, 6, 0x6fa87e4fu);
  PD_MD5_STEP(PD_I, d, a, b, c, 0u, 10, 0xfe2ce6e0u);
  PD_MD5_STEP(PD_I, c, d, a, b, 0u, 15, 0xa3014314u);
  PD_MD5_STEP(PD_I, b, c, d, a, 0u, 21, 0x4e0811a1u);
  PD_MD5_STEP(PD_I, a, b, c, d, m4, 6, 0xf7537e82u);
  PD_MD5_STEP(PD_I, d, a, b, c, 0u, 10, 0xbd3af235u);
  PD_MD5_STEP(PD_I, c, d, a, b, m2, 15, 0x2ad7d2bbu);
  PD_MD5_STEP(PD_I, b, c, d, a, 0u, 21, 0xeb86d391u);
  U4 o;
  o.x = a + 0x67452301u;
  o.y = b + 0xefcdab89u;
  o.z = c + 0x98badcfeu;
  o.w = d + 0x10325476u;
  return o;
}
#undef PD_MD5_STEP
#undef PD_F
#undef PD_G
#undef PD_H
#undef PD_I

PD_HD uint32_t u4_get(const U4 &s, uint32_t i)
{
  return i == 0 ? s.x : (i == 1 ? s.y : (i == 2 ? s.z : s.w));
}

// ---------------------------------------------------------------------------
// Candidate sampling: lo + (T)(((float)w)/0xFFFFFFFF * (hi - lo)).
// (float)0xFFFFFFFF is 2^32, and dividing by a power of two is exact, so the
// quotient is computed as w * 2^-32.  For an integer FeatType the span is
// an int, the product a float, and the cast truncates (C conversion rules,
// the same as OpenCL C's for these scalar types).
// ---------------------------------------------------------------------------
template <typename T>
PD_HD T sample_value(uint32_t w, T lo, T hi)
{
  float u = (float)w * 0x1p-32f;
  auto span = hi - lo;       // float for float, int for short
  auto q = u * span;         // float
  return (T)(lo + (T)q);
}

// Parameters of candidate f of node `node` in tree `tree`:
// chain from seed (tree, node, f, 0), one md5 per four entries.
template <typename T, int FD>
PD_HD void candidate_params(uint32_t tree, uint32_t node, uint32_t f, const T *lo, const T *hi,
                            T *out)
{
  U4 s = {tree, node, f, 0u};
#pragma unroll
  for (int j = 0; j < FD; j += 4) {
    s = md5rand(s);
#pragma unroll
    for (int q = 0; q < 4 && j + q < FD; q++) out[j + q] = sample_value<T>(u4_get(s, q), lo[j + q], hi[j + q]);
  }
}

// Threshold t of candidate f: chain from seed (tree, node, f, 1), md5 number
// t/4 + 1, word t % 4.
template <typename T>
PD_HD T candidate_threshold(uint32_t tree, uint32_t node, uint32_t f, uint32_t t, T lo, T hi)
{
  U4 s = {tree, node, f, 1u};
  for (uint32_t j = 0; j <= t / 4; j++) s = md5rand(s);
  return sample_value<T>(u4_get(s, t & 3), lo, hi);
}

// ---------------------------------------------------------------------------
// Deterministic log2 for the gain.  x = m * 2^e with m in [sqrt(1/2), sqrt(2));
// s = (m-1)/(m+1); ln m = 2 s (1 + s^2/3 + s^4/5 + ... + s^22/23) in FP64
// (Horner, |s| < 0.172, truncation < 1e-18); log2 x = (float)(e + ln m * log2 e).
// Only IEEE add / mul / div in a fixed order: bit-identical on host and device.
// Valid for positive normal x (the gain calls it for x > 1e-6 only).
// ---------------------------------------------------------------------------
PD_HD float log2_det(float x)
{
  uint32_t b;
  memcpy(&b, &x, 4);
  int e = (int)((b >> 23) & 0xffu) - 127;
  uint32_t mb = (b & 0x007fffffu) | 0x3f800000u;
  float mf;
  memcpy(&mf, &mb, 4);
  double m = (double)mf;
  if (m > 1.4142135623730951) {
    m = m * 0.5;
    e += 1;
  }
  double s = (m - 1.0) / (m + 1.0);
  double s2 = s * s;
  double p = 1.0 / 23.0;
  p = p * s2 + 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  p = p * s2 + 1.0;
  double lnm = 2.0 * s * p;
  double r = (double)e + lnm * 1.4426950408889634;
  return (float)r;
}

// -p*log2(p) contribution with the reference's EPS rule (EPS is the double
// literal 1.e-6, so the float is compared after promotion).
PD_HD float plogp(float tmp) { return ((double)tmp <= 1.e-6) ? 0.0f : tmp * log2_det(tmp); }

// Node entropy h = -sum_c plogp(tot_c / N), in the order learnBestFeature
// accumulates it (learn_best_feature.cl:63-76).
PD_HD float slot_entropy(const uint32_t *tot, uint32_t nC)
{
  uint32_t N = 0;
  for (uint32_t c = 0; c < nC; c++) N += tot[c];
  const float fN = (float)(int)N;
  float h = 0.0f;
  for (uint32_t c = 0; c < nC; c++) h -= plogp((float)tot[c] / fN);
  return h;
}

// ---------------------------------------------------------------------------
// FP32 information gain of one split, op for op as learnBestFeature
// (learn_best_feature.cl:42-93): left[c] = histogram count, tot[c] = node's
// per-class totals.
// ---------------------------------------------------------------------------
template <int NC_MAX>
PD_HD float gain_f32(const uint32_t *left, const uint32_t *tot, uint32_t nC)
{
  int nL = 0, nR = 0;
  for (uint32_t l = 0; l < nC; l++) {
    nL += (int)left[l];
    nR += (int)(tot[l] - left[l]);
  }
  float h = 0.0f, hL = 0.0f, hR = 0.0f;
  float fN = (float)(nL + nR), fL = (float)nL, fR = (float)nR;
  for (uint32_t l = 0; l < nC; l++) {
    float tmp = (float)tot[l] / fN;
    h -= plogp(tmp);
    tmp = nL ? (float)left[l] / fL : 0.0f;
    hL -= plogp(tmp);
    tmp = nR ? (float)(tot[l] - left[l]) / fR : 0.0f;
    hR -= plogp(tmp);
  }
  float a = fL / fN * hL;
  float b = fR / fN * hR;
  return h - (a + b);
}

// Orderable key for an argmax with lowest-index tie-break: larger key wins.
// High word: gain bits mapped so that unsigned order = float order (-0 is
// folded onto +0 first); low word: ~k so that the smaller k wins a tie.
PD_HD uint64_t split_key(float gain, uint32_t k)
{
  gain = gain + 0.0f;
  uint32_t b;
  memcpy(&b, &gain, 4);
  uint32_t o = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  return ((uint64_t)o << 32) | (uint64_t)(0xffffffffu - k);
}
PD_HD float key_gain(uint64_t key)
{
  uint32_t o = (uint32_t)(key >> 32);
  uint32_t b = (o & 0x80000000u) ? (o & 0x7fffffffu) : ~o;
  float g;
  memcpy(&g, &b, 4);
  return g;
}
PD_HD uint32_t key_index(uint64_t key) { return 0xffffffffu - (uint32_t)(key & 0xffffffffu); }

}  // namespace pd
