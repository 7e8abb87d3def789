// Dataset ingestion on the device: the two pixel samplers of the reference
// (UniformImageSampler, src/padenti/uniform_image_sampler_impl.hpp:34-63;
// ClassWeightedImageSampler, src/padenti/classweighted_image_sampler_impl.hpp:
// 57-123), the CRC-32 of each image's bytes that seeds them, and the RGB ->
// label mapping of the labels loader (src/padenti/cv_image_loader_impl.hpp:
// 235-267).
//
// Both samplers seed a Mersenne twister MT19937 per image with
// crc32(image bytes) ^ seed and draw from it in pixel-scan order; the device
// path reproduces every draw, so the sample lists equal the reference's.
//   uniform:        labelled pixels L (label != 0) in scan order; n_samples
//                   times r = uniform_int(0, |L|-1) (Boost's rejection form:
//                   bucket = floor(2^32 / |L|), accept x / bucket <= |L| - 1),
//                   sample L[r]; sort ascending (duplicates kept).
//   class-weighted: bucket size B_c = ceilf(w_c * (float)n_samples); per
//                   labelled pixel i of class c with class ordinal k: k < B_c
//                   fills slot k; otherwise one draw u = x * 2^-32 (double)
//                   and j = roundf((float)(u * k)), slot j <- i when j < B_c;
//                   output the filled slots of every class, sorted.
// The reservoir's last-writer-wins rule becomes a max over pixel indices,
// which is what makes the scan parallel: the draw index of a pixel is the
// number of earlier labelled pixels whose class ordinal is past its bucket.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pd {
namespace ingest {

// ---- CRC-32 (ISO-HDLC: reflected polynomial 0x04C11DB7, init and final xor
// 0xFFFFFFFF; the check value of "123456789" is 0xCBF43926).  Polynomials are
// held bit-reflected: bit 31 is the coefficient of x^0.
constexpr uint32_t CRC_POLY = 0xEDB88320u;
constexpr uint32_t CRC_ONE = 0x80000000u;  // the polynomial 1

// a * b mod P
__host__ __device__ constexpr uint32_t gf2_mul(uint32_t a, uint32_t b)
{
  uint32_t p = 0;
  for (int i = 0; i < 32; i++) {
    p ^= b & (0u - ((a >> (31 - i)) & 1u));
    b = (b >> 1) ^ (CRC_POLY & (0u - (b & 1u)));  // b *= x
  }
  return p;
}

struct CrcTables {
  uint32_t t[4][256];  // t[0]: one byte; t[s]: a byte followed by s zero bytes
  uint32_t x2k[72];    // x^(2^k) mod P
  uint32_t w64[64];    // x^(8 * 64 * j) mod P: j segments of 64 bytes
};

constexpr CrcTables make_crc_tables()
{
  CrcTables T{};
  for (uint32_t i = 0; i < 256; i++) {
    uint32_t c = i;
    for (int k = 0; k < 8; k++) c = (c >> 1) ^ (CRC_POLY & (0u - (c & 1u)));
    T.t[0][i] = c;
  }
  for (int s = 1; s < 4; s++)
    for (uint32_t i = 0; i < 256; i++) T.t[s][i] = (T.t[s - 1][i] >> 8) ^ T.t[0][T.t[s - 1][i] & 0xFFu];
  T.x2k[0] = CRC_ONE >> 1;  // x
  for (int k = 1; k < 72; k++) T.x2k[k] = gf2_mul(T.x2k[k - 1], T.x2k[k - 1]);
  T.w64[0] = CRC_ONE;
  for (int j = 1; j < 64; j++) T.w64[j] = gf2_mul(T.w64[j - 1], T.x2k[9]);  // * x^512
  return T;
}

__constant__ const CrcTables c_crc = make_crc_tables();

// x^(8n) mod P: the effect of n zero bytes on a CRC register
__host__ __device__ inline uint32_t crc_x8n(const CrcTables &T, uint64_t n)
{
  uint32_t r = CRC_ONE;
  for (int k = 3; n; n >>= 1, k++)
    if (n & 1u) r = gf2_mul(r, T.x2k[k]);
  return r;
}

// register form with zero initial value ("pure" CRC), four bytes at once
__device__ __forceinline__ uint32_t crc_word(const uint32_t (*T)[256], uint32_t crc, uint32_t w)
{
  crc ^= w;
  return T[3][crc & 0xFFu] ^ T[2][(crc >> 8) & 0xFFu] ^ T[1][(crc >> 16) & 0xFFu] ^ T[0][crc >> 24];
}

__device__ __forceinline__ uint32_t crc_byte(const uint32_t (*T)[256], uint32_t crc, uint32_t b)
{
  return T[0][(crc ^ b) & 0xFFu] ^ (crc >> 8);
}

// One image's byte range [start, end) (absolute addresses) cut into 4 KiB
// blocks from a0 = start & ~63; block m covers [a0 + 4096 m, a0 + 4096 (m+1)).
struct CrcJob {
  uint64_t start, end, a0;
  uint32_t first_block, n_blocks;
};

// Wave per 4 KiB block, lane per 64-byte segment (four 16-byte loads).
// Lane CRCs are shifted to the end of the block's last whole segment and
// xor-reduced; the last block continues the register over its tail bytes.
// Blocks before the last one are shifted by P[n_blocks - 2 - m] =
// x^(8 * 4096 (n_blocks - 2 - m)) into acc_a (the remaining shift, the
// last block's length, is applied once per image by k_crc_final); the last
// block goes to acc_b.  xor is order free, so the result is deterministic.
__global__ __launch_bounds__(256) void k_crc32(const CrcJob *__restrict__ jobs, const uint32_t *__restrict__ blk_img,
                                               uint32_t n_blocks, const uint32_t *__restrict__ P, uint32_t *acc_a,
                                               uint32_t *acc_b)
{
  __shared__ uint32_t T[4][256];
  for (uint32_t i = threadIdx.x; i < 1024; i += 256) T[i >> 8][i & 255] = c_crc.t[i >> 8][i & 255];
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t b = blockIdx.x * 4u + (threadIdx.x >> 6); b < n_blocks; b += gridDim.x * 4u) {
    const uint32_t im = blk_img[b];
    const CrcJob J = jobs[im];
    const uint32_t m = b - J.first_block;
    const bool last = (m + 1u == J.n_blocks);
    const uint64_t blo = J.a0 + (uint64_t)m * 4096u;
    const uint32_t L = last ? (uint32_t)((J.end - blo) >> 6) : 64u;  // lanes whose segment ends in the image
    uint32_t crc = 0;
    if (lane < L) {
      const uint64_t lo = blo + 64ull * lane;
      if (lo >= J.start) {
        const uint4 *p = reinterpret_cast<const uint4 *>(lo);
        const uint4 v0 = p[0], v1 = p[1], v2 = p[2], v3 = p[3];
        crc = crc_word(T, crc, v0.x); crc = crc_word(T, crc, v0.y);
        crc = crc_word(T, crc, v0.z); crc = crc_word(T, crc, v0.w);
        crc = crc_word(T, crc, v1.x); crc = crc_word(T, crc, v1.y);
        crc = crc_word(T, crc, v1.z); crc = crc_word(T, crc, v1.w);
        crc = crc_word(T, crc, v2.x); crc = crc_word(T, crc, v2.y);
        crc = crc_word(T, crc, v2.z); crc = crc_word(T, crc, v2.w);
        crc = crc_word(T, crc, v3.x); crc = crc_word(T, crc, v3.y);
        crc = crc_word(T, crc, v3.z); crc = crc_word(T, crc, v3.w);
      } else {  // the image's first segment starts inside it
        for (uint64_t a = J.start; a < lo + 64u; a++) crc = crc_byte(T, crc, *reinterpret_cast<const uint8_t *>(a));
      }
      crc = gf2_mul(crc, c_crc.w64[L - 1u - lane]);
    }
    for (int o = 32; o >= 1; o >>= 1) crc ^= __shfl_xor(crc, o);
    if (last) {
      const uint64_t t0 = blo + 64ull * L;
      for (uint64_t a = (t0 > J.start ? t0 : J.start); a < J.end; a++)
        crc = crc_byte(T, crc, *reinterpret_cast<const uint8_t *>(a));
      if (lane == 0) atomicXor(&acc_b[im], crc);
    } else {
      crc = gf2_mul(crc, P[J.n_blocks - 2u - m]);
      if (lane == 0) atomicXor(&acc_a[im], crc);
    }
  }
}

// crc32 = pure ^ (0xFFFFFFFF * x^(8 len)) ^ 0xFFFFFFFF, pure = acc_a * x^(8 q) ^ acc_b
// with q the length of the last block from a0.
__global__ __launch_bounds__(64) void k_crc_final(const CrcJob *__restrict__ jobs, uint32_t n_images,
                                                  const uint32_t *__restrict__ acc_a,
                                                  const uint32_t *__restrict__ acc_b, uint32_t *crc_out)
{
  const uint32_t i = blockIdx.x * 64u + threadIdx.x;
  if (i >= n_images) return;
  const CrcJob J = jobs[i];
  uint32_t pure = 0;
  if (J.n_blocks) {
    const uint64_t q = J.end - J.a0 - 4096ull * (J.n_blocks - 1u);
    pure = gf2_mul(acc_a[i], crc_x8n(c_crc, q)) ^ acc_b[i];
  }
  crc_out[i] = pure ^ gf2_mul(0xFFFFFFFFu, crc_x8n(c_crc, J.end - J.start)) ^ 0xFFFFFFFFu;
}

// ---- MT19937 (Matsumoto & Nishimura 1998), 256 threads share one state ----
constexpr uint32_t MT_N = 624, MT_M = 397;

__device__ __forceinline__ uint32_t mt_mix(uint32_t u, uint32_t v)
{
  const uint32_t y = (u & 0x80000000u) | (v & 0x7FFFFFFFu);
  return (y >> 1) ^ (0x9908B0DFu & (0u - (y & 1u)));
}

__device__ __forceinline__ uint32_t mt_temper(uint32_t y)
{
  y ^= y >> 11;
  y ^= (y << 7) & 0x9D2C5680u;
  y ^= (y << 15) & 0xEFC60000u;
  return y ^ (y >> 18);
}

// init_genrand(s): one thread (the recurrence is serial)
__device__ inline void mt_seed(uint32_t *mt, uint32_t s)
{
  mt[0] = s;
  for (uint32_t i = 1; i < MT_N; i++) {
    s = 1812433253u * (s ^ (s >> 30)) + i;
    mt[i] = s;
  }
}

// One twist of the whole state by the block (>= 256 threads, every thread
// calls it), then the 624 tempered outputs into out[0..623].  The serial
// recurrence new[i] = new/old[i+397 mod 624] ^ mix(old[i], old/new[i+1]) runs
// in three dependent ranges: [0,227) reads old words only, [227,454) and
// [454,624) read new[i-227] of the range before (i = 623 reads new[0]).
__device__ inline void mt_twist(uint32_t *mt, uint32_t *out, uint32_t tid)
{
  uint32_t a1 = 0, a2 = 0, a3 = 0, last_old = 0;
  if (tid < 227) {
    a1 = mt[tid + MT_M] ^ mt_mix(mt[tid], mt[tid + 1]);
    a2 = mt_mix(mt[227 + tid], mt[228 + tid]);
  }
  if (tid < 169) a3 = mt_mix(mt[454 + tid], mt[455 + tid]);
  if (tid == 169) last_old = mt[623];
  __syncthreads();
  if (tid < 227) mt[tid] = a1;
  __syncthreads();
  if (tid < 227) mt[227 + tid] = mt[tid] ^ a2;
  __syncthreads();
  if (tid < 169) mt[454 + tid] = mt[227 + tid] ^ a3;
  if (tid == 169) mt[623] = mt[396] ^ mt_mix(last_old, mt[0]);
  __syncthreads();
  for (uint32_t i = tid; i < MT_N; i += blockDim.x) out[i] = mt_temper(mt[i]);
  __syncthreads();
}

// ---- labelled-pixel rank index (uniform sampler) --------------------------
struct SampleJob {
  uint64_t lab_off;   // byte offset of the image's labels (16-byte aligned)
  uint32_t n_px;      // width * height
  uint32_t word0;     // first 32-pixel word of the image in the bitmask
  uint32_t out_base;  // first output slot of the image (image * capacity)
  uint32_t pad;
};

// bits[word0 + w] = labelled-pixel mask of pixels [32 w, 32 w + 32); cnt = its popcount
__global__ __launch_bounds__(256) void k_label_bits(const uint8_t *__restrict__ labels,
                                                    const SampleJob *__restrict__ jobs,
                                                    const uint32_t *__restrict__ word_img, uint32_t n_words,
                                                    uint32_t *bits, uint32_t *cnt)
{
  const uint32_t g = blockIdx.x * 256u + threadIdx.x;
  if (g >= n_words) return;
  const SampleJob J = jobs[word_img[g]];
  const uint32_t w = g - J.word0;
  const uint8_t *lab = labels + J.lab_off + 32ull * w;
  uint32_t m = 0;
  if (32u * w + 32u <= J.n_px) {
    const uint4 a = reinterpret_cast<const uint4 *>(lab)[0], b = reinterpret_cast<const uint4 *>(lab)[1];
    const uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int k = 0; k < 8; k++)
#pragma unroll
      for (int j = 0; j < 4; j++) m |= (((v[k] >> (8 * j)) & 0xFFu) != 0u ? 1u : 0u) << (4 * k + j);
  } else {
    for (uint32_t j = 0; 32u * w + j < J.n_px; j++) m |= (lab[j] != 0u ? 1u : 0u) << j;
  }
  bits[g] = m;
  cnt[g] = __popc(m);
}

// pixel of labelled rank r: largest word with pre <= r, then the bit
__device__ inline uint32_t select_rank(const uint32_t *bits, const uint32_t *pre, uint32_t word0, uint32_t n_words,
                                       uint32_t base, uint32_t r)
{
  uint32_t lo = 0, hi = n_words;  // pre[word0 + lo] - base <= r
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (pre[word0 + mid] - base <= r) lo = mid;
    else hi = mid;
  }
  uint32_t b = bits[word0 + lo];
  for (uint32_t k = r - (pre[word0 + lo] - base); k; k--) b &= b - 1u;
  return 32u * lo + (uint32_t)__builtin_ctz(b);
}

__device__ inline uint32_t block_excl_scan(uint32_t v, uint32_t *ws, uint32_t &total)
{
  // 256 threads: inclusive wave scan, wave totals in ws[0..3]
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  uint32_t s = v;
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(s, o);
    if (lane >= (uint32_t)o) s += t;
  }
  if (lane == 63) ws[wave] = s;
  __syncthreads();
  uint32_t before = 0;
  total = 0;
  for (uint32_t w = 0; w < 4; w++) {
    const uint32_t x = ws[w];
    if (w < wave) before += x;
    total += x;
  }
  __syncthreads();
  return before + s - v;
}

// One block of 256 threads per image.  pre = exclusive prefix of cnt over the
// whole batch.  Writes the unsorted samples and the segment end for the sort.
__global__ __launch_bounds__(256) void k_sample_uniform(const SampleJob *__restrict__ jobs,
                                                        const uint32_t *__restrict__ bits,
                                                        const uint32_t *__restrict__ pre,
                                                        const uint32_t *__restrict__ crc, uint32_t seed,
                                                        uint32_t n_samples, uint32_t *out, uint32_t *n_out,
                                                        uint32_t *seg_end)
{
  __shared__ uint32_t mt[MT_N], y[MT_N], ws[4];
  const uint32_t im = blockIdx.x, tid = threadIdx.x;
  const SampleJob J = jobs[im];
  const uint32_t n_words = (J.n_px + 31u) >> 5;
  const uint32_t base = pre[J.word0];
  const uint32_t n_lab = pre[J.word0 + n_words] - base;  // pre has one entry past the batch
  uint32_t *o = out + J.out_base;
  if (n_lab == 0u) {  // no labelled pixel: no sample (the reference's range would be empty)
    if (tid == 0) { n_out[im] = 0; seg_end[im] = J.out_base; }
    return;
  }
  if (n_lab == 1u) {  // uniform_int over one value draws nothing
    const uint32_t px = select_rank(bits, pre, J.word0, n_words, base, 0u);
    for (uint32_t s = tid; s < n_samples; s += 256u) o[s] = px;
    if (tid == 0) { n_out[im] = n_samples; seg_end[im] = J.out_base + n_samples; }
    return;
  }
  const uint32_t range = n_lab - 1u;
  const uint32_t bucket = (uint32_t)(0x100000000ull / n_lab);
  if (tid == 0) mt_seed(mt, crc[im] ^ seed);
  __syncthreads();
  uint32_t filled = 0;
  while (filled < n_samples) {
    mt_twist(mt, y, tid);
    uint32_t r[3], acc = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const uint32_t i = 3u * tid + k;
      r[k] = (i < MT_N) ? y[i] / bucket : 0xFFFFFFFFu;
      acc += (r[k] <= range) ? 1u : 0u;
    }
    uint32_t total;
    uint32_t slot = filled + block_excl_scan(acc, ws, total);
#pragma unroll
    for (int k = 0; k < 3; k++)
      if (r[k] <= range) {
        if (slot < n_samples) o[slot] = select_rank(bits, pre, J.word0, n_words, base, r[k]);
        slot++;
      }
    filled += total;
  }
  if (tid == 0) { n_out[im] = n_samples; seg_end[im] = J.out_base + n_samples; }
}

// ---- class-weighted reservoir --------------------------------------------
// One block of 256 threads per image, 256 pixels per step in scan order.
// bucket: zero-initialised slots [image][sum B] (slot of class c at boff[c]),
// filled with atomicMax (the last writer in scan order has the largest pixel
// index).  Labels above n_classes set *bad.
__global__ __launch_bounds__(256) void k_sample_weighted(const uint8_t *__restrict__ labels,
                                                         const SampleJob *__restrict__ jobs,
                                                         const uint32_t *__restrict__ crc, uint32_t seed,
                                                         uint32_t n_classes, const uint32_t *__restrict__ B,
                                                         const uint32_t *__restrict__ boff, uint32_t *bucket,
                                                         uint32_t *class_count, uint32_t *bad)
{
  __shared__ uint32_t mt[MT_N], y[2 * MT_N];
  __shared__ uint32_t cnt[256], cntw[4][256], sB[256], dw[4];
  __shared__ uint4 tile[256];
  const uint32_t im = blockIdx.x, tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
  const SampleJob J = jobs[im];
  const uint8_t *lab = labels + J.lab_off;
  uint32_t *bk = bucket + J.out_base;
  cnt[tid] = 0;
  cntw[0][tid] = cntw[1][tid] = cntw[2][tid] = cntw[3][tid] = 0;
  sB[tid] = tid < n_classes ? B[tid] : 0u;
  if (tid == 0) mt_seed(mt, crc[im] ^ seed);
  __syncthreads();
  uint32_t D = 0, gen_end = 0;  // draws consumed; outputs generated so far
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
  for (uint32_t t0 = 0; t0 < J.n_px; t0 += 4096u) {
    {  // stage 4096 label bytes
      const uint32_t b0 = t0 + 16u * tid;
      uint4 v = make_uint4(0, 0, 0, 0);
      if (b0 + 16u <= J.n_px) {
        v = reinterpret_cast<const uint4 *>(lab + b0)[0];
      } else if (b0 < J.n_px) {
        uint32_t w[4] = {0, 0, 0, 0};
        for (uint32_t j = 0; b0 + j < J.n_px; j++) w[j >> 2] |= (uint32_t)lab[b0 + j] << (8 * (j & 3));
        v = make_uint4(w[0], w[1], w[2], w[3]);
      }
      tile[tid] = v;
    }
    __syncthreads();
    const uint8_t *tb = reinterpret_cast<const uint8_t *>(tile);
    for (uint32_t s = 0; s < 16u && t0 + 256u * s < J.n_px; s++) {
      const uint32_t p = t0 + 256u * s + tid;
      uint32_t l = (p < J.n_px) ? tb[256u * s + tid] : 0u;
      if (l > n_classes) { atomicOr(bad, 1u); l = 0; }
      const bool valid = l != 0u;
      const uint32_t c = l - 1u;
      // class ordinal inside the wave: one ballot per distinct class
      uint64_t pending = __ballot(valid);
      uint32_t rank = 0;
      bool leader = false;
      while (pending) {
        const uint32_t lead = (uint32_t)__builtin_ctzll(pending);
        const uint32_t cl = __shfl(c, lead);
        const uint64_t mm = __ballot(valid && c == cl);
        if (valid && c == cl) rank = (uint32_t)__popcll(mm & lt);
        if (lane == lead) { cntw[wave][cl] = (uint32_t)__popcll(mm); leader = true; }
        pending &= ~mm;
      }
      __syncthreads();
      uint32_t k = 0;
      if (valid) {
        k = cnt[c] + rank;
        for (uint32_t w = 0; w < wave; w++) k += cntw[w][c];
      }
      const bool draw = valid && k >= sB[c];
      const uint64_t dm = __ballot(draw);
      if (lane == 0) dw[wave] = (uint32_t)__popcll(dm);
      __syncthreads();
      uint32_t d = D + (uint32_t)__popcll(dm & lt), total = 0;
      for (uint32_t w = 0; w < 4; w++) {
        if (w < wave) d += dw[w];
        total += dw[w];
      }
      if (D + total > gen_end) {  // block-uniform: one twist covers the <= 256 draws of a step
        mt_twist(mt, y + (gen_end % (2u * MT_N)), tid);
        gen_end += MT_N;
      }
      if (draw) {
        const double u = (double)y[d % (2u * MT_N)] * 2.3283064365386962890625e-10;  // x / 2^32
        const uint32_t j = (uint32_t)roundf((float)(u * (double)k));
        if (j < sB[c]) atomicMax(&bk[boff[c] + j], p);
      } else if (valid) {
        atomicMax(&bk[boff[c] + k], p);
      }
      D += total;
      __syncthreads();
      if (leader) {
        atomicAdd(&cnt[c], cntw[wave][c]);
        cntw[wave][c] = 0;
      }
    }
    __syncthreads();
  }
  if (tid < n_classes) class_count[(uint64_t)im * n_classes + tid] = cnt[tid];
}

// filled slots of every class -> contiguous unsorted samples; segment end
__global__ __launch_bounds__(256) void k_compact_buckets(const SampleJob *__restrict__ jobs, uint32_t n_classes,
                                                         const uint32_t *__restrict__ B,
                                                         const uint32_t *__restrict__ boff,
                                                         const uint32_t *__restrict__ class_count,
                                                         const uint32_t *__restrict__ bucket, uint32_t *out,
                                                         uint32_t *n_out, uint32_t *seg_end)
{
  const uint32_t im = blockIdx.x, tid = threadIdx.x;
  const SampleJob J = jobs[im];
  uint32_t dst = 0;
  for (uint32_t c = 0; c < n_classes; c++) {
    const uint32_t cc = class_count[(uint64_t)im * n_classes + c];
    const uint32_t n = cc < B[c] ? cc : B[c];
    for (uint32_t j = tid; j < n; j += 256u) out[J.out_base + dst + j] = bucket[J.out_base + boff[c] + j];
    dst += n;
  }
  if (tid == 0) { n_out[im] = dst; seg_end[im] = J.out_base + dst; }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v)
{
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += (uint64_t)gridDim.x * 256u) p[i] = v;
}

// ---- RGB(A) -> label: the first class whose colour equals the pixel's RGB
// (alpha ignored) gives label class + 1, no match gives 0.  keys: the map's
// distinct colours (r << 16 | g << 8 | b) ascending with the first class of
// each, so a binary search finds the first match.
__global__ __launch_bounds__(256) void k_rgb_labels(const uint8_t *__restrict__ rgb, uint64_t n_px, uint32_t ch,
                                                    const uint32_t *__restrict__ keys,
                                                    const uint8_t *__restrict__ cls, uint32_t n_keys,
                                                    uint8_t *__restrict__ out)
{
  __shared__ uint32_t sk[256];
  __shared__ uint8_t sc[256];
  if (threadIdx.x < n_keys) { sk[threadIdx.x] = keys[threadIdx.x]; sc[threadIdx.x] = cls[threadIdx.x]; }
  __syncthreads();
  for (uint64_t i = blockIdx.x * 256ull + threadIdx.x; i < n_px; i += (uint64_t)gridDim.x * 256u) {
    const uint8_t *p = rgb + i * ch;
    const uint32_t key = ((uint32_t)p[0] << 16) | ((uint32_t)p[1] << 8) | p[2];
    uint32_t lo = 0, hi = n_keys;  // first key >= key
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (sk[mid] < key) lo = mid + 1;
      else hi = mid;
    }
    out[i] = (lo < n_keys && sk[lo] == key) ? (uint8_t)(sc[lo] + 1u) : (uint8_t)0;
  }
}

}  // namespace ingest
}  // namespace pd
