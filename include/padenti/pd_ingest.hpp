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
// A wave hashes a chunk of up to CRC_CHUNK consecutive blocks.
constexpr uint32_t CRC_CHUNK = 8;
constexpr uint32_t CRC_GAP = 4096 - 64;  // bytes between a lane's segments in consecutive blocks

struct CrcJob {
  uint64_t start, end, a0;
  uint32_t first_chunk, n_blocks;
};

// Largest i < n with first[i] <= v (first ascending, first[0] = 0).
__device__ __forceinline__ uint32_t find_owner(const uint32_t *__restrict__ first, uint32_t n, uint32_t v)
{
  uint32_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (first[mid] <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint32_t crc_seg(const uint32_t (*T)[256], uint32_t crc, const uint4 (&v)[4])
{
#pragma unroll
  for (int q = 0; q < 4; q++) {
    crc = crc_word(T, crc, v[q].x);
    crc = crc_word(T, crc, v[q].y);
    crc = crc_word(T, crc, v[q].z);
    crc = crc_word(T, crc, v[q].w);
  }
  return crc;
}

// a * x^(8 CRC_GAP) mod P from G[k][v] = (v at byte k) * x^(8 CRC_GAP)
__device__ __forceinline__ uint32_t mul_gap(const uint32_t (*G)[256], uint32_t a)
{
  return G[0][a & 0xFFu] ^ G[1][(a >> 8) & 0xFFu] ^ G[2][(a >> 16) & 0xFFu] ^ G[3][a >> 24];
}

// Wave per chunk, lane per 64-byte segment of every block (four 16-byte
// loads, the next block's prefetched while the current one is hashed).  A
// lane carries one register across the chunk's whole blocks: before each
// new segment it is shifted over the 4032 bytes of the other lanes (table
// multiply).  At the end of the whole blocks the lane registers are shifted
// to the block end (x^(8 * 64 * j)), xor-reduced, and shifted by
// P[n_blocks - 1 - e] = x^(8 * 4096 (n_blocks - 1 - e)) (e = the chunk's end
// block) into acc_a; the image's last block (lanes whose segment ends in the
// image, then the tail bytes) goes to acc_b.  k_crc_final applies the last
// block's length.  xor is order free, so the result is deterministic.
__global__ __launch_bounds__(256) void k_crc32(const CrcJob *__restrict__ jobs, const uint32_t *__restrict__ chunk0,
                                               uint32_t n_images, uint32_t n_chunks, const uint32_t *__restrict__ P,
                                               const uint32_t *__restrict__ gap, uint32_t *acc_a, uint32_t *acc_b)
{
  __shared__ uint32_t T[4][256], G[4][256];
  for (uint32_t i = threadIdx.x; i < 1024; i += 256) {
    T[i >> 8][i & 255] = c_crc.t[i >> 8][i & 255];
    G[i >> 8][i & 255] = gap[i];
  }
  __syncthreads();
  const uint32_t lane = threadIdx.x & 63u;
  for (uint32_t ch = blockIdx.x * 4u + (threadIdx.x >> 6); ch < n_chunks; ch += gridDim.x * 4u) {
    const uint32_t im = find_owner(chunk0, n_images, ch);
    const CrcJob J = jobs[im];
    const uint32_t m0 = (ch - J.first_chunk) * CRC_CHUNK;
    const uint32_t m1 = (m0 + CRC_CHUNK < J.n_blocks) ? m0 + CRC_CHUNK : J.n_blocks;
    const bool has_last = (m1 == J.n_blocks);
    const uint32_t mf = has_last ? J.n_blocks - 1u : m1;  // whole blocks [m0, mf)
    const bool head = (J.start > J.a0) && lane == 0;        // lane 0 of block 0 starts inside its segment
    if (m0 < mf) {
      uint4 v[4] = {make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0), make_uint4(0, 0, 0, 0)};
      if (!(m0 == 0 && head)) {
        const uint4 *p = reinterpret_cast<const uint4 *>(J.a0 + (uint64_t)m0 * 4096u + 64ull * lane);
        v[0] = p[0]; v[1] = p[1]; v[2] = p[2]; v[3] = p[3];
      }
      uint32_t reg = 0;
      for (uint32_t m = m0; m < mf; m++) {
        const uint4 cur[4] = {v[0], v[1], v[2], v[3]};
        if (m + 1u < mf) {
          const uint4 *p = reinterpret_cast<const uint4 *>(J.a0 + (uint64_t)(m + 1u) * 4096u + 64ull * lane);
          v[0] = p[0]; v[1] = p[1]; v[2] = p[2]; v[3] = p[3];
        }
        if (m > m0) reg = mul_gap(G, reg);
        if (m == 0 && head) {
          for (uint64_t a = J.start; a < J.a0 + 64u; a++) reg = crc_byte(T, reg, *reinterpret_cast<const uint8_t *>(a));
        } else {
          reg = crc_seg(T, reg, cur);
        }
      }
      reg = gf2_mul(reg, c_crc.w64[63u - lane]);
      for (int o = 32; o >= 1; o >>= 1) reg ^= __shfl_xor(reg, o);
      if (lane == 0) atomicXor(&acc_a[im], gf2_mul(reg, P[J.n_blocks - 1u - mf]));
    }
    if (has_last) {
      const uint32_t m = J.n_blocks - 1u;
      const uint64_t blo = J.a0 + (uint64_t)m * 4096u;
      const uint32_t L = (uint32_t)((J.end - blo) >> 6);  // lanes whose segment ends in the image
      uint32_t crc = 0;
      if (lane < L) {
        const uint64_t lo = blo + 64ull * lane;
        if (lo >= J.start) {
          const uint4 *p = reinterpret_cast<const uint4 *>(lo);
          const uint4 w[4] = {p[0], p[1], p[2], p[3]};
          crc = crc_seg(T, crc, w);
        } else {  // the image's first segment starts inside it
          for (uint64_t a = J.start; a < lo + 64u; a++) crc = crc_byte(T, crc, *reinterpret_cast<const uint8_t *>(a));
        }
        crc = gf2_mul(crc, c_crc.w64[L - 1u - lane]);
      }
      for (int o = 32; o >= 1; o >>= 1) crc ^= __shfl_xor(crc, o);
      const uint64_t t0 = blo + 64ull * L;
      for (uint64_t a = (t0 > J.start ? t0 : J.start); a < J.end; a++)
        crc = crc_byte(T, crc, *reinterpret_cast<const uint8_t *>(a));
      if (lane == 0) atomicXor(&acc_b[im], crc);
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

// bits[word0 + w] = labelled-pixel mask of pixels [32 w, 32 w + 32); cnt = its
// popcount.  A block per tile of 256 words of one image; tile0[i] = first
// tile of image i (binary search, n_images entries).
__global__ __launch_bounds__(256) void k_label_bits(const uint8_t *__restrict__ labels,
                                                    const SampleJob *__restrict__ jobs,
                                                    const uint32_t *__restrict__ tile0, uint32_t n_images,
                                                    uint32_t *bits, uint32_t *cnt)
{
  __shared__ uint32_t s_im;
  if (threadIdx.x == 0) s_im = find_owner(tile0, n_images, blockIdx.x);
  __syncthreads();
  const uint32_t im = s_im;
  const SampleJob J = jobs[im];
  const uint32_t w = (blockIdx.x - tile0[im]) * 256u + threadIdx.x;
  if (32ull * w >= J.n_px) return;
  const uint32_t g = J.word0 + w;
  const uint8_t *lab = labels + J.lab_off + 32ull * w;
  uint32_t m = 0;
  if (32ull * w + 32u <= J.n_px) {
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
// Ordering inside one wave: LDS operations of a wave execute in issue order,
// so a fence + wave barrier is all the synchronisation lanes of one wave need.
__device__ __forceinline__ void wave_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One twist of a state by one wave (no tempering: outputs are tempered when
// drawn).  Same three dependent ranges as mt_twist.
__device__ inline void mt_twist_wave(uint32_t *mt, uint32_t lane)
{
  uint32_t a1[4], a2[4], a3[3];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const uint32_t i = lane + 64u * r;
    a1[r] = a2[r] = 0;
    if (i < 227) {
      a1[r] = mt[i + MT_M] ^ mt_mix(mt[i], mt[i + 1]);
      a2[r] = mt_mix(mt[227 + i], mt[228 + i]);
    }
  }
#pragma unroll
  for (int r = 0; r < 3; r++) {
    const uint32_t i = lane + 64u * r;
    a3[r] = (i < 169) ? mt_mix(mt[454 + i], mt[455 + i]) : 0u;
  }
  const uint32_t last_old = mt[623];
  wave_sync();
#pragma unroll
  for (int r = 0; r < 4; r++)
    if (lane + 64u * r < 227) mt[lane + 64u * r] = a1[r];
  wave_sync();
#pragma unroll
  for (int r = 0; r < 4; r++)
    if (lane + 64u * r < 227) mt[227 + lane + 64u * r] = mt[lane + 64u * r] ^ a2[r];
  wave_sync();
#pragma unroll
  for (int r = 0; r < 3; r++)
    if (lane + 64u * r < 169) mt[454 + lane + 64u * r] = mt[227 + lane + 64u * r] ^ a3[r];
  if (lane == 0) mt[623] = mt[396] ^ mt_mix(last_old, mt[0]);
  wave_sync();
}

// One wave per image (four images per block), 64 pixels per step in scan
// order (lane l: pixel t0 + 64 s + l); the label bytes of the next 16 steps
// are prefetched from global memory.  A pixel's class ordinal k is the class
// count so far (LDS) plus its rank among the wave's lanes of that class (one
// ballot per distinct class, the class read with readlane); it draws iff
// k >= B_c, and its draw index is the wave prefix of those flags, so the MT
// stream is consumed in the serial loop's order.  bucket: zero-initialised
// slots [image][sum B] (class c at boff[c]) filled with atomicMax (the last
// writer in scan order has the largest pixel index).  Labels above
// n_classes set *bad.
//
// Bound by VALU issue: about 85 VALU + 43 SALU instructions per 64-pixel step
// (class ranks, the draw's FP64 product and roundf, tempering, a twist per
// 624 draws).  Keeping the slots in LDS instead of global atomics cut the
// occupancy and ran slower.
__global__ __launch_bounds__(256) void k_sample_weighted(const uint8_t *__restrict__ labels,
                                                         const SampleJob *__restrict__ jobs, uint32_t n_images,
                                                         const uint32_t *__restrict__ crc, uint32_t seed,
                                                         uint32_t n_classes, const uint32_t *__restrict__ B,
                                                         const uint32_t *__restrict__ boff, uint32_t *bucket,
                                                         uint32_t *class_count, uint32_t *bad)
{
  __shared__ uint32_t sB[256], sO[256];
  __shared__ uint32_t mt_all[4][MT_N], cnt_all[4][256];
  const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
  sB[threadIdx.x] = threadIdx.x < n_classes ? B[threadIdx.x] : 0u;
  sO[threadIdx.x] = threadIdx.x < n_classes ? boff[threadIdx.x] : 0u;
  __syncthreads();  // the only block-wide barrier: from here on the waves are independent
  const uint32_t im = blockIdx.x * 4u + wave;
  if (im >= n_images) return;
  uint32_t *mt = mt_all[wave], *cnt = cnt_all[wave];
  for (uint32_t c = lane; c < 256u; c += 64u) cnt[c] = 0;
  const SampleJob J = jobs[im];
  const uint8_t *lab = labels + J.lab_off;
  uint32_t *bk = bucket + J.out_base;
  if (lane == 0) mt_seed(mt, crc[im] ^ seed);
  wave_sync();
  const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64u - lane));
  uint32_t D = 0, gen = 0;  // draws consumed; outputs generated ([gen - 624, gen) are in mt)
  uint32_t nx[16];
#pragma unroll
  for (int s = 0; s < 16; s++) nx[s] = (64u * s + lane < J.n_px) ? lab[64u * s + lane] : 0u;
  for (uint32_t t0 = 0; t0 < J.n_px; t0 += 1024u) {
    uint32_t cur[16];
#pragma unroll
    for (int s = 0; s < 16; s++) {
      cur[s] = nx[s];
      const uint32_t q = t0 + 1024u + 64u * s + lane;
      nx[s] = (q < J.n_px) ? lab[q] : 0u;  // prefetch the next tile
    }
#pragma unroll
    for (int s = 0; s < 16; s++) {
      if (t0 + 64u * s >= J.n_px) break;
      const uint32_t p = t0 + 64u * s + lane;
      uint32_t l = cur[s];
      if (l > n_classes) { atomicOr(bad, 1u); l = 0; }
      const bool valid = l != 0u;
      const uint32_t c = l - 1u;
      uint32_t k = 0, Bc = 0, Oc = 0;
      uint64_t pending = __ballot(valid);
      while (pending) {
        const uint32_t lead = (uint32_t)__builtin_ctzll(pending);
        const uint32_t cl = __builtin_amdgcn_readlane(c, lead);
        const uint64_t mm = __ballot(valid && c == cl);
        const uint32_t before = cnt[cl], bcl = sB[cl], ocl = sO[cl];
        if (valid && c == cl) {
          k = before + (uint32_t)__popcll(mm & lt);
          Bc = bcl;
          Oc = ocl;
        }
        if (lane == lead) cnt[cl] = before + (uint32_t)__popcll(mm);
        pending &= ~mm;
      }
      const bool draw = valid && k >= Bc;
      const uint64_t dm = __ballot(draw);
      const uint32_t d = D + (uint32_t)__popcll(dm & lt);
      D += (uint32_t)__popcll(dm);
      uint32_t x = 0;
      if (D > gen) {  // wave-uniform: this step crosses into the next twist
        if (draw && d < gen) x = mt[d + MT_N - gen];
        mt_twist_wave(mt, lane);
        gen += MT_N;
        if (draw && d + MT_N >= gen) x = mt[d + MT_N - gen];
      } else if (draw) {
        x = mt[d + MT_N - gen];
      }
      if (draw) {
        const double u = (double)mt_temper(x) * 2.3283064365386962890625e-10;  // x / 2^32
        const uint32_t j = (uint32_t)roundf((float)(u * (double)k));
        if (j < Bc) atomicMax(&bk[Oc + j], p);
      } else if (valid) {
        atomicMax(&bk[Oc + k], p);
      }
    }
  }
  wave_sync();
  for (uint32_t c = lane; c < n_classes; c += 64u) class_count[(uint64_t)im * n_classes + c] = cnt[c];
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
