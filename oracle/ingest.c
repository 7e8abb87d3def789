/*
 * CPU restatement of the reference's dataset ingestion: the two image
 * samplers, the CRC-32 that seeds them, MT19937 and the two Boost.Random
 * distributions they draw through, and the RGB -> label mapping.
 *
 * TEST INFRASTRUCTURE ONLY (linked into liboracle.so): used by tests/ as the
 * checker and by the CPU-baseline leg of the ingestion benchmark.
 *
 * Follows, in the reference (paths relative to its root):
 *   src/padenti/uniform_image_sampler_impl.hpp:34-63      or_sample_uniform
 *   src/padenti/classweighted_image_sampler_impl.hpp:26-44  or_bucket_sizes
 *   src/padenti/classweighted_image_sampler_impl.hpp:57-123 or_sample_weighted
 *   src/padenti/cv_image_loader_impl.hpp:235-267          or_rgb_to_labels
 * Third-party arithmetic the reference calls (absent here, restated from the
 * published algorithms):
 *   boost::crc_32_type            = CRC-32/ISO-HDLC (reflected 0x04C11DB7,
 *                                   init and xorout 0xFFFFFFFF)
 *   boost::random::mt19937        = MT19937 (Matsumoto & Nishimura 1998),
 *                                   seed(v): init_genrand(v)
 *   uniform_int_distribution<>    Boost's generate_uniform_int for an
 *                                   engine range 2^32 - 1 above the wanted
 *                                   range: bucket = brange / (range + 1),
 *                                   plus one when brange % (range + 1) ==
 *                                   range; draw x / bucket until <= range;
 *                                   a zero range draws nothing
 *   uniform_01<double>            x * (1 / (brange + 1.0)), redrawn while >= 1
 * Pinned by tests/test_oracle_ingest.py: zlib.crc32 and the CRC-32 check
 * value; numpy's MT19937 under legacy seeding and the C++ standard's
 * 10000th-output requirement for the default seed; the closed form
 * floor(2^32 / n) of the bucket; hand-worked sampler cases on given streams.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* ---- CRC-32, one byte at a time ---------------------------------------- */
static uint32_t crc_table[256];
static int crc_ready = 0;

static void crc_init(void)
{
  for (uint32_t i = 0; i < 256; i++) {
    uint32_t c = i;
    for (int k = 0; k < 8; k++) c = (c & 1u) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
    crc_table[i] = c;
  }
  crc_ready = 1;
}

uint32_t or_crc32(const uint8_t *p, uint64_t n)
{
  if (!crc_ready) crc_init();
  uint32_t c = 0xFFFFFFFFu;
  for (uint64_t i = 0; i < n; i++) c = crc_table[(c ^ p[i]) & 0xFFu] ^ (c >> 8);
  return c ^ 0xFFFFFFFFu;
}

/* ---- generators: MT19937, or a given stream (tests) -------------------- */
typedef struct {
  uint32_t mt[624];
  uint32_t i;
  const uint32_t *stream; /* non-NULL: replay these outputs instead */
  uint64_t n_stream, used;
  int overrun;
} or_gen;

static void mt_seed(or_gen *g, uint32_t s)
{
  g->mt[0] = s;
  for (uint32_t i = 1; i < 624; i++) g->mt[i] = 1812433253u * (g->mt[i - 1] ^ (g->mt[i - 1] >> 30)) + i;
  g->i = 624;
  g->stream = NULL;
  g->used = 0;
  g->overrun = 0;
}

static uint32_t gen_next(or_gen *g)
{
  if (g->stream) {
    if (g->used >= g->n_stream) {
      g->overrun = 1;
      return 0;
    }
    return g->stream[g->used++];
  }
  if (g->i >= 624) {
    for (uint32_t k = 0; k < 624; k++) {
      const uint32_t y = (g->mt[k] & 0x80000000u) | (g->mt[(k + 1) % 624] & 0x7FFFFFFFu);
      g->mt[k] = g->mt[(k + 397) % 624] ^ (y >> 1) ^ ((y & 1u) ? 0x9908B0DFu : 0u);
    }
    g->i = 0;
  }
  uint32_t y = g->mt[g->i++];
  y ^= y >> 11;
  y ^= (y << 7) & 0x9D2C5680u;
  y ^= (y << 15) & 0xEFC60000u;
  y ^= y >> 18;
  g->used++;
  return y;
}

void or_mt_outputs(uint32_t seed, uint64_t n, uint32_t *out)
{
  or_gen g;
  mt_seed(&g, seed);
  for (uint64_t i = 0; i < n; i++) out[i] = gen_next(&g);
}

/* Boost generate_uniform_int for an engine of range [0, 2^32 - 1] */
static uint32_t uniform_int(or_gen *g, uint32_t lo, uint32_t hi)
{
  const uint32_t range = hi - lo, brange = 0xFFFFFFFFu;
  if (range == 0) return lo;
  uint32_t bucket = brange / (range + 1u);
  if (brange % (range + 1u) == range) ++bucket;
  for (;;) {
    const uint32_t r = gen_next(g) / bucket;
    if (r <= range) return r + lo;
    if (g->overrun) return lo;
  }
}

static double uniform_01(or_gen *g)
{
  const double factor = 1.0 / ((double)0xFFFFFFFFu + 1.0);
  for (;;) {
    const double r = (double)gen_next(g) * factor;
    if (r < 1.0) return r;
  }
}

static int cmp_u32(const void *a, const void *b)
{
  const uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return (x > y) - (x < y);
}

/* ---- UniformImageSampler::sample ---------------------------------------- */
static uint32_t sample_uniform(or_gen *g, const uint8_t *labels, uint32_t w, uint32_t h, uint32_t n_samples,
                               uint32_t *out)
{
  const uint64_t npx = (uint64_t)w * h;
  uint32_t *nn = (uint32_t *)malloc((npx ? npx : 1) * sizeof(uint32_t));
  uint32_t m = 0;
  for (uint64_t id = 0; id < npx; id++)
    if (labels[id]) nn[m++] = (uint32_t)id;
  if (m == 0) { /* the reference's range would be empty: no samples */
    free(nn);
    return 0;
  }
  for (uint32_t i = 0; i < n_samples; i++) out[i] = nn[uniform_int(g, 0, m - 1)];
  qsort(out, n_samples, sizeof(uint32_t), cmp_u32);
  free(nn);
  return n_samples;
}

uint32_t or_sample_uniform(const uint8_t *data, uint64_t data_bytes, const uint8_t *labels, uint32_t w, uint32_t h,
                           uint32_t n_samples, uint32_t seed, uint32_t *out)
{
  or_gen g;
  mt_seed(&g, or_crc32(data, data_bytes) ^ seed);
  return sample_uniform(&g, labels, w, h, n_samples, out);
}

uint32_t or_sample_uniform_stream(const uint8_t *labels, uint32_t w, uint32_t h, uint32_t n_samples,
                                  const uint32_t *stream, uint64_t n_stream, uint32_t *out, uint64_t *used)
{
  or_gen g;
  mt_seed(&g, 0);
  g.stream = stream;
  g.n_stream = n_stream;
  const uint32_t r = sample_uniform(&g, labels, w, h, n_samples, out);
  *used = g.overrun ? (uint64_t)-1 : g.used;
  return r;
}

/* ---- ClassWeightedImageSampler ------------------------------------------ */
void or_bucket_sizes(uint32_t n_samples, uint32_t n_classes, const float *weights, uint32_t *B)
{
  const float p = 1.f / (float)n_classes;
  for (uint32_t c = 0; c < n_classes; c++) {
    const float w = weights ? weights[c] : p;
    B[c] = (uint32_t)ceilf(w * (float)n_samples);
  }
}

static uint32_t sample_weighted(or_gen *g, const uint8_t *labels, uint32_t w, uint32_t h, uint32_t n_samples,
                                uint32_t n_classes, const float *weights, uint32_t *out)
{
  uint32_t *B = (uint32_t *)calloc(n_classes, sizeof(uint32_t));
  or_bucket_sizes(n_samples, n_classes, weights, B);
  uint32_t **bucket = (uint32_t **)calloc(n_classes, sizeof(uint32_t *));
  uint32_t *n_in = (uint32_t *)calloc(n_classes, sizeof(uint32_t));
  uint32_t *counter = (uint32_t *)calloc(n_classes, sizeof(uint32_t));
  for (uint32_t c = 0; c < n_classes; c++) bucket[c] = (uint32_t *)malloc((B[c] ? B[c] : 1) * sizeof(uint32_t));
  const uint64_t npx = (uint64_t)w * h;
  for (uint64_t i = 0; i < npx; i++) {
    uint32_t l = labels[i];
    if (!l) continue;
    l -= 1;
    if (l >= n_classes) continue; /* out of range: the device path reports an error */
    if (n_in[l] < B[l]) {
      bucket[l][n_in[l]++] = (uint32_t)i;
    } else {
      const uint32_t j = (uint32_t)roundf((float)(uniform_01(g) * (double)counter[l]));
      if (j < B[l]) bucket[l][j] = (uint32_t)i;
    }
    counter[l]++;
  }
  uint32_t tot = 0;
  for (uint32_t c = 0; c < n_classes; c++) {
    memcpy(out + tot, bucket[c], n_in[c] * sizeof(uint32_t));
    tot += n_in[c];
  }
  qsort(out, tot, sizeof(uint32_t), cmp_u32);
  for (uint32_t c = 0; c < n_classes; c++) free(bucket[c]);
  free(bucket);
  free(n_in);
  free(counter);
  free(B);
  return tot;
}

uint32_t or_sample_weighted(const uint8_t *data, uint64_t data_bytes, const uint8_t *labels, uint32_t w, uint32_t h,
                            uint32_t n_samples, uint32_t n_classes, const float *weights, uint32_t seed,
                            uint32_t *out)
{
  or_gen g;
  mt_seed(&g, or_crc32(data, data_bytes) ^ seed);
  return sample_weighted(&g, labels, w, h, n_samples, n_classes, weights, out);
}

uint32_t or_sample_weighted_stream(const uint8_t *labels, uint32_t w, uint32_t h, uint32_t n_samples,
                                   uint32_t n_classes, const float *weights, const uint32_t *stream,
                                   uint64_t n_stream, uint32_t *out, uint64_t *used)
{
  or_gen g;
  mt_seed(&g, 0);
  g.stream = stream;
  g.n_stream = n_stream;
  const uint32_t r = sample_weighted(&g, labels, w, h, n_samples, n_classes, weights, out);
  *used = g.overrun ? (uint64_t)-1 : g.used;
  return r;
}

/* Batch form for the CPU baseline: images in parallel (OpenMP), each one
 * exactly as above; out[i * cap ...]. kind 0 uniform, 1 class-weighted. */
void or_sample_images(int kind, uint32_t n_images, const uint8_t *data, const uint64_t *data_off,
                      const uint64_t *data_bytes, const uint8_t *labels, const uint64_t *lab_off,
                      const uint32_t *widths, const uint32_t *heights, uint32_t n_samples, uint32_t n_classes,
                      const float *weights, uint32_t seed, uint32_t cap, uint32_t *out, uint32_t *n_out,
                      uint32_t *crc_out)
{
  if (!crc_ready) crc_init();
#pragma omp parallel for schedule(dynamic, 1)
  for (uint32_t i = 0; i < n_images; i++) {
    const uint32_t crc = or_crc32(data + data_off[i], data_bytes[i]);
    if (crc_out) crc_out[i] = crc;
    or_gen g;
    mt_seed(&g, crc ^ seed);
    n_out[i] = kind == 0 ? sample_uniform(&g, labels + lab_off[i], widths[i], heights[i], n_samples,
                                          out + (uint64_t)i * cap)
                         : sample_weighted(&g, labels + lab_off[i], widths[i], heights[i], n_samples, n_classes,
                                           weights, out + (uint64_t)i * cap);
  }
}

/* ---- _rgb2label ----------------------------------------------------------- */
void or_rgb_to_labels(const uint8_t *rgb, uint64_t n_px, uint32_t ch, const uint8_t *map, uint32_t n_classes,
                      uint8_t *out)
{
  for (uint64_t i = 0; i < n_px; i++) {
    const uint8_t *p = rgb + i * ch;
    uint32_t l = 0;
    for (; l < n_classes; l++)
      if (p[0] == map[3 * l] && p[1] == map[3 * l + 1] && p[2] == map[3 * l + 2]) break;
    out[i] = (uint8_t)(l < n_classes ? l + 1 : 0);
  }
}
