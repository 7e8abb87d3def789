// C ABI of the dataset-ingestion path (include/padenti_hip.h): device
// samplers, device CRC-32 and RGB -> label mapping.  The kernels are in
// include/padenti/pd_ingest.hpp.
#include "padenti/pd_runtime.hpp"
#include "padenti/pd_ingest.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

using namespace pd;
using namespace pd::engine;
using namespace pd::ingest;

namespace {

// One context per device: a stream and grow-only buffers.  Calls on one
// device are serialised by the mutex.
// Host arrays that feed asynchronous copies live here until the call's
// stream synchronisation.
struct IngestDev {
  std::recursive_mutex mu;
  hipStream_t stream = nullptr;
  DevBuf jobs, blk0, P, gap, acc, crc, sjobs, tile0, bits, cnt, pre, scan_tmp, unsorted, bucket, class_count, B,
      boff, seg_begin, seg_end, sort_tmp, bad, data, labels, samples, n_out, crc_out, keys, cls, rgb, lab_out;
  std::vector<uint32_t> hP;  // x^(8 * 4096 j) mod P
  std::vector<uint32_t> hgap;  // [4][256]: byte k of a register times x^(8 * CRC_GAP)
  std::vector<CrcJob> h_jobs;
  std::vector<SampleJob> h_sjobs;
  std::vector<uint32_t> h_blk0, h_tile0, h_seg_begin, h_B, h_boff, h_keys;
  std::vector<uint8_t> h_cls, h_stage;
};

constexpr CrcTables HT = make_crc_tables();
using Lock = std::lock_guard<std::recursive_mutex>;

IngestDev &ctx(int device)
{
  static std::mutex mu;
  static std::map<int, std::unique_ptr<IngestDev>> all;
  std::lock_guard<std::mutex> g(mu);
  std::unique_ptr<IngestDev> &p = all[device];
  if (!p) {
    p.reset(new IngestDev);
    HIPCHK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  }
  return *p;
}

template <class T>
void upload(DevBuf &b, const std::vector<T> &v, hipStream_t s)
{
  b.ensure(std::max<size_t>(v.size(), 1) * sizeof(T));
  if (!v.empty()) HIPCHK(hipMemcpyAsync(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
}

uint32_t grid_for(uint64_t n, uint32_t per_block, uint32_t cap = 65535u * 16u)
{
  return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + per_block - 1) / per_block, cap));
}

// Bucket sizes B_c = ceilf(w_c * (float)n) (classweighted_image_sampler_impl.hpp:31-33, 40-41).
std::vector<uint32_t> bucket_sizes(const pd_sampler &s)
{
  std::vector<uint32_t> B(s.n_classes);
  const float n = (float)s.n_samples;
  const float p = 1.f / (float)s.n_classes;
  for (uint32_t c = 0; c < s.n_classes; c++) {
    const float w = s.weights ? s.weights[c] : p;
    require(std::isfinite(w) && w >= 0.f, "class weights must be finite and >= 0");
    const float b = std::ceil(w * n);
    require(b < 4294967296.f, "bucket size overflows");
    B[c] = (uint32_t)b;
  }
  return B;
}

uint32_t capacity_of(const pd_sampler &s)
{
  require(s.kind == PD_SAMPLER_UNIFORM || s.kind == PD_SAMPLER_CLASS_WEIGHTED, "unknown sampler kind");
  if (s.kind == PD_SAMPLER_UNIFORM) return s.n_samples;
  require(s.n_classes >= 1 && s.n_classes <= 255, "n_classes must be in [1, 255]");
  uint64_t t = 0;
  for (uint32_t b : bucket_sizes(s)) t += b;
  require(t < (1ull << 31), "sampler capacity too large");
  return (uint32_t)t;
}

// every [off, off + bytes) inside a buffer of `size` bytes
void check_ranges(uint32_t n, const uint64_t *off, const uint64_t *bytes, uint64_t size, const char *what)
{
  for (uint32_t i = 0; i < n; i++)
    require(off[i] <= size && bytes[i] <= size - off[i],
            std::string(what) + " range " + std::to_string(i) + " lies outside its buffer");
}

// CRC jobs of n byte ranges of `base` (checked by the caller); launches
// zero-fill, the block kernel and the finalisation on the context stream,
// crc -> d_crc.
void launch_crc(IngestDev &X, uint32_t n, const uint8_t *base, const uint64_t *off, const uint64_t *bytes,
                uint32_t *d_crc)
{
  std::vector<CrcJob> &jobs = X.h_jobs;
  std::vector<uint32_t> &chunk0 = X.h_blk0;
  jobs.assign(n, CrcJob{});
  chunk0.assign(n, 0u);
  uint64_t total = 0;  // chunks
  uint32_t max_nb = 0;
  for (uint32_t i = 0; i < n; i++) {
    CrcJob &J = jobs[i];
    J.start = (uint64_t)(uintptr_t)base + off[i];
    J.end = J.start + bytes[i];
    J.a0 = J.start & ~63ull;
    const uint64_t nb = bytes[i] ? (J.end - J.a0 + 4095u) / 4096u : 0u;
    require(nb < (1ull << 32) && total + nb < (1ull << 32), "too many bytes in one CRC batch");
    J.first_chunk = chunk0[i] = (uint32_t)total;
    J.n_blocks = (uint32_t)nb;
    total += (nb + CRC_CHUNK - 1) / CRC_CHUNK;
    max_nb = std::max<uint32_t>(max_nb, (uint32_t)nb);
  }
  if (X.hP.empty()) X.hP.push_back(CRC_ONE);
  while (X.hP.size() < max_nb) X.hP.push_back(gf2_mul(X.hP.back(), HT.x2k[15]));  // * x^(8 * 4096)
  if (X.hgap.empty()) {
    const uint32_t C = crc_x8n(HT, CRC_GAP);
    X.hgap.resize(1024);
    for (uint32_t k = 0; k < 4; k++)
      for (uint32_t v = 0; v < 256; v++) X.hgap[k * 256 + v] = gf2_mul(v << (8 * k), C);
  }
  upload(X.jobs, jobs, X.stream);
  upload(X.blk0, chunk0, X.stream);
  upload(X.P, X.hP, X.stream);
  upload(X.gap, X.hgap, X.stream);
  X.acc.ensure(2ull * std::max<uint32_t>(n, 1) * 4u);
  uint32_t *acc = X.acc.as<uint32_t>();
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(2ull * n, 256)), dim3(256), 0, X.stream, acc, 2ull * n, 0u);
  if (total)
    hipLaunchKernelGGL(k_crc32, dim3(grid_for(total, 4, 1u << 20)), dim3(256), 0, X.stream, X.jobs.as<CrcJob>(),
                       X.blk0.as<uint32_t>(), n, (uint32_t)total, X.P.as<uint32_t>(), X.gap.as<uint32_t>(), acc,
                       acc + n);
  hipLaunchKernelGGL(k_crc_final, dim3(grid_for(n, 64)), dim3(64), 0, X.stream, X.jobs.as<CrcJob>(), n, acc,
                     acc + n, d_crc);
  HIPCHK(hipGetLastError());
}

struct Timer {
  hipStream_t s;
  std::vector<Event> ev;
  explicit Timer(hipStream_t st, int n) : s(st), ev(n) {}
  void mark(int i) { HIPCHK(hipEventRecord(ev[i].e, s)); }
  float ms(int a, int b) const { return (float)elapsed_ms(ev[a], ev[b]); }
};

void sample_dev(int device, const pd_sampler &s, uint32_t n, const void *d_data, uint64_t data_size,
                const uint64_t *data_off, const uint64_t *data_bytes, const uint8_t *d_labels, uint64_t labels_size,
                const uint64_t *lab_off, const uint32_t *widths, const uint32_t *heights, uint32_t *d_samples,
                uint32_t *d_n_out, uint32_t *d_crc, pd_ingest_stats *stats)
{
  require(n >= 1, "n_images must be >= 1");
  require(d_data && data_off && data_bytes && d_labels && lab_off && widths && heights && d_samples && d_n_out,
          "null pointer in sample_images");
  check_ranges(n, data_off, data_bytes, data_size, "image data");
  {
    std::vector<uint64_t> lb(n);
    for (uint32_t i = 0; i < n; i++) lb[i] = (uint64_t)widths[i] * heights[i];
    check_ranges(n, lab_off, lb.data(), labels_size, "label image");
  }
  const uint32_t cap = capacity_of(s);
  require((uint64_t)n * cap < (1ull << 31), "n_images * capacity must be < 2^31");
  set_device(device);
  IngestDev &X = ctx(device);
  Lock g(X.mu);
  const hipStream_t st = X.stream;

  std::vector<SampleJob> &jobs = X.h_sjobs;
  std::vector<uint32_t> &tile0 = X.h_tile0, &seg_begin = X.h_seg_begin;
  jobs.assign(n, SampleJob{});
  tile0.assign(n, 0u);
  seg_begin.assign(n, 0u);
  uint64_t n_words = 0, n_tiles = 0;
  uint64_t label_bytes = 0, crc_bytes = 0, max_px = 0;
  for (uint32_t i = 0; i < n; i++) {
    const uint64_t px = (uint64_t)widths[i] * heights[i];
    require(px < (1ull << 31), "an image must have fewer than 2^31 pixels");
    require(((uintptr_t)d_labels + lab_off[i]) % 16 == 0, "label images must start at 16-byte aligned addresses");
    SampleJob &J = jobs[i];
    J.lab_off = lab_off[i];
    J.n_px = (uint32_t)px;
    J.out_base = i * cap;
    J.pad = 0;
    seg_begin[i] = J.out_base;
    label_bytes += px;
    crc_bytes += data_bytes[i];
    max_px = std::max(max_px, px);
    const uint64_t nw = (px + 31) / 32;
    J.word0 = (uint32_t)n_words;
    tile0[i] = (uint32_t)n_tiles;
    n_words += nw;
    n_tiles += (nw + 255) / 256;
  }
  require(n_words < (1ull << 31) && n_tiles < (1ull << 31), "too many pixels in one batch");

  Timer tm(st, 5);
  tm.mark(0);
  X.crc.ensure((size_t)n * 4u);
  uint32_t *crc = d_crc ? d_crc : X.crc.as<uint32_t>();
  launch_crc(X, n, (const uint8_t *)d_data, data_off, data_bytes, crc);
  tm.mark(1);
  upload(X.sjobs, jobs, st);
  upload(X.seg_begin, seg_begin, st);
  X.seg_end.ensure((size_t)n * 4u);
  X.unsorted.ensure((size_t)n * cap * 4u + 4u);
  uint32_t *uns = X.unsorted.as<uint32_t>();
  if (s.kind == PD_SAMPLER_UNIFORM) {
    const uint32_t nw = (uint32_t)n_words;
    upload(X.tile0, tile0, st);
    X.bits.ensure((size_t)nw * 4u + 4u);
    X.cnt.ensure(((size_t)nw + 1) * 4u);
    X.pre.ensure(((size_t)nw + 1) * 4u);
    uint32_t *cnt = X.cnt.as<uint32_t>();
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, cnt + nw, 1ull, 0u);
    if (n_tiles)
      hipLaunchKernelGGL(k_label_bits, dim3((uint32_t)n_tiles), dim3(256), 0, st, d_labels, X.sjobs.as<SampleJob>(),
                         X.tile0.as<uint32_t>(), n, X.bits.as<uint32_t>(), cnt);
    size_t tb = 0;
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt, X.pre.as<uint32_t>(), (int)nw + 1, st));
    X.scan_tmp.ensure(tb);
    HIPCHK(hipcub::DeviceScan::ExclusiveSum(X.scan_tmp.p, tb, cnt, X.pre.as<uint32_t>(), (int)nw + 1, st));
    tm.mark(2);
    hipLaunchKernelGGL(k_sample_uniform, dim3(n), dim3(256), 0, st, X.sjobs.as<SampleJob>(), X.bits.as<uint32_t>(),
                       X.pre.as<uint32_t>(), crc, s.seed, s.n_samples, uns, d_n_out, X.seg_end.as<uint32_t>());
    HIPCHK(hipGetLastError());
  } else {
    std::vector<uint32_t> &B = X.h_B, &boff = X.h_boff;
    B = bucket_sizes(s);
    boff.assign(s.n_classes, 0u);
    uint32_t o = 0;
    for (uint32_t c = 0; c < s.n_classes; c++) { boff[c] = o; o += B[c]; }
    upload(X.B, B, st);
    upload(X.boff, boff, st);
    X.bucket.ensure((size_t)n * cap * 4u + 4u);
    X.class_count.ensure((size_t)n * s.n_classes * 4u);
    X.bad.ensure(4);
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for((uint64_t)n * cap, 256)), dim3(256), 0, st,
                       X.bucket.as<uint32_t>(), (uint64_t)n * cap, 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, st, X.bad.as<uint32_t>(), 1ull, 0u);
    tm.mark(2);
    hipLaunchKernelGGL(k_sample_weighted, dim3((n + 3) / 4), dim3(256), 0, st, d_labels, X.sjobs.as<SampleJob>(), n,
                       crc, s.seed, s.n_classes, X.B.as<uint32_t>(), X.boff.as<uint32_t>(), X.bucket.as<uint32_t>(),
                       X.class_count.as<uint32_t>(), X.bad.as<uint32_t>());
    hipLaunchKernelGGL(k_compact_buckets, dim3(n), dim3(256), 0, st, X.sjobs.as<SampleJob>(), s.n_classes,
                       X.B.as<uint32_t>(), X.boff.as<uint32_t>(), X.class_count.as<uint32_t>(),
                       X.bucket.as<uint32_t>(), uns, d_n_out, X.seg_end.as<uint32_t>());
    HIPCHK(hipGetLastError());
  }
  tm.mark(3);
  int end_bit = 1;
  while (end_bit < 32 && (1ull << end_bit) < max_px) end_bit++;
  size_t sb = 0;
  HIPCHK(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, sb, uns, d_samples, (int)((uint64_t)n * cap), (int)n,
                                                     X.seg_begin.as<uint32_t>(), X.seg_end.as<uint32_t>(), 0,
                                                     end_bit, st));
  X.sort_tmp.ensure(sb);
  HIPCHK(hipcub::DeviceSegmentedRadixSort::SortKeys(X.sort_tmp.p, sb, uns, d_samples, (int)((uint64_t)n * cap),
                                                     (int)n, X.seg_begin.as<uint32_t>(), X.seg_end.as<uint32_t>(),
                                                     0, end_bit, st));
  tm.mark(4);
  HIPCHK(hipStreamSynchronize(st));
  if (s.kind == PD_SAMPLER_CLASS_WEIGHTED) {
    uint32_t bad = 0;
    HIPCHK(hipMemcpy(&bad, X.bad.p, 4, hipMemcpyDeviceToHost));
    require(bad == 0, "a label exceeds n_classes");
  }
  if (stats) {
    stats->crc_ms = tm.ms(0, 1);
    stats->index_ms = s.kind == PD_SAMPLER_UNIFORM ? tm.ms(1, 2) : 0.f;
    stats->sample_ms = tm.ms(2, 3);
    stats->sort_ms = tm.ms(3, 4);
    stats->total_ms = tm.ms(0, 4);
    stats->crc_bytes = crc_bytes;
    stats->label_bytes = label_bytes;
  }
}

// host array -> device buffer with every piece at a 16-byte aligned offset
const uint8_t *stage(DevBuf &b, std::vector<uint8_t> &h, const uint8_t *src, uint32_t n, const uint64_t *off,
                     const uint64_t *bytes, std::vector<uint64_t> &new_off, hipStream_t st)
{
  new_off.resize(n);
  uint64_t t = 0;
  for (uint32_t i = 0; i < n; i++) {
    new_off[i] = t;
    t += (bytes[i] + 15) & ~15ull;
  }
  h.assign((size_t)std::max<uint64_t>(t, 16), 0);
  for (uint32_t i = 0; i < n; i++) std::memcpy(h.data() + new_off[i], src + off[i], (size_t)bytes[i]);
  b.ensure(h.size());
  HIPCHK(hipMemcpyAsync(b.p, h.data(), h.size(), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return b.as<uint8_t>();
}

void rgb_keys(const uint8_t *map, uint32_t nC, std::vector<uint32_t> &keys, std::vector<uint8_t> &cls)
{
  require(map != nullptr && nC >= 1 && nC <= 255, "rgb2label needs 1..255 classes");
  std::map<uint32_t, uint8_t> first;
  for (uint32_t c = 0; c < nC; c++) {
    const uint32_t k = ((uint32_t)map[3 * c] << 16) | ((uint32_t)map[3 * c + 1] << 8) | map[3 * c + 2];
    first.emplace(k, (uint8_t)c);  // keeps the first class of a repeated colour
  }
  keys.clear();
  cls.clear();
  for (auto &kv : first) {
    keys.push_back(kv.first);
    cls.push_back(kv.second);
  }
}

}  // namespace

extern "C" {

int pd_sampler_capacity(const pd_sampler *s, uint32_t *capacity)
{
  return guarded([&] {
    require(s && capacity, "null argument");
    *capacity = capacity_of(*s);
  });
}

int pd_sample_images_dev(int device, const pd_sampler *s, uint32_t n_images, const void *d_data, uint64_t data_size,
                         const uint64_t *data_offsets, const uint64_t *data_bytes, const uint8_t *d_labels,
                         uint64_t labels_size, const uint64_t *label_offsets, const uint32_t *widths,
                         const uint32_t *heights, uint32_t *d_samples, uint32_t *d_n_out, uint32_t *d_crc,
                         pd_ingest_stats *stats)
{
  return guarded([&] {
    require(s != nullptr, "null sampler");
    sample_dev(device, *s, n_images, d_data, data_size, data_offsets, data_bytes, d_labels, labels_size,
               label_offsets, widths, heights, d_samples, d_n_out, d_crc, stats);
  });
}

int pd_sample_images(int device, const pd_sampler *s, uint32_t n_images, const void *data,
                     const uint64_t *data_offsets, const uint64_t *data_bytes, const uint8_t *labels,
                     const uint64_t *label_offsets, const uint32_t *widths, const uint32_t *heights,
                     uint32_t *samples, uint32_t *n_out, uint32_t *crc)
{
  return guarded([&] {
    require(s && data && data_offsets && data_bytes && labels && label_offsets && widths && heights && samples &&
                n_out,
            "null pointer in pd_sample_images");
    require(n_images >= 1, "n_images must be >= 1");
    const uint32_t cap = capacity_of(*s);
    set_device(device);
    IngestDev &X = ctx(device);
    std::vector<uint64_t> doff, loff, lbytes(n_images);
    for (uint32_t i = 0; i < n_images; i++) lbytes[i] = (uint64_t)widths[i] * heights[i];
    Lock g(X.mu);
    const uint8_t *dd = stage(X.data, X.h_stage, (const uint8_t *)data, n_images, data_offsets, data_bytes, doff,
                              X.stream);
    const uint8_t *dl = stage(X.labels, X.h_stage, labels, n_images, label_offsets, lbytes.data(), loff, X.stream);
    X.samples.ensure((size_t)n_images * cap * 4u + 4u);
    X.n_out.ensure((size_t)n_images * 4u);
    X.crc_out.ensure((size_t)n_images * 4u);
    uint32_t *ds = X.samples.as<uint32_t>(), *dn = X.n_out.as<uint32_t>(), *dc = X.crc_out.as<uint32_t>();
    sample_dev(device, *s, n_images, dd, X.data.cap, doff.data(), data_bytes, dl, X.labels.cap, loff.data(), widths,
               heights, ds, dn, dc, nullptr);
    HIPCHK(hipMemcpy(samples, ds, (size_t)n_images * cap * 4u, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(n_out, dn, (size_t)n_images * 4u, hipMemcpyDeviceToHost));
    if (crc) HIPCHK(hipMemcpy(crc, dc, (size_t)n_images * 4u, hipMemcpyDeviceToHost));
  });
}

int pd_crc32_images_dev(int device, uint32_t n, const void *d_data, uint64_t data_size, const uint64_t *offsets,
                        const uint64_t *bytes, uint32_t *d_crc, float *kernel_ms)
{
  return guarded([&] {
    require(n >= 1 && d_data && offsets && bytes && d_crc, "bad CRC arguments");
    check_ranges(n, offsets, bytes, data_size, "CRC");
    set_device(device);
    IngestDev &X = ctx(device);
    Lock g(X.mu);
    Timer tm(X.stream, 2);
    tm.mark(0);
    launch_crc(X, n, (const uint8_t *)d_data, offsets, bytes, d_crc);
    tm.mark(1);
    HIPCHK(hipStreamSynchronize(X.stream));
    if (kernel_ms) *kernel_ms = tm.ms(0, 1);
  });
}

int pd_rgb_to_labels_dev(int device, uint64_t n_pixels, uint32_t channels, const uint8_t *d_rgb,
                         const uint8_t *rgb2label, uint32_t n_classes, uint8_t *d_labels, float *kernel_ms)
{
  return guarded([&] {
    require(channels == 3 || channels == 4, "channels must be 3 or 4");
    require(d_rgb && d_labels, "null pointer");
    set_device(device);
    IngestDev &X = ctx(device);
    Lock g(X.mu);
    std::vector<uint32_t> &keys = X.h_keys;
    std::vector<uint8_t> &cls = X.h_cls;
    rgb_keys(rgb2label, n_classes, keys, cls);
    upload(X.keys, keys, X.stream);
    upload(X.cls, cls, X.stream);
    Timer tm(X.stream, 2);
    tm.mark(0);
    if (n_pixels)
      hipLaunchKernelGGL(k_rgb_labels, dim3(grid_for(n_pixels, 256, 1u << 16)), dim3(256), 0, X.stream, d_rgb,
                         n_pixels, channels, X.keys.as<uint32_t>(), X.cls.as<uint8_t>(), (uint32_t)keys.size(),
                         d_labels);
    HIPCHK(hipGetLastError());
    tm.mark(1);
    HIPCHK(hipStreamSynchronize(X.stream));
    if (kernel_ms) *kernel_ms = tm.ms(0, 1);
  });
}

int pd_rgb_to_labels(int device, uint64_t n_pixels, uint32_t channels, const uint8_t *rgb, const uint8_t *rgb2label,
                     uint32_t n_classes, uint8_t *labels)
{
  return guarded([&] {
    require(channels == 3 || channels == 4, "channels must be 3 or 4");
    require(rgb && labels, "null pointer");
    set_device(device);
    IngestDev &X = ctx(device);
    Lock g(X.mu);
    X.rgb.ensure(n_pixels * channels + 16);
    X.lab_out.ensure(n_pixels + 16);
    if (n_pixels) HIPCHK(hipMemcpy(X.rgb.p, rgb, n_pixels * channels, hipMemcpyHostToDevice));
    const int rc = pd_rgb_to_labels_dev(device, n_pixels, channels, X.rgb.as<uint8_t>(), rgb2label, n_classes,
                                        X.lab_out.as<uint8_t>(), nullptr);
    if (rc != PD_OK) throw ArgError(g_err);
    if (n_pixels) HIPCHK(hipMemcpy(labels, X.lab_out.p, n_pixels, hipMemcpyDeviceToHost));
  });
}

}  // extern "C"
