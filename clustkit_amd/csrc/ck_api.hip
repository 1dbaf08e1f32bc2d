// extern "C" entry points of libclustkit_hip.so (declared in
// include/clustkit_hip.h).  Host-side orchestration only: argument checks,
// per-device workspace, kernel launches on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/clustkit_hip.h"
#include "ck_internal.h"

namespace ck {
__global__ void sw_plain_kernel(const int32_t*, const int32_t*, const int32_t*, int64_t,
                                int64_t, int64_t, const uint8_t*, const int64_t*,
                                const int32_t*, const int32_t*, int32_t, const int8_t*, float,
                                uint32_t, int4*, int32_t, float*, int32_t*, int32_t*,
                                uint8_t*);
__global__ void sw_wave_kernel(const int32_t*, Counters*, const int32_t*, const uint8_t*,
                               const int64_t*, const int32_t*, const int32_t*, int32_t,
                               const int8_t*, float, uint32_t, int32_t, int32_t, float*,
                               int32_t*, int32_t*, uint8_t*, int32_t*);
__global__ void ck_reset_kernel(Counters*, int32_t*, int32_t);
__global__ void ck_prep_kernel(const int32_t*, int64_t, const int32_t*, const int32_t*,
                               int32_t, float, uint32_t, int32_t, int32_t, int32_t*,
                               int32_t*, Counters*, int32_t*, float*, int32_t*, int32_t*,
                               uint8_t*);
__global__ void ck_scan_kernel(int32_t*, Counters*);
__global__ void ck_scatter_kernel(const int32_t*, int64_t, int32_t*, int32_t*);
__global__ void ck_jaccard_kernel(const int32_t*, int64_t, const uint64_t*, int32_t, float*);
}  // namespace ck

using namespace ck;

namespace {

thread_local std::string g_err = "no error";
thread_local int64_t g_stats[3] = {0, 0, 0};
thread_local Counters* g_stats_ctr = nullptr;
thread_local hipEvent_t g_ev[4] = {nullptr, nullptr, nullptr, nullptr};
thread_local bool g_ev_valid = false;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define CK_HIP(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(CK_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(e_));          \
  } while (0)

// Grow-only device buffer.
struct DevBuf {
  void* ptr = nullptr;
  size_t bytes = 0;
  int ensure(size_t need) {
    if (need <= bytes) return CK_OK;
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    bytes = 0;
    size_t want = std::max<size_t>(need, 256);
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
      ptr = nullptr;
      return fail(CK_ERR_NOMEM, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    bytes = want;
    return CK_OK;
  }
  template <class T> T* as() const { return static_cast<T*>(ptr); }
};

struct DevCtx {
  bool ready = false;
  int device = -1;
  int num_cu = 0;
  int8_t sub_host[400];
  bool sub_valid = false;
  DevBuf sub, ctr, hist, bin_of, queue, rerun, plain_ws;
  // staging for the host-pointer entry points
  DevBuf h_pairs, h_flat, h_off, h_len, h_hint, o_sims, o_scores, o_matches, o_mask,
      j_sk, j_out;
  int8_t* sub_pinned = nullptr;
  std::mutex mu;
};

std::mutex g_mu;
std::vector<DevCtx*> g_ctx;

int get_ctx(int device, DevCtx** out) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(CK_ERR_NODEVICE, "no GPU visible to HIP");
  if (device < 0 || device >= n) return fail(CK_ERR_INVALID, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_mu);
  if ((int)g_ctx.size() < n) g_ctx.resize(n, nullptr);
  if (!g_ctx[device]) g_ctx[device] = new DevCtx();
  DevCtx* c = g_ctx[device];
  CK_HIP(hipSetDevice(device));
  if (!c->ready) {
    hipDeviceProp_t prop;
    CK_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cu = prop.multiProcessorCount;
    c->device = device;
    CK_HIP(hipHostMalloc((void**)&c->sub_pinned, 400, hipHostMallocDefault));
    int rc;
    if ((rc = c->sub.ensure(400)) || (rc = c->ctr.ensure(sizeof(Counters))) ||
        (rc = c->hist.ensure(NBINS * sizeof(int32_t))))
      return rc;
    c->ready = true;
  }
  *out = c;
  return CK_OK;
}

int current_ctx(DevCtx** out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  return get_ctx(dev, out);
}

int upload_sub(DevCtx* c, const int8_t* h_sub, hipStream_t st) {
  if (c->sub_valid && std::memcmp(c->sub_host, h_sub, 400) == 0) return CK_OK;
  // The pinned staging copy must not change while an earlier upload is in flight.
  CK_HIP(hipStreamSynchronize(st));
  std::memcpy(c->sub_pinned, h_sub, 400);
  std::memcpy(c->sub_host, h_sub, 400);
  CK_HIP(hipMemcpyAsync(c->sub.ptr, c->sub_pinned, 400, hipMemcpyHostToDevice, st));
  c->sub_valid = true;
  return CK_OK;
}

int32_t fast_w_cap(int32_t band, int32_t max_len) {
  int64_t w = band > 0 ? 2 * (int64_t)std::max(band, 50) + 1 : 2 * (int64_t)max_len + 1;
  return (int32_t)std::min<int64_t>(w, MAX_FAST_W);
}

int32_t plain_width(int32_t band, int32_t max_len) {
  return band > 0 ? 2 * std::max(band, 50) + 2 : 2 * std::max(max_len, 1) + 2;
}

constexpr int64_t PLAIN_BATCH = 500000;           // pairs per plain-form launch
constexpr size_t PLAIN_WS_BUDGET = size_t(1) << 30;  // bytes of int32 band rows

}  // namespace

extern "C" {

const char* ck_last_error(void) { return g_err.c_str(); }

const char* ck_version(void) { return "clustkit_hip 0.1 (gfx950)"; }

int ck_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int ck_init(int device) {
  DevCtx* c;
  return get_ctx(device, &c);
}

int ck_sw_align_device(const int32_t* d_pairs, int64_t M, const uint8_t* d_flat,
                       const int64_t* d_offsets, const int32_t* d_lengths, int32_t max_len,
                       int32_t band_width, const int8_t* h_sub_matrix,
                       const int32_t* d_diag_hints, float threshold, uint32_t flags,
                       float* d_out_sims, int32_t* d_out_scores, int32_t* d_out_matches,
                       uint8_t* d_out_mask, void* stream) {
  if (M < 0) return fail(CK_ERR_INVALID, "negative pair count");
  if (M > INT32_MAX) return fail(CK_ERR_INVALID, "pair count above 2^31-1");
  if (M == 0) return CK_OK;
  if (!d_pairs || !d_flat || !d_offsets || !d_lengths || !h_sub_matrix || !d_out_sims ||
      !d_out_scores)
    return fail(CK_ERR_INVALID, "null pointer argument");
  if (max_len < 0) return fail(CK_ERR_INVALID, "negative max_len");
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  if ((rc = upload_sub(c, h_sub_matrix, st))) return rc;
  if ((rc = c->bin_of.ensure(M * sizeof(int32_t))) || (rc = c->queue.ensure(M * sizeof(int32_t))) ||
      (rc = c->rerun.ensure(M * sizeof(int32_t))))
    return rc;
  Counters* ctr = c->ctr.as<Counters>();
  int32_t* hist = c->hist.as<int32_t>();
  const bool plain = (flags & CK_SW_PLAIN) != 0;
  const int32_t w_cap = fast_w_cap(band_width, max_len);
  const int32_t pw = plain_width(band_width, max_len);

  // Plain-kernel workspace: T threads x pw band entries x 16 bytes.
  int64_t t_plain = (int64_t)(PLAIN_WS_BUDGET / ((size_t)pw * sizeof(int4)));
  t_plain = std::min<int64_t>(t_plain, 256 * 1024);
  t_plain = std::max<int64_t>(256, (t_plain / 256) * 256);
  if ((rc = c->plain_ws.ensure((size_t)t_plain * pw * sizeof(int4)))) return rc;

  const bool timing = (flags & CK_SW_TIMING) != 0;
  if (timing) {
    for (auto& e : g_ev)
      if (!e) CK_HIP(hipEventCreate(&e));
    CK_HIP(hipEventRecord(g_ev[0], st));
  }
  g_ev_valid = false;
  const int prep_grid = (int)std::min<int64_t>((M + 255) / 256, 4 * (int64_t)c->num_cu);
  hipLaunchKernelGGL(ck_reset_kernel, dim3((NBINS + 255) / 256), dim3(256), 0, st, ctr, hist,
                     NBINS);
  hipLaunchKernelGGL(ck_prep_kernel, dim3(prep_grid), dim3(256), 0, st, d_pairs, M, d_lengths,
                     d_diag_hints, band_width, threshold, flags, w_cap, plain ? 1 : 0,
                     c->bin_of.as<int32_t>(), hist, ctr, c->rerun.as<int32_t>(), d_out_sims,
                     d_out_scores, d_out_matches, d_out_mask);
  hipLaunchKernelGGL(ck_scan_kernel, dim3(1), dim3(1024), 0, st, hist, ctr);
  hipLaunchKernelGGL(ck_scatter_kernel, dim3(prep_grid), dim3(256), 0, st,
                     c->bin_of.as<int32_t>(), M, hist, c->queue.as<int32_t>());
  const int plain_grid = (int)(t_plain / 256);
  if (timing) CK_HIP(hipEventRecord(g_ev[1], st));
  if (plain) {
    for (int64_t base = 0; base < M; base += PLAIN_BATCH) {
      hipLaunchKernelGGL(sw_plain_kernel, dim3(plain_grid), dim3(256), 0, st, d_pairs,
                         c->queue.as<int32_t>(), &ctr->queue_len, (int64_t)0, base,
                         PLAIN_BATCH, d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                         c->sub.as<int8_t>(), threshold, flags, c->plain_ws.as<int4>(), pw,
                         d_out_sims, d_out_scores, d_out_matches, d_out_mask);
    }
    if (timing) {
      CK_HIP(hipEventRecord(g_ev[2], st));
      CK_HIP(hipEventRecord(g_ev[3], st));
    }
  } else {
    const int32_t rb_stride = w_cap + 2 * SEG;
    const int32_t bs_stride = ((w_cap + 3 * SEG) + 15) & ~15;
    const size_t lds = TABLE_DIM * TABLE_DIM * 4 + (size_t)(WAVE_BLOCK / SEG) *
                           (rb_stride * sizeof(int2) + bs_stride);
    if (lds > 64 * 1024)
      CK_HIP(hipFuncSetAttribute((const void*)sw_wave_kernel,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    int per_cu = 0;
    CK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, sw_wave_kernel, WAVE_BLOCK,
                                                        lds));
    if (per_cu < 1) per_cu = 1;
    const int64_t need = (M + WAVE_BLOCK / SEG - 1) / (WAVE_BLOCK / SEG);
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)per_cu * c->num_cu));
    hipLaunchKernelGGL(sw_wave_kernel, dim3(grid), dim3(WAVE_BLOCK), lds, st,
                       c->queue.as<int32_t>(), ctr, d_pairs, d_flat, d_offsets, d_lengths,
                       d_diag_hints, band_width, c->sub.as<int8_t>(), threshold, flags,
                       rb_stride, bs_stride, d_out_sims, d_out_scores, d_out_matches,
                       d_out_mask, c->rerun.as<int32_t>());
    if (timing) CK_HIP(hipEventRecord(g_ev[2], st));
    hipLaunchKernelGGL(sw_plain_kernel, dim3(plain_grid), dim3(256), 0, st, d_pairs,
                       c->rerun.as<int32_t>(), &ctr->rerun_len, (int64_t)0, (int64_t)0,
                       (int64_t)INT32_MAX, d_flat, d_offsets, d_lengths, d_diag_hints,
                       band_width, c->sub.as<int8_t>(), threshold, flags,
                       c->plain_ws.as<int4>(), pw, d_out_sims, d_out_scores, d_out_matches,
                       d_out_mask);
    if (timing) CK_HIP(hipEventRecord(g_ev[3], st));
  }
  CK_HIP(hipGetLastError());
  g_stats_ctr = ctr;
  g_ev_valid = timing;
  return CK_OK;
}

int ck_sw_last_timing(float* queue_ms, float* main_ms, float* rerun_ms) {
  if (!g_ev_valid) return fail(CK_ERR_INVALID, "no timed ck_sw_align_device call on this thread");
  CK_HIP(hipEventSynchronize(g_ev[3]));
  float a = 0, b = 0, d = 0;
  CK_HIP(hipEventElapsedTime(&a, g_ev[0], g_ev[1]));
  CK_HIP(hipEventElapsedTime(&b, g_ev[1], g_ev[2]));
  CK_HIP(hipEventElapsedTime(&d, g_ev[2], g_ev[3]));
  if (queue_ms) *queue_ms = a;
  if (main_ms) *main_ms = b;
  if (rerun_ms) *rerun_ms = d;
  return CK_OK;
}

int ck_sw_last_stats(int64_t* batched_pairs, int64_t* rerun_pairs, int64_t* skipped_pairs) {
  if (!g_stats_ctr) return fail(CK_ERR_INVALID, "no ck_sw_align_device call on this thread");
  Counters h;
  CK_HIP(hipMemcpy(&h, g_stats_ctr, sizeof h, hipMemcpyDeviceToHost));
  if (batched_pairs) *batched_pairs = h.queue_len;
  if (rerun_pairs) *rerun_pairs = h.rerun_len;
  if (skipped_pairs) *skipped_pairs = h.skipped;
  return CK_OK;
}

int ck_sw_align_host(const int32_t* h_pairs, int64_t M, const uint8_t* h_flat, int64_t flat_len,
                     const int64_t* h_offsets, const int32_t* h_lengths, int64_t n_seqs,
                     int32_t band_width, const int8_t* h_sub_matrix, const int32_t* h_diag_hints,
                     float threshold, uint32_t flags, float* h_out_sims, int32_t* h_out_scores,
                     int32_t* h_out_matches, uint8_t* h_out_mask) {
  if (M < 0 || n_seqs < 0 || flat_len < 0) return fail(CK_ERR_INVALID, "negative size");
  if (M == 0) return CK_OK;
  if (!h_pairs || !h_offsets || !h_lengths || !h_sub_matrix || !h_out_sims || !h_out_scores)
    return fail(CK_ERR_INVALID, "null pointer argument");
  if (n_seqs == 0) return fail(CK_ERR_INVALID, "pairs given but no sequences");
  int32_t max_len = 0;
  for (int64_t i = 0; i < n_seqs; ++i) {
    if (h_lengths[i] < 0) return fail(CK_ERR_INVALID, "negative length at %lld", (long long)i);
    if (h_lengths[i] > 0 && (h_offsets[i] < 0 || h_offsets[i] + h_lengths[i] > flat_len))
      return fail(CK_ERR_INVALID, "sequence %lld outside the flat buffer", (long long)i);
    max_len = std::max(max_len, h_lengths[i]);
  }
  for (int64_t p = 0; p < 2 * M; ++p)
    if (h_pairs[p] < 0 || h_pairs[p] >= n_seqs)
      return fail(CK_ERR_INVALID, "pair index %d out of range", h_pairs[p]);
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    if ((rc = c->h_pairs.ensure(2 * M * sizeof(int32_t))) ||
        (rc = c->h_flat.ensure(std::max<int64_t>(flat_len, 1))) ||
        (rc = c->h_off.ensure(n_seqs * sizeof(int64_t))) ||
        (rc = c->h_len.ensure(n_seqs * sizeof(int32_t))) ||
        (rc = c->o_sims.ensure(M * sizeof(float))) ||
        (rc = c->o_scores.ensure(M * sizeof(int32_t))) ||
        (rc = c->o_matches.ensure(M * sizeof(int32_t))) || (rc = c->o_mask.ensure(M)))
      return rc;
    if (h_diag_hints && (rc = c->h_hint.ensure(M * sizeof(int32_t)))) return rc;
    CK_HIP(hipMemcpy(c->h_pairs.ptr, h_pairs, 2 * M * sizeof(int32_t), hipMemcpyHostToDevice));
    if (flat_len) CK_HIP(hipMemcpy(c->h_flat.ptr, h_flat, flat_len, hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(c->h_off.ptr, h_offsets, n_seqs * sizeof(int64_t), hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(c->h_len.ptr, h_lengths, n_seqs * sizeof(int32_t), hipMemcpyHostToDevice));
    if (h_diag_hints)
      CK_HIP(hipMemcpy(c->h_hint.ptr, h_diag_hints, M * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  rc = ck_sw_align_device(c->h_pairs.as<int32_t>(), M, c->h_flat.as<uint8_t>(),
                          c->h_off.as<int64_t>(), c->h_len.as<int32_t>(), max_len, band_width,
                          h_sub_matrix, h_diag_hints ? c->h_hint.as<int32_t>() : nullptr,
                          threshold, flags, c->o_sims.as<float>(), c->o_scores.as<int32_t>(),
                          c->o_matches.as<int32_t>(), c->o_mask.as<uint8_t>(), nullptr);
  if (rc) return rc;
  CK_HIP(hipDeviceSynchronize());
  std::lock_guard<std::mutex> lk(c->mu);
  CK_HIP(hipMemcpy(h_out_sims, c->o_sims.ptr, M * sizeof(float), hipMemcpyDeviceToHost));
  CK_HIP(hipMemcpy(h_out_scores, c->o_scores.ptr, M * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (h_out_matches)
    CK_HIP(hipMemcpy(h_out_matches, c->o_matches.ptr, M * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (h_out_mask) CK_HIP(hipMemcpy(h_out_mask, c->o_mask.ptr, M, hipMemcpyDeviceToHost));
  return CK_OK;
}

int ck_batch_sw_align(const int32_t* pairs, const uint8_t* flat_sequences, const int64_t* offsets,
                      const int32_t* lengths, int32_t M, int32_t band_width,
                      const int8_t* sub_matrix, float threshold, const int32_t* diag_hints,
                      float* out_sims, int32_t* out_scores, uint8_t* out_mask) {
  if (M <= 0) return M < 0 ? fail(CK_ERR_INVALID, "negative pair count") : CK_OK;
  if (!pairs || !offsets || !lengths) return fail(CK_ERR_INVALID, "null pointer argument");
  int64_t n = 0;
  for (int64_t p = 0; p < 2 * (int64_t)M; ++p) {
    if (pairs[p] < 0) return fail(CK_ERR_INVALID, "negative pair index");
    n = std::max<int64_t>(n, (int64_t)pairs[p] + 1);
  }
  int64_t flat_len = 0;
  for (int64_t i = 0; i < n; ++i)
    if (lengths[i] > 0) flat_len = std::max<int64_t>(flat_len, offsets[i] + lengths[i]);
  (void)threshold;  // unused by the reference kernel as well
  return ck_sw_align_host(pairs, M, flat_sequences, flat_len, offsets, lengths, n, band_width,
                          sub_matrix, diag_hints, threshold, 0u, out_sims, out_scores, nullptr,
                          out_mask);
}

int ck_jaccard_device(const int32_t* d_pairs, int64_t M, const uint64_t* d_sketches,
                      int32_t sketch_size, float* d_out_sims, void* stream) {
  if (M < 0 || sketch_size <= 0) return fail(CK_ERR_INVALID, "bad size");
  if (M == 0) return CK_OK;
  if (!d_pairs || !d_sketches || !d_out_sims) return fail(CK_ERR_INVALID, "null pointer argument");
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  const int grid = (int)std::min<int64_t>((M + 255) / 256, 16 * (int64_t)c->num_cu);
  hipLaunchKernelGGL(ck_jaccard_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, d_pairs, M,
                     d_sketches, sketch_size, d_out_sims);
  CK_HIP(hipGetLastError());
  return CK_OK;
}

int ck_jaccard_host(const int32_t* h_pairs, int64_t M, const uint64_t* h_sketches, int64_t n_seqs,
                    int32_t sketch_size, float* h_out_sims) {
  if (M < 0 || n_seqs < 0 || sketch_size <= 0) return fail(CK_ERR_INVALID, "bad size");
  if (M == 0) return CK_OK;
  if (!h_pairs || !h_sketches || !h_out_sims) return fail(CK_ERR_INVALID, "null pointer argument");
  for (int64_t p = 0; p < 2 * M; ++p)
    if (h_pairs[p] < 0 || h_pairs[p] >= n_seqs)
      return fail(CK_ERR_INVALID, "pair index %d out of range", h_pairs[p]);
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    const size_t skb = (size_t)n_seqs * sketch_size * sizeof(uint64_t);
    if ((rc = c->h_pairs.ensure(2 * M * sizeof(int32_t))) || (rc = c->j_sk.ensure(skb)) ||
        (rc = c->j_out.ensure(M * sizeof(float))))
      return rc;
    CK_HIP(hipMemcpy(c->h_pairs.ptr, h_pairs, 2 * M * sizeof(int32_t), hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(c->j_sk.ptr, h_sketches, skb, hipMemcpyHostToDevice));
  }
  rc = ck_jaccard_device(c->h_pairs.as<int32_t>(), M, c->j_sk.as<uint64_t>(), sketch_size,
                         c->j_out.as<float>(), nullptr);
  if (rc) return rc;
  CK_HIP(hipDeviceSynchronize());
  CK_HIP(hipMemcpy(h_out_sims, c->j_out.ptr, M * sizeof(float), hipMemcpyDeviceToHost));
  return CK_OK;
}

}  // extern "C"
