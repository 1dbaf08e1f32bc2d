// extern "C" entry points of libclustkit_hip.so (declared in
// include/clustkit_hip.h).  Host-side orchestration only: argument checks,
// per-device workspace, kernel launches on the caller's stream.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <atomic>
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
                                uint8_t*, int32_t*);
template <int S, bool WIDE>
__global__ void sw_wave_kernel(const int32_t*, int32_t*, const int32_t*, const int32_t*,
                               const uint8_t*, const int64_t*, const int32_t*, const int32_t*,
                               int32_t, const int8_t*, float, uint32_t, int32_t, int32_t,
                               float*, int32_t*, int32_t*, uint8_t*, int32_t*, int32_t*);
#define CK_WAVE_EXTERN(S_, W_)                                                                 \
  extern template __global__ void sw_wave_kernel<S_, W_>(                                      \
      const int32_t*, int32_t*, const int32_t*, const int32_t*, const uint8_t*, const int64_t*, \
      const int32_t*, const int32_t*, int32_t, const int8_t*, float, uint32_t, int32_t, int32_t, \
      float*, int32_t*, int32_t*, uint8_t*, int32_t*, int32_t*);
CK_WAVE_EXTERN(16, false)
CK_WAVE_EXTERN(8, false)
#undef CK_WAVE_EXTERN
template <int S, bool WIDE, int MINB, int NT, bool SHORTW>
__global__ void sw_pipe_kernel(const int32_t*, int32_t*, const int32_t*, const int32_t*,
                               const uint8_t*, const int64_t*, const int32_t*, const int32_t*,
                               int32_t, const int8_t*, float, uint32_t, int32_t, int32_t,
                               float*, int32_t*, int32_t*, uint8_t*, int32_t*, int32_t*);
#define CK_PIPE_EXTERN(S_, W_, B_, NT_, SH_)                                                   \
  extern template __global__ void sw_pipe_kernel<S_, W_, B_, NT_, SH_>(                        \
      const int32_t*, int32_t*, const int32_t*, const int32_t*, const uint8_t*, const int64_t*, \
      const int32_t*, const int32_t*, int32_t, const int8_t*, float, uint32_t, int32_t, int32_t, \
      float*, int32_t*, int32_t*, uint8_t*, int32_t*, int32_t*);
CK_PIPE_EXTERN(16, false, 1, WAVE_BLOCK, false)
CK_PIPE_EXTERN(16, false, 4, WAVE_BLOCK, false)
CK_PIPE_EXTERN(16, false, 1, WAVE_BLOCK, true)
CK_PIPE_EXTERN(16, false, 4, WAVE_BLOCK, true)
CK_PIPE_EXTERN(16, false, 1, 1024, false)
CK_PIPE_EXTERN(8, false, 1, WAVE_BLOCK, false)
CK_PIPE_EXTERN(8, false, 4, WAVE_BLOCK, false)
CK_PIPE_EXTERN(64, false, 1, WAVE_BLOCK, false)
CK_PIPE_EXTERN(64, true, 1, WAVE_BLOCK, false)
#undef CK_PIPE_EXTERN
__global__ void nw_prep_kernel(const int32_t*, int64_t, const int64_t*, const int32_t*, int32_t,
                               float, int32_t, int32_t, int32_t, int32_t, int32_t, int32_t*,
                               int32_t*, Counters*, int32_t*, float*, int32_t*, uint8_t*);
template <bool SCORED, int S>
__global__ void nw_wave_kernel(const int32_t*, const int32_t*, const int32_t*, int32_t*,
                               const int32_t*, const uint8_t*, const int64_t*, const int32_t*,
                               int32_t, float, int32_t, const int8_t*, int32_t, float*, int32_t*,
                               uint8_t*);
#define CK_NW_WAVE_INST(B_, S_)                                                                  \
  extern template __global__ void nw_wave_kernel<B_, S_>(                                      \
      const int32_t*, const int32_t*, const int32_t*, int32_t*, const int32_t*,                 \
      const uint8_t*, const int64_t*, const int32_t*, int32_t, float, int32_t, const int8_t*,   \
      int32_t, float*, int32_t*, uint8_t*);
CK_NW_WAVE_INST(false, 16)
CK_NW_WAVE_INST(true, 16)
template <bool SCORED, bool PIPE>
__global__ void nw_pack_kernel(const int32_t*, const int32_t*, const int32_t*, const int32_t*,
                               int32_t*, int32_t*, const int32_t*, const uint8_t*,
                               const int64_t*, const int32_t*, int32_t, float, int32_t,
                               const int8_t*, int32_t, int32_t, float*, int32_t*, uint8_t*);
#define CK_NW_PACK_INST(B_, P_)                                                                  \
  extern template __global__ void nw_pack_kernel<B_, P_>(                                      \
      const int32_t*, const int32_t*, const int32_t*, const int32_t*, int32_t*, int32_t*,      \
      const int32_t*, const uint8_t*, const int64_t*, const int32_t*, int32_t, float, int32_t,  \
      const int8_t*, int32_t, int32_t, float*, int32_t*, uint8_t*);
CK_NW_PACK_INST(false, false)
CK_NW_PACK_INST(true, false)
CK_NW_PACK_INST(false, true)
CK_NW_PACK_INST(true, true)
template <bool SCORED>
__global__ void nw_plain_kernel(const int32_t*, const int32_t*, const int32_t*, const uint8_t*,
                                const int64_t*, const int32_t*, int32_t, float, int32_t,
                                const int8_t*, int4*, int32_t, float*, int32_t*, uint8_t*,
                                int32_t*);
#define CK_NW_PLAIN_INST(B_)                                                                    \
  extern template __global__ void nw_plain_kernel<B_>(                                          \
      const int32_t*, const int32_t*, const int32_t*, const uint8_t*, const int64_t*,           \
      const int32_t*, int32_t, float, int32_t, const int8_t*, int4*, int32_t, float*, int32_t*, \
      uint8_t*, int32_t*);
CK_NW_PLAIN_INST(false)
CK_NW_PLAIN_INST(true)
int32_t nw_lds_bytes(int32_t w_cap, int32_t lanes);
int32_t nw_pack_lds_bytes(int32_t w_cap16, int32_t w_cap8, bool pipe);
__global__ void ck_reset_kernel(Counters*, int32_t*, int32_t);
__global__ void ck_prep_kernel(const int32_t*, int64_t, const int32_t*, const int32_t*,
                               int32_t, float, uint32_t, int32_t, int32_t, int32_t, int32_t,
                               int32_t*, int32_t*, Counters*, int32_t*, float*, int32_t*,
                               int32_t*, uint8_t*);
__global__ void ck_scan_kernel(int32_t*, Counters*);
__global__ void ck_scatter_kernel(const int32_t*, int64_t, int32_t*, int32_t*);
__global__ void ck_jaccard_kernel(const int32_t*, int64_t, const uint64_t*, int32_t, float*);
__global__ void km_qcodes_kernel(const uint8_t*, const int64_t*, const int32_t*, int32_t, int32_t,
                                 int32_t, const int32_t*, int32_t, int32_t*);
__global__ void km_fill_kernel(uint32_t*, int64_t, uint32_t);
__global__ void km_similar_kernel(const uint8_t*, const int64_t*, const int32_t*, const int64_t*,
                                  const int32_t*, const int32_t*, int32_t, KmSub400, int32_t,
                                  int32_t, int32_t, uint32_t*);
__global__ void km_phase_a_tiled_kernel(const int32_t*, const int64_t*, const int32_t*, int32_t,
                                        const int64_t*, int32_t, const int32_t*, int32_t,
                                        uint32_t*);
__global__ void km_phase_a_global_kernel(const int32_t*, const int64_t*, const int32_t*, int32_t,
                                         const int64_t*, const int32_t*, int32_t, uint32_t*);
__global__ void km_sel_hist_kernel(SelectArgs);
__global__ void km_sel_cut_kernel(SelectArgs);
__global__ void km_sel_count_kernel(SelectArgs);
__global__ void km_sel_emit_kernel(SelectArgs);
__global__ void km_sel_hoff_kernel(SelectArgs);
template <bool SLOT16>
__global__ void km_phase_b_tiled_kernel(const int32_t*, const int64_t*, const int32_t*, int32_t,
                                        const int64_t*, int32_t, const int32_t*, const int32_t*,
                                        int32_t, const uint32_t*, const int32_t*, int32_t,
                                        int32_t, int32_t*, int64_t);
extern template __global__ void km_phase_b_tiled_kernel<true>(
    const int32_t*, const int64_t*, const int32_t*, int32_t, const int64_t*, int32_t,
    const int32_t*, const int32_t*, int32_t, const uint32_t*, const int32_t*, int32_t, int32_t,
    int32_t*, int64_t);
extern template __global__ void km_phase_b_tiled_kernel<false>(
    const int32_t*, const int64_t*, const int32_t*, int32_t, const int64_t*, int32_t,
    const int32_t*, const int32_t*, int32_t, const uint32_t*, const int32_t*, int32_t, int32_t,
    int32_t*, int64_t);
__global__ void km_phase_b_global_kernel(const int32_t*, const int64_t*, const int32_t*, int32_t,
                                         const int64_t*, const int32_t*, const int32_t*, int32_t,
                                         const uint32_t*, const int32_t*, int32_t, int32_t,
                                         int32_t*, int64_t);
__global__ void km_phase_b_best_kernel(const int32_t*, const int32_t*, int32_t, int32_t*, int64_t,
                                       uint32_t*, int32_t*);
template <int K, bool FDIV, bool DOT, bool CODES>
__global__ void km_phase_b_target_kernel(const int32_t*, const int64_t*, const int32_t*, int32_t,
                                         int32_t, const uint8_t*, int64_t, const int16_t*,
                                         const int64_t*, const int32_t*, const int32_t*,
                                         const int32_t*, int32_t, int32_t, int32_t, int32_t,
                                         uint32_t*, int32_t*);
#define KM_TGT_EXTERN(K, F, D, C)                                                             \
  extern template __global__ void km_phase_b_target_kernel<K, F, D, C>(                       \
      const int32_t*, const int64_t*, const int32_t*, int32_t, int32_t, const uint8_t*,       \
      int64_t, const int16_t*, const int64_t*, const int32_t*, const int32_t*, const int32_t*, \
      int32_t, int32_t, int32_t, int32_t, uint32_t*, int32_t*);
KM_TGT_EXTERN(1, false, false, false)
KM_TGT_EXTERN(2, false, false, false)
KM_TGT_EXTERN(3, false, false, false)
KM_TGT_EXTERN(4, false, false, false)
KM_TGT_EXTERN(5, false, false, false)
KM_TGT_EXTERN(1, true, false, false)
KM_TGT_EXTERN(2, true, false, false)
KM_TGT_EXTERN(3, true, false, false)
KM_TGT_EXTERN(4, true, false, false)
KM_TGT_EXTERN(5, true, false, false)
KM_TGT_EXTERN(3, false, true, false)
KM_TGT_EXTERN(3, true, true, false)
KM_TGT_EXTERN(1, false, false, true)
KM_TGT_EXTERN(2, false, false, true)
KM_TGT_EXTERN(3, false, false, true)
KM_TGT_EXTERN(4, false, false, true)
KM_TGT_EXTERN(5, false, false, true)
KM_TGT_EXTERN(1, true, false, true)
KM_TGT_EXTERN(2, true, false, true)
KM_TGT_EXTERN(3, true, false, true)
KM_TGT_EXTERN(4, true, false, true)
KM_TGT_EXTERN(5, true, false, true)
#undef KM_TGT_EXTERN
__global__ void km_finalize_kernel(const int32_t*, const int32_t*, const int32_t*, int32_t,
                                   const uint32_t*, int64_t, const int32_t*, int32_t,
                                   const int32_t*, const int32_t*, int32_t, int32_t, int32_t*,
                                   int32_t*, int32_t*, int32_t*);
__global__ void ck_search_pairs_kernel(const int32_t*, const int32_t*, int32_t, int32_t, int32_t,
                                       int32_t, int32_t*);
constexpr int KM_TILE_HOST = 8192;
constexpr int KM_TGT_CODES = 8192;      // LDS code table of the target-major Phase B
constexpr int KM_TGT_BINS_HOST = 512;   // per-wavefront histogram bins of that kernel
constexpr int KM_TGT_MAX_KMERS = 16383; // query k-mer positions of the target-major table
hipError_t kmer_index_build(const uint8_t*, const int64_t*, const int32_t*, int32_t, int32_t,
                            int32_t, int64_t, int64_t*, int32_t*, int32_t*, int32_t*, int64_t*,
                            int*, hipStream_t);
hipError_t kmer_tile_offsets(const int64_t*, const int32_t*, int64_t, int32_t, int64_t*,
                             hipStream_t);
hipError_t kmer_db_codes(const uint8_t*, const int64_t*, const int32_t*, int32_t, int32_t, int32_t,
                         const int64_t*, int16_t*, hipStream_t);
hipError_t sketch_build(const uint8_t*, const int64_t*, const int32_t*, int64_t, int32_t,
                        int32_t, int32_t, uint64_t, uint64_t*, hipStream_t);
int32_t sketch_max_size();
hipError_t search_topk(const int32_t*, const int32_t*, const float*, const int32_t*,
                       const uint8_t*, const int32_t*, const int32_t*, const float*, int64_t,
                       int32_t, int32_t, int32_t, int32_t, int32_t*, float*, float*, int32_t*,
                       int32_t*, hipStream_t);
int32_t topk_max_candidates();
struct GraphWs;
GraphWs* graph_ws_new();
void graph_ws_free(GraphWs*);
hipError_t graph_csr(GraphWs*, const int32_t*, int64_t, const float*, int64_t, int32_t*, int32_t*,
                     float*, int64_t*, hipStream_t);
struct LshWs;
LshWs* lsh_ws_new();
void lsh_ws_free(LshWs*);
hipError_t lsh_build(LshWs*, const uint64_t*, int64_t, int32_t, const int32_t*, const int64_t*,
                     int32_t, int32_t, int32_t, int64_t, int64_t*, hipStream_t);
hipError_t lsh_unpack(const LshWs*, int64_t, int32_t*, hipStream_t);
}  // namespace ck

using namespace ck;

// Scratch memory of ck_sw_align_device_noalloc: one fixed array per GPU
// (each GPU the library runs on holds its own copy until the process ends).
// Layout, offsets in bytes: Counters at 0, the 400-byte substitution table
// at 256, the cost-bin histogram at 1024, four int32[NOALLOC_MAX_PAIRS]
// per-pair arrays (cost bin, queue, wide-score list, int32 list), then the
// band rows of the one-pair-per-thread int32 kernel.  In host code the name
// `scratch` is not a device address: the per-device address comes from
// hipGetSymbolAddress (DevCtx::scratch).
constexpr int64_t NOALLOC_MAX_PAIRS = int64_t(1) << 20;
constexpr size_t NOALLOC_PLAIN_BYTES = size_t(32) << 20;
constexpr size_t NA_SUB = 256;
constexpr size_t NA_HIST = 1024;
constexpr size_t NA_PAIRS = NA_HIST + NBINS * sizeof(int32_t);
constexpr size_t NA_PLAIN = NA_PAIRS + 4 * NOALLOC_MAX_PAIRS * sizeof(int32_t);
constexpr size_t NOALLOC_BYTES = NA_PLAIN + NOALLOC_PLAIN_BYTES;
static_assert(sizeof(Counters) <= NA_SUB && NA_SUB + 400 <= NA_HIST && NA_PAIRS % 256 == 0 &&
              NA_PLAIN % 256 == 0, "scratch layout");

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

// The substitution table travels as a kernel argument, so the no-allocation
// form needs no staging copy from host memory.
struct Sub400 {
  int8_t v[400];
};

__global__ void ck_sub_store_kernel(Sub400 s, int8_t* __restrict__ dst) {
  for (int x = threadIdx.x; x < 400; x += blockDim.x) dst[x] = s.v[x];
}

namespace {

thread_local std::string g_err = "no error";
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

// Workspace of one stream: work that is enqueued on different streams of a
// device uses different workspaces, so batches may overlap on the GPU.
struct StreamWs {
  int8_t sub_host[400];
  bool sub_valid = false;
  int8_t* sub_pinned = nullptr;
  DevBuf sub, ctr, hist, bin_of, queue, rerun, plain_list, plain_ws;
  DevBuf km_qcode, km_counts, km_surv, km_surv_n, km_surv_tr, km_hoff, km_hist, km_best,
      km_bbin, km_sel, km_sel_n, km_sel_tr, km_ghist, km_plan, km_bcnt;
  size_t km_hist_zeroed = 0;   // bytes of km_hist known to be zero
  size_t km_ghist_zeroed = 0;  // bytes of km_ghist known to be zero
  // Batched SW call captured as a hipGraph (captured on a private stream,
  // replayed on the caller's); sw_key identifies the arguments it was
  // captured with.
  // LSH workspace; it holds the last LSH call's pairs (lsh_m of them, packed
  // lo * lsh_n + hi).
  ck::LshWs* lsh_ws = nullptr;
  ck::GraphWs* graph_ws = nullptr;
  int64_t lsh_m = 0, lsh_n = 0;
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t sw_exec = nullptr;
  std::vector<uint64_t> sw_key;
  bool graph_broken = false;
};

struct DevCtx {
  bool ready = false;
  int device = -1;
  int num_cu = 0;
  unsigned char* scratch = nullptr;   // this device's copy of `scratch`
  // staging for the host-pointer entry points
  DevBuf h_pairs, h_flat, h_off, h_len, h_hint, o_sims, o_scores, o_matches, o_mask,
      j_sk, j_out;
  // host drop-in staging for k-mer scoring
  DevBuf kh_qflat, kh_qoff, kh_qlen, kh_koff, kh_tid, kh_pos, kh_freq, kh_tile, kh_len,
      kh_tg, kh_cn, kh_sc, kh_dg;
  std::vector<std::pair<hipStream_t, StreamWs*>> ws;
  std::mutex mu;
  // Held for the whole of a host-pointer call (copy-in, kernels, sync,
  // copy-out): those calls share the staging buffers above and the null
  // stream's workspace, so two of them on one device run one after the
  // other.  Lock order: host_mu before mu.
  std::mutex host_mu;
  int ws_for(hipStream_t st, StreamWs** out) {
    for (auto& e : ws)
      if (e.first == st) { *out = e.second; return CK_OK; }
    StreamWs* w = new StreamWs();
    if (hipHostMalloc((void**)&w->sub_pinned, 400, hipHostMallocDefault) != hipSuccess)
      return fail(CK_ERR_NOMEM, "pinned allocation failed");
    int rc;
    if ((rc = w->sub.ensure(400)) || (rc = w->ctr.ensure(sizeof(Counters))) ||
        (rc = w->hist.ensure(NBINS * sizeof(int32_t))))
      return rc;
    ws.emplace_back(st, w);
    *out = w;
    return CK_OK;
  }
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
    void* sp = nullptr;
    CK_HIP(hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)));
    c->scratch = static_cast<unsigned char*>(sp);
    c->device = device;
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

int upload_sub(StreamWs* c, const int8_t* h_sub, hipStream_t st) {
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

// Capture `enqueue` into a graph on the workspace's private stream and make
// it the workspace's executable graph (updated in place when the topology
// allows, re-instantiated otherwise).  A capture the runtime refuses marks
// the workspace graph_broken: later calls launch directly.
template <class F>
int capture_sw(StreamWs* w, const std::vector<uint64_t>& key, F&& enqueue) {
  if (!w->cap_stream) CK_HIP(hipStreamCreateWithFlags(&w->cap_stream, hipStreamNonBlocking));
  hipGraph_t g = nullptr;
  if (hipStreamBeginCapture(w->cap_stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    (void)hipGetLastError();
    w->graph_broken = true;
    return CK_OK;
  }
  enqueue(w->cap_stream);
  hipError_t e = hipStreamEndCapture(w->cap_stream, &g);
  if (e != hipSuccess || !g) {
    (void)hipGetLastError();
    if (g) (void)hipGraphDestroy(g);
    w->graph_broken = true;
    return CK_OK;
  }
  bool ok = false;
  if (w->sw_exec) {
    hipGraphNode_t err_node = nullptr;
    hipGraphExecUpdateResult res;
    ok = hipGraphExecUpdate(w->sw_exec, g, &err_node, &res) == hipSuccess &&
         res == hipGraphExecUpdateSuccess;
    if (!ok) {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(w->sw_exec);
      w->sw_exec = nullptr;
    }
  }
  if (!ok) {
    e = hipGraphInstantiate(&w->sw_exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(g);
      w->sw_exec = nullptr;
      return fail(CK_ERR_HIP, "hipGraphInstantiate failed: %s", hipGetErrorString(e));
    }
  }
  (void)hipGraphDestroy(g);
  w->sw_key = key;
  return CK_OK;
}

// Batched-kernel configuration (ck_sw_set_config): lanes per pair (0 = by
// band: 8 for bands of 1..32, else 16), cost-bin width 2^shift steps, and
// whether 16- and 8-lane segments run the stripe-pipelined kernel.
std::atomic<int32_t> g_cfg_lanes{0};
std::atomic<int32_t> g_cfg_shift{DEFAULT_BIN_SHIFT};
std::atomic<int32_t> g_cfg_pipe{1};

int32_t lanes_for(int32_t band) {
  const int32_t l = g_cfg_lanes.load();
  if (l == 8 || l == 16) return l;
  return (band >= 1 && band <= 32) ? 8 : 16;
}

constexpr int64_t PLAIN_BATCH = 500000;           // pairs per plain-form launch
constexpr size_t PLAIN_WS_BUDGET = size_t(1) << 30;  // bytes of int32 band rows

// Buffers one SW call works in: a stream workspace's grow-only buffers
// (ck_sw_align_device) or the fixed scratch array
// (ck_sw_align_device_noalloc).  pw: entries of one int32-kernel band row;
// t_plain: threads of that kernel (its rows take t_plain * pw * 16 bytes).
struct SwBufs {
  Counters* ctr;
  int32_t* hist;
  const int8_t* sub;      // 400-byte table in device memory
  int32_t* bin_of;        // [M]
  int32_t* queue;         // [M]
  int32_t* rerun;         // [M]
  int32_t* plain_list;    // [M]
  int4* plain_ws;         // [t_plain * pw]
  int64_t t_plain;
};

// Enqueue one batched (or plain) SW call on `st`: work-queue construction,
// the batched kernel, the wide-score and int32 re-runs.  w: the stream
// workspace whose captured hipGraph is replayed (nullptr: launch directly).
int sw_run(DevCtx* c, StreamWs* w, const SwBufs& b, const int32_t* d_pairs, int64_t M,
           const uint8_t* d_flat, const int64_t* d_offsets, const int32_t* d_lengths,
           int32_t band_width, const int32_t* d_diag_hints, float threshold, uint32_t flags,
           int32_t w_cap, int32_t pw, float* d_out_sims, int32_t* d_out_scores,
           int32_t* d_out_matches, uint8_t* d_out_mask, hipStream_t st) {
  int rc = CK_OK;
  Counters* ctr = b.ctr;
  int32_t* hist = b.hist;
  const bool plain = (flags & CK_SW_PLAIN) != 0;
  const bool timing = (flags & CK_SW_TIMING) != 0;
  if (timing) {
    for (auto& e : g_ev)
      if (!e) CK_HIP(hipEventCreate(&e));
  }
  g_ev_valid = false;
  const int prep_grid = (int)std::min<int64_t>((M + 255) / 256, 4 * (int64_t)c->num_cu);
  const int scatter_grid = (int)std::min<int64_t>((M + 2047) / 2048, 4 * (int64_t)c->num_cu);
  const int plain_grid = (int)(b.t_plain / 256);
  // Latency form: one pair per wavefront, its 64 lanes one chain
  // (sw_pipe_kernel<64, false>).
  const bool latency = (flags & CK_SW_LATENCY) != 0;
  const int32_t S = latency ? 64 : lanes_for(band_width);
  const int32_t bin_shift = g_cfg_shift.load();
  // Staging buffers of the pipelined kernel with s-lane chains: two per
  // chain, each a group's window (at least 2s steps) plus 3s codes, at the
  // smallest size that is 16 modulo 64 (16-lane chains of one wavefront then
  // sit 32 bytes apart modulo 128, on different LDS banks, and three
  // workgroups fit a CU up to bands of 128).
  auto pipe_bs = [&](int32_t s) {
    return ((std::max(w_cap, 2 * s) + 3 * s + 47) & ~63) + 16;
  };
  const int32_t rb_stride = w_cap + 2 * S;
  // The per-stripe kernel's staging buffers sit 128k + 32 bytes apart.
  auto lds_of = [&](bool piped, int32_t nt = WAVE_BLOCK) -> size_t {
    const int32_t bs = piped ? pipe_bs(S) : (((w_cap + 3 * S) + 127) & ~127) + 32;
    return TABLE_BYTES + (size_t)(nt / S) *
                             (rb_stride * sizeof(int2) + (piped ? 2 : 1) * bs);
  };
  // 8-lane segments run pipelined where their buffers fit a CU's LDS (with
  // full-matrix windows they may not), else one stripe at a time.
  const bool pipe = !latency && (S == 16 || (S == 8 && lds_of(true) <= 160 * 1024)) &&
                    g_cfg_pipe.load() != 0;
  const bool piped = pipe || latency;
  const int32_t bs_stride = piped ? pipe_bs(S) : (((w_cap + 3 * S) + 127) & ~127) + 32;
  // Where four workgroups fit a CU's 160 KB of LDS, the pipelined kernel
  // built for four (registers capped at 128) runs.  Where they do not but one
  // 1024-thread workgroup does (sixteen wavefronts sharing one table), that
  // form runs: also four wavefronts per SIMD, registers capped at 128.
  const bool pipe4 = pipe && lds_of(true) * 4 <= 160 * 1024;
  const bool big = pipe && S == 16 && !pipe4 && lds_of(true, 1024) <= 160 * 1024;
  const int32_t nt = big ? 1024 : WAVE_BLOCK;
  const size_t lds = lds_of(piped, nt);
  const int32_t pairs_per_block = nt / S;
  // Bands whose windows are shorter than 2S + 7 = 39 positions (16-lane
  // segments) take the build whose reduced transitions read and store the
  // row buffer step by step.
  const bool shortw = S == 16 && band_width >= 1 && 2 * band_width + 1 < 39;
  const void* wave_fn =
      latency ? (const void*)sw_pipe_kernel<64, false, 1, WAVE_BLOCK, false>
      : pipe4 ? (S == 8 ? (const void*)sw_pipe_kernel<8, false, 4, WAVE_BLOCK, false>
                 : shortw ? (const void*)sw_pipe_kernel<16, false, 4, WAVE_BLOCK, true>
                          : (const void*)sw_pipe_kernel<16, false, 4, WAVE_BLOCK, false>)
      : big   ? (const void*)sw_pipe_kernel<16, false, 1, 1024, false>
      : pipe  ? (S == 8 ? (const void*)sw_pipe_kernel<8, false, 1, WAVE_BLOCK, false>
                 : shortw ? (const void*)sw_pipe_kernel<16, false, 1, WAVE_BLOCK, true>
                          : (const void*)sw_pipe_kernel<16, false, 1, WAVE_BLOCK, false>)
                        : S == 8 ? (const void*)sw_wave_kernel<8, false>
                                 : (const void*)sw_wave_kernel<16, false>;
  // Wide-score re-run: few, long pairs, one per wavefront with its 64 lanes
  // as one chain (the pass lasts as long as its longest pair).
  const void* wide_fn = (const void*)sw_pipe_kernel<64, true, 1, WAVE_BLOCK, false>;
  const int32_t rb16 = w_cap + 2 * 64;
  const int32_t bs16 = pipe_bs(64);
  const size_t lds16 = TABLE_BYTES + (size_t)(WAVE_BLOCK / 64) *
                           (rb16 * sizeof(int2) + 2 * bs16);
  int grid = 1, wide_grid = 1;
  if (!plain) {
    if (lds > 64 * 1024)
      CK_HIP(hipFuncSetAttribute(wave_fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (lds16 > 64 * 1024)
      CK_HIP(hipFuncSetAttribute(wide_fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)lds16));
    int per_cu = 0, wide_per_cu = 0;
    CK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, wave_fn, nt, lds));
    CK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&wide_per_cu, wide_fn, WAVE_BLOCK,
                                                        lds16));
    if (per_cu < 1) per_cu = 1;
    if (wide_per_cu < 1) wide_per_cu = 1;
    const int64_t need = (M + pairs_per_block - 1) / pairs_per_block;
    grid = (int)std::max<int64_t>(1, std::min<int64_t>(need, (int64_t)per_cu * c->num_cu));
    const int64_t wneed = (M + WAVE_BLOCK / WAVE - 1) / (WAVE_BLOCK / WAVE);
    wide_grid = (int)std::max<int64_t>(
        1, std::min<int64_t>(wneed, (int64_t)wide_per_cu * c->num_cu));
  }

  // The whole call: work-queue construction, the batched (or plain) kernel,
  // the int32 re-run.  ev: timing events (null: none).
  auto enqueue = [&](hipStream_t s2, bool ev) {
    if (ev) (void)hipEventRecord(g_ev[0], s2);
    hipLaunchKernelGGL(ck_reset_kernel, dim3((NBINS + 255) / 256), dim3(256), 0, s2, ctr, hist,
                       NBINS);
    hipLaunchKernelGGL(ck_prep_kernel, dim3(prep_grid), dim3(256), 0, s2, d_pairs, M, d_lengths,
                       d_diag_hints, band_width, threshold, flags, w_cap, plain ? 1 : 0, S,
                       bin_shift, b.bin_of, hist, ctr,
                       b.plain_list, d_out_sims,
                       d_out_scores, d_out_matches, d_out_mask);
    hipLaunchKernelGGL(ck_scan_kernel, dim3(1), dim3(1024), 0, s2, hist, ctr);
    hipLaunchKernelGGL(ck_scatter_kernel, dim3(scatter_grid), dim3(256), 0, s2,
                       b.bin_of, M, hist, b.queue);
    if (ev) (void)hipEventRecord(g_ev[1], s2);
    if (plain) {
      for (int64_t base = 0; base < M; base += PLAIN_BATCH) {
        hipLaunchKernelGGL(sw_plain_kernel, dim3(plain_grid), dim3(256), 0, s2, d_pairs,
                           b.queue, &ctr->queue_len, (int64_t)0, base,
                           PLAIN_BATCH, d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                           b.sub, threshold, flags, b.plain_ws, pw,
                           d_out_sims, d_out_scores, d_out_matches, d_out_mask, &ctr->err);
      }
      if (ev) {
        (void)hipEventRecord(g_ev[2], s2);
        (void)hipEventRecord(g_ev[3], s2);
      }
    } else {
      if (pipe4 && S == 8)
        hipLaunchKernelGGL((sw_pipe_kernel<8, false, 4, WAVE_BLOCK, false>), dim3(grid), dim3(WAVE_BLOCK), lds, s2,
                           b.queue, &ctr->queue_head, &ctr->queue_len, d_pairs,
                           d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                           b.sub, threshold, flags, rb_stride, bs_stride,
                           d_out_sims, d_out_scores, d_out_matches, d_out_mask,
                           b.rerun, &ctr->rerun_len);
      else if (pipe && S == 8)
        hipLaunchKernelGGL((sw_pipe_kernel<8, false, 1, WAVE_BLOCK, false>), dim3(grid), dim3(WAVE_BLOCK), lds, s2,
                           b.queue, &ctr->queue_head, &ctr->queue_len, d_pairs,
                           d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                           b.sub, threshold, flags, rb_stride, bs_stride,
                           d_out_sims, d_out_scores, d_out_matches, d_out_mask,
                           b.rerun, &ctr->rerun_len);
      else if (big)
        hipLaunchKernelGGL((sw_pipe_kernel<16, false, 1, 1024, false>), dim3(grid), dim3(1024), lds, s2,
                           b.queue, &ctr->queue_head, &ctr->queue_len, d_pairs,
                           d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                           b.sub, threshold, flags, rb_stride, bs_stride,
                           d_out_sims, d_out_scores, d_out_matches, d_out_mask,
                           b.rerun, &ctr->rerun_len);
      else if (pipe4 && shortw)
        hipLaunchKernelGGL((sw_pipe_kernel<16, false, 4, WAVE_BLOCK, true>), dim3(grid), dim3(WAVE_BLOCK), lds, s2,
                           b.queue, &ctr->queue_head, &ctr->queue_len, d_pairs,
                           d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                           b.sub, threshold, flags, rb_stride, bs_stride,
                           d_out_sims, d_out_scores, d_out_matches, d_out_mask,
                           b.rerun, &ctr->rerun_len);
      else if (pipe4)
        hipLaunchKernelGGL((sw_pipe_kernel<16, false, 4, WAVE_BLOCK, false>), dim3(grid), dim3(WAVE_BLOCK), lds, s2,
                           b.queue, &ctr->queue_head, &ctr->queue_len, d_pairs,
                           d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                           b.sub, threshold, flags, rb_stride, bs_stride,
                           d_out_sims, d_out_scores, d_out_matches, d_out_mask,
                           b.rerun, &ctr->rerun_len);
      else if (pipe && shortw)
        hipLaunchKernelGGL((sw_pipe_kernel<16, false, 1, WAVE_BLOCK, true>), dim3(grid), dim3(WAVE_BLOCK), lds, s2,
                           b.queue, &ctr->queue_head, &ctr->queue_len, d_pairs,
                           d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                           b.sub, threshold, flags, rb_stride, bs_stride,
                           d_out_sims, d_out_scores, d_out_matches, d_out_mask,
                           b.rerun, &ctr->rerun_len);
      else if (pipe)
        hipLaunchKernelGGL((sw_pipe_kernel<16, false, 1, WAVE_BLOCK, false>), dim3(grid), dim3(WAVE_BLOCK), lds, s2,
                           b.queue, &ctr->queue_head, &ctr->queue_len, d_pairs,
                           d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                           b.sub, threshold, flags, rb_stride, bs_stride,
                           d_out_sims, d_out_scores, d_out_matches, d_out_mask,
                           b.rerun, &ctr->rerun_len);
      else if (latency)
        hipLaunchKernelGGL((sw_pipe_kernel<64, false, 1, WAVE_BLOCK, false>), dim3(grid), dim3(WAVE_BLOCK), lds, s2,
                           b.queue, &ctr->queue_head, &ctr->queue_len, d_pairs,
                           d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                           b.sub, threshold, flags, rb_stride, bs_stride,
                           d_out_sims, d_out_scores, d_out_matches, d_out_mask,
                           b.rerun, &ctr->rerun_len);
      else if (S == 8)
        hipLaunchKernelGGL((sw_wave_kernel<8, false>), dim3(grid), dim3(WAVE_BLOCK), lds, s2,
                           b.queue, &ctr->queue_head, &ctr->queue_len, d_pairs,
                           d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                           b.sub, threshold, flags, rb_stride, bs_stride,
                           d_out_sims, d_out_scores, d_out_matches, d_out_mask,
                           b.rerun, &ctr->rerun_len);
      else
        hipLaunchKernelGGL((sw_wave_kernel<16, false>), dim3(grid), dim3(WAVE_BLOCK), lds, s2,
                           b.queue, &ctr->queue_head, &ctr->queue_len, d_pairs,
                           d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                           b.sub, threshold, flags, rb_stride, bs_stride,
                           d_out_sims, d_out_scores, d_out_matches, d_out_mask,
                           b.rerun, &ctr->rerun_len);
      if (ev) (void)hipEventRecord(g_ev[2], s2);
      // Pairs that reached 32700: wide-score wavefronts; what they cannot
      // finish (>= 65500, raw codes > 30) joins the int32 list.
      hipLaunchKernelGGL((sw_pipe_kernel<64, true, 1, WAVE_BLOCK, false>), dim3(wide_grid), dim3(WAVE_BLOCK), lds16,
                         s2, b.rerun, &ctr->wide_head, &ctr->rerun_len, d_pairs,
                         d_flat, d_offsets, d_lengths, d_diag_hints, band_width,
                         b.sub, threshold, flags, rb16, bs16, d_out_sims,
                         d_out_scores, d_out_matches, d_out_mask, b.plain_list,
                         &ctr->plain_len);
      hipLaunchKernelGGL(sw_plain_kernel, dim3(plain_grid), dim3(256), 0, s2, d_pairs,
                         b.plain_list, &ctr->plain_len, (int64_t)0, (int64_t)0,
                         (int64_t)INT32_MAX, d_flat, d_offsets, d_lengths, d_diag_hints,
                         band_width, b.sub, threshold, flags,
                         b.plain_ws, pw, d_out_sims, d_out_scores, d_out_matches,
                         d_out_mask, &ctr->err);
      if (ev) (void)hipEventRecord(g_ev[3], s2);
    }
  };

  const bool use_graph = w && !timing && !(flags & CK_SW_NO_GRAPH) && !w->graph_broken;
  if (!use_graph) {
    enqueue(st, timing);
  } else {
    uint32_t tbits;
    std::memcpy(&tbits, &threshold, 4);
    const std::vector<uint64_t> key = {
        (uint64_t)d_pairs, (uint64_t)M, (uint64_t)d_flat, (uint64_t)d_offsets,
        (uint64_t)d_lengths, (uint64_t)(uint32_t)band_width, (uint64_t)d_diag_hints, tbits,
        flags, (uint64_t)d_out_sims, (uint64_t)d_out_scores, (uint64_t)d_out_matches,
        (uint64_t)d_out_mask, (uint64_t)b.bin_of, (uint64_t)b.queue,
        (uint64_t)b.rerun, (uint64_t)b.plain_ws, (uint64_t)(uint32_t)w_cap,
        (uint64_t)(uint32_t)pw, (uint64_t)grid, (uint64_t)lds, (uint64_t)plain_grid, (uint64_t)scatter_grid,
        (uint64_t)S, (uint64_t)bin_shift, (uint64_t)b.plain_list, (uint64_t)wide_grid,
        (uint64_t)pipe, (uint64_t)pipe4, (uint64_t)big};
    if (!(w->sw_exec && key == w->sw_key)) {
      if ((rc = capture_sw(w, key, [&](hipStream_t s2) { enqueue(s2, false); }))) return rc;
    }
    if (w->graph_broken) enqueue(st, false);
    else CK_HIP(hipGraphLaunch(w->sw_exec, st));
  }
  CK_HIP(hipGetLastError());
  g_stats_ctr = ctr;
  g_ev_valid = timing;
  return CK_OK;
}

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

int ck_device_props(int device, int32_t* num_cu, int32_t* clock_khz, int64_t* hbm_bytes,
                    char* arch_name, int32_t arch_name_len) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
    return fail(CK_ERR_NODEVICE, "no GPU visible to HIP");
  if (device < 0 || device >= n) return fail(CK_ERR_INVALID, "device %d out of range", device);
  hipDeviceProp_t prop;
  CK_HIP(hipGetDeviceProperties(&prop, device));
  if (num_cu) *num_cu = prop.multiProcessorCount;
  if (clock_khz) *clock_khz = prop.clockRate;
  if (hbm_bytes) *hbm_bytes = (int64_t)prop.totalGlobalMem;
  if (arch_name && arch_name_len > 0) {
    std::strncpy(arch_name, prop.gcnArchName, (size_t)arch_name_len - 1);
    arch_name[arch_name_len - 1] = 0;
  }
  return CK_OK;
}

int ck_sw_set_config(int32_t lanes_per_pair, int32_t bin_shift, int32_t pipelined) {
  if (lanes_per_pair != 0 && lanes_per_pair != 8 && lanes_per_pair != 16)
    return fail(CK_ERR_INVALID, "lanes per pair must be 0 (by band), 8 or 16");
  if (bin_shift < 0 || bin_shift > 12)
    return fail(CK_ERR_INVALID, "bin shift must be within 0..12");
  if (pipelined != 0 && pipelined != 1)
    return fail(CK_ERR_INVALID, "pipelined must be 0 or 1");
  g_cfg_lanes.store(lanes_per_pair);
  g_cfg_shift.store(bin_shift);
  g_cfg_pipe.store(pipelined);
  return CK_OK;
}

int ck_sw_get_config(int32_t* lanes_per_pair, int32_t* bin_shift, int32_t* pipelined) {
  if (!lanes_per_pair || !bin_shift || !pipelined) return fail(CK_ERR_INVALID, "null pointer");
  *lanes_per_pair = g_cfg_lanes.load();
  *bin_shift = g_cfg_shift.load();
  *pipelined = g_cfg_pipe.load();
  return CK_OK;
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
  StreamWs* w;
  if ((rc = c->ws_for(st, &w))) return rc;
  if ((rc = upload_sub(w, h_sub_matrix, st))) return rc;
  if ((rc = w->bin_of.ensure(M * sizeof(int32_t))) || (rc = w->queue.ensure(M * sizeof(int32_t))) ||
      (rc = w->rerun.ensure(M * sizeof(int32_t))) ||
      (rc = w->plain_list.ensure(M * sizeof(int32_t))))
    return rc;
  const int32_t w_cap = fast_w_cap(band_width, max_len);
  const int32_t pw = plain_width(band_width, max_len);

  // Plain-kernel workspace: T threads x pw band entries x 16 bytes.
  int64_t t_plain = (int64_t)(PLAIN_WS_BUDGET / ((size_t)pw * sizeof(int4)));
  t_plain = std::min<int64_t>(t_plain, 256 * 1024);
  t_plain = std::max<int64_t>(256, (t_plain / 256) * 256);
  if ((rc = w->plain_ws.ensure((size_t)t_plain * pw * sizeof(int4)))) return rc;
  const SwBufs b{w->ctr.as<Counters>(), w->hist.as<int32_t>(), w->sub.as<int8_t>(),
                 w->bin_of.as<int32_t>(), w->queue.as<int32_t>(), w->rerun.as<int32_t>(),
                 w->plain_list.as<int32_t>(), w->plain_ws.as<int4>(), t_plain};
  return sw_run(c, w, b, d_pairs, M, d_flat, d_offsets, d_lengths, band_width, d_diag_hints,
                threshold, flags, w_cap, pw, d_out_sims, d_out_scores, d_out_matches,
                d_out_mask, st);
}

int ck_sw_align_device_noalloc(const int32_t* d_pairs, int64_t M, const uint8_t* d_flat,
                               const int64_t* d_offsets, const int32_t* d_lengths,
                               int32_t max_len, int32_t band_width, const int8_t* h_sub_matrix,
                               const int32_t* d_diag_hints, float threshold, uint32_t flags,
                               float* d_out_sims, int32_t* d_out_scores,
                               int32_t* d_out_matches, uint8_t* d_out_mask, void* stream) {
  if (M < 0) return fail(CK_ERR_INVALID, "negative pair count");
  if (M > NOALLOC_MAX_PAIRS)
    return fail(CK_ERR_INVALID, "%lld pairs: the no-allocation form takes at most %lld",
                (long long)M, (long long)NOALLOC_MAX_PAIRS);
  if (M == 0) return CK_OK;
  if (!d_pairs || !d_flat || !d_offsets || !d_lengths || !h_sub_matrix || !d_out_sims ||
      !d_out_scores)
    return fail(CK_ERR_INVALID, "null pointer argument");
  if (max_len < 0) return fail(CK_ERR_INVALID, "negative max_len");
  if (flags & CK_SW_TIMING)
    return fail(CK_ERR_INVALID, "CK_SW_TIMING is not available in the no-allocation form");
  const int32_t w_cap = fast_w_cap(band_width, max_len);
  const int32_t pw = plain_width(band_width, max_len);
  int64_t t_plain = (int64_t)(NOALLOC_PLAIN_BYTES / ((size_t)pw * sizeof(int4)));
  t_plain = (std::min<int64_t>(t_plain, 256 * 1024) / 256) * 256;
  if (t_plain < 256)
    return fail(CK_ERR_INVALID,
                "band rows of %d entries exceed the no-allocation workspace (at most %lld: "
                "band_width or, with band_width <= 0, max_len at most %lld)",
                pw, (long long)(NOALLOC_PLAIN_BYTES / (256 * sizeof(int4))),
                (long long)(NOALLOC_PLAIN_BYTES / (256 * sizeof(int4)) / 2 - 1));
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  unsigned char* base = c->scratch;
  int32_t* per_pair = reinterpret_cast<int32_t*>(base + NA_PAIRS);
  const SwBufs b{reinterpret_cast<Counters*>(base), reinterpret_cast<int32_t*>(base + NA_HIST),
                 reinterpret_cast<int8_t*>(base + NA_SUB), per_pair,
                 per_pair + NOALLOC_MAX_PAIRS, per_pair + 2 * NOALLOC_MAX_PAIRS,
                 per_pair + 3 * NOALLOC_MAX_PAIRS, reinterpret_cast<int4*>(base + NA_PLAIN),
                 t_plain};
  Sub400 sv;
  std::memcpy(sv.v, h_sub_matrix, sizeof sv.v);
  hipStream_t st = (hipStream_t)stream;
  std::lock_guard<std::mutex> lk(c->mu);
  hipLaunchKernelGGL(ck_sub_store_kernel, dim3(1), dim3(256), 0, st, sv,
                     reinterpret_cast<int8_t*>(base + NA_SUB));
  return sw_run(c, nullptr, b, d_pairs, M, d_flat, d_offsets, d_lengths, band_width,
                d_diag_hints, threshold, flags, w_cap, pw, d_out_sims, d_out_scores,
                d_out_matches, d_out_mask, st);
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
  if (rerun_pairs) *rerun_pairs = h.rerun_len + h.prep_plain;
  if (skipped_pairs) *skipped_pairs = h.skipped;
  return CK_OK;
}

int ck_sw_last_paths(int64_t* wide_pairs, int64_t* int32_pairs) {
  if (!g_stats_ctr) return fail(CK_ERR_INVALID, "no ck_sw_align_device call on this thread");
  Counters h;
  CK_HIP(hipMemcpy(&h, g_stats_ctr, sizeof h, hipMemcpyDeviceToHost));
  if (wide_pairs) *wide_pairs = h.rerun_len;
  if (int32_pairs) *int32_pairs = h.plain_len;
  return CK_OK;
}

int ck_sw_last_status(void) {
  if (!g_stats_ctr) return fail(CK_ERR_INVALID, "no alignment call on this thread");
  Counters h;
  CK_HIP(hipMemcpy(&h, g_stats_ctr, sizeof h, hipMemcpyDeviceToHost));
  if (h.err) return fail(CK_ERR_INVALID, "band row wider than the workspace (max_len too small)");
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
  if (flat_len > 0 && !h_flat) return fail(CK_ERR_INVALID, "null flat buffer");
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
  std::lock_guard<std::mutex> host_lk(c->host_mu);
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
  Counters h;
  CK_HIP(hipMemcpy(&h, g_stats_ctr, sizeof h, hipMemcpyDeviceToHost));
  if (h.err) return fail(CK_ERR_INVALID, "band row wider than the workspace (max_len too small)");
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
  if (!pairs || !offsets || !lengths || !sub_matrix || !out_sims || !out_scores || !out_mask)
    return fail(CK_ERR_INVALID, "null pointer argument");
  int64_t n = 0;
  for (int64_t p = 0; p < 2 * (int64_t)M; ++p) {
    if (pairs[p] < 0) return fail(CK_ERR_INVALID, "negative pair index");
    n = std::max<int64_t>(n, (int64_t)pairs[p] + 1);
  }
  int64_t flat_len = 0;
  for (int64_t i = 0; i < n; ++i)
    if (lengths[i] > 0) flat_len = std::max<int64_t>(flat_len, offsets[i] + lengths[i]);
  if (flat_len > 0 && !flat_sequences) return fail(CK_ERR_INVALID, "null flat_sequences");
  (void)threshold;  // unused by the reference kernel as well
  return ck_sw_align_host(pairs, M, flat_sequences, flat_len, offsets, lengths, n, band_width,
                          sub_matrix, diag_hints, threshold, 0u, out_sims, out_scores, nullptr,
                          out_mask);
}

// Banded global alignment (NW identity), nw_wave.hip.  Band windows of up
// to NW_MAX_W positions run out of LDS (the int32 form's 16 segments of
// (W + 1) * 16 + (W + 51) * 4 bytes per 256-thread workgroup fit 160 KB):
// windows of up to NW_NARROW_W positions 8 lanes per pair, wider ones 16;
// wider windows one pair per thread.
constexpr int32_t NW_MAX_W = 479;
constexpr int32_t NW_NARROW_W = 63;

int ck_nw_align_device(const int32_t* d_pairs, int64_t M, const uint8_t* d_flat,
                       const int64_t* d_offsets, const int32_t* d_lengths, int32_t max_len,
                       int32_t band_width, const int8_t* h_sub_matrix, float threshold,
                       uint32_t flags, float* d_out_sims, int32_t* d_out_scores,
                       uint8_t* d_out_mask, void* stream) {
  if (M < 0) return fail(CK_ERR_INVALID, "negative pair count");
  if (M > INT32_MAX) return fail(CK_ERR_INVALID, "pair count above 2^31-1");
  if (M == 0) return CK_OK;
  if (!d_pairs || !d_flat || !d_offsets || !d_lengths || !d_out_sims || !d_out_mask)
    return fail(CK_ERR_INVALID, "null pointer argument");
  if (max_len < 0) return fail(CK_ERR_INVALID, "negative max_len");
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  StreamWs* w;
  if ((rc = c->ws_for(st, &w))) return rc;
  const bool scored = h_sub_matrix != nullptr;
  if (scored && (rc = upload_sub(w, h_sub_matrix, st))) return rc;
  if ((rc = w->bin_of.ensure(M * sizeof(int32_t))) || (rc = w->queue.ensure(M * sizeof(int32_t))) ||
      (rc = w->plain_list.ensure(M * sizeof(int32_t))))
    return rc;
  Counters* ctr = w->ctr.as<Counters>();
  int32_t* hist = w->hist.as<int32_t>();
  const bool plain = (flags & CK_NW_PLAIN) != 0;
  const bool pipe = (flags & CK_NW_STRIPE) == 0;
  const int32_t raw = (flags & CK_NW_RAW) ? 1 : 0;
  // Widest band row any pair can take: the adaptive band is at most
  // band_width; pairs of at most 50 residues take the full matrix.
  const int32_t ml = std::max(max_len, 1);
  const int64_t wmax = band_width > 0
                           ? 2 * (int64_t)std::max(std::min(band_width, ml), std::min(ml, 50)) + 1
                           : 2 * (int64_t)ml + 1;
  const int32_t w_cap = plain ? 0 : (int32_t)std::min<int64_t>(wmax, NW_MAX_W);
  // Bands of up to NW_NARROW_W positions run 8 lanes per pair (shorter ramp
  // per stripe, smaller LDS buffers), wider ones 16.
  const int32_t w_narrow = std::min(w_cap, NW_NARROW_W);
  const int32_t pw = (int32_t)wmax + 1;
  int64_t t_plain = (int64_t)(PLAIN_WS_BUDGET / ((size_t)pw * sizeof(int4)));
  t_plain = std::min<int64_t>(t_plain, 256 * 1024);
  t_plain = std::max<int64_t>(256, (t_plain / 256) * 256);
  if ((rc = w->plain_ws.ensure((size_t)t_plain * pw * sizeof(int4)))) return rc;
  const int prep_grid = (int)std::min<int64_t>((M + 255) / 256, 4 * (int64_t)c->num_cu);
  const int scatter_grid = (int)std::min<int64_t>((M + 2047) / 2048, 4 * (int64_t)c->num_cu);
  const int plain_grid = (int)(t_plain / 256);
  // Score range of the table (plain form: +2 / -1), for the packed-key test.
  int32_t smax = 2, smin = -1;
  if (scored) {
    smax = -4;
    smin = -4;
    for (int x = 0; x < 400; ++x) {
      smax = std::max<int32_t>(smax, h_sub_matrix[x]);
      smin = std::min<int32_t>(smin, h_sub_matrix[x]);
    }
  }
  // The pipelined schedule needs groups of at least 2S + 1 steps: window
  // capacities of at least 33 (16 lanes) and 17 (8 lanes).
  const int32_t wn = pipe ? std::max(w_narrow, 17) : std::max(w_narrow, 1);
  const int32_t wc = pipe ? std::max(w_cap, 33) : w_cap;
  const void* fnw = scored ? (const void*)nw_wave_kernel<true, 16>
                           : (const void*)nw_wave_kernel<false, 16>;
  const void* fp = scored ? (pipe ? (const void*)nw_pack_kernel<true, true>
                                  : (const void*)nw_pack_kernel<true, false>)
                          : (pipe ? (const void*)nw_pack_kernel<false, true>
                                  : (const void*)nw_pack_kernel<false, false>);
  const size_t ldsw = plain ? 0 : (size_t)nw_lds_bytes(w_cap, 16);
  const size_t ldsp = plain ? 0 : (size_t)nw_pack_lds_bytes(wc, wn, pipe);
  int gridw = 1, gridp = 1;
  if (!plain) {
    auto grid_for = [&](const void* fn, size_t lds, int pairs_per_block, int* g) -> int {
      if (lds > 64 * 1024)
        CK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      int pc = 0;
      CK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, fn, WAVE_BLOCK, lds));
      pc = std::max(pc, 1);
      *g = (int)std::max<int64_t>(
          1, std::min<int64_t>((M + pairs_per_block - 1) / pairs_per_block,
                               (int64_t)pc * c->num_cu));
      return CK_OK;
    };
    if ((rc = grid_for(fnw, ldsw, WAVE_BLOCK / 16, &gridw)) ||
        (rc = grid_for(fp, ldsp, WAVE_BLOCK / 16, &gridp)))
      return rc;
  }
  const int8_t* dsub = scored ? w->sub.as<int8_t>() : nullptr;
  hipLaunchKernelGGL(ck_reset_kernel, dim3((NBINS + 255) / 256), dim3(256), 0, st, ctr, hist,
                     NBINS);
  hipLaunchKernelGGL(nw_prep_kernel, dim3(prep_grid), dim3(256), 0, st, d_pairs, M, d_offsets,
                     d_lengths, band_width, threshold, raw, w_cap, w_narrow, smax, smin,
                     w->bin_of.as<int32_t>(), hist, ctr,
                     w->plain_list.as<int32_t>(), d_out_sims, d_out_scores, d_out_mask);
  if (!plain) {
    hipLaunchKernelGGL(ck_scan_kernel, dim3(1), dim3(1024), 0, st, hist, ctr);
    hipLaunchKernelGGL(ck_scatter_kernel, dim3(scatter_grid), dim3(256), 0, st,
                       w->bin_of.as<int32_t>(), M, hist, w->queue.as<int32_t>());
    // After the scatter hist[b] is the end of bin b: the three kernel classes
    // hold queue ranges [0, e0), [e0, e1), [e1, queue_len).
    const int32_t* e0 = hist + NBINS / 4 - 1;
    const int32_t* e1 = hist + NBINS / 2 - 1;
    const int32_t* qend = &ctr->queue_len;
#define NW_LAUNCH(K_, G_, L_, BEG_, END_, HEAD_, W_)                                             \
    hipLaunchKernelGGL(K_, dim3(G_), dim3(WAVE_BLOCK), L_, st, w->queue.as<int32_t>(), BEG_,    \
                       END_, HEAD_, d_pairs, d_flat, d_offsets, d_lengths, band_width,          \
                       threshold, raw, dsub, W_, d_out_sims, d_out_scores, d_out_mask)
#define NW_PACK_LAUNCH(K_)                                                                       \
    hipLaunchKernelGGL(K_, dim3(gridp), dim3(WAVE_BLOCK), ldsp, st, w->queue.as<int32_t>(), e0,   \
                       e1, qend, &ctr->wide_head, &ctr->aux_head, d_pairs, d_flat, d_offsets,  \
                       d_lengths, band_width, threshold, raw, dsub, wc, wn, d_out_sims,         \
                       d_out_scores, d_out_mask)
    if (scored) {
      NW_LAUNCH((nw_wave_kernel<true, 16>), gridw, ldsw, (const int32_t*)nullptr, e0,
                &ctr->queue_head, w_cap);
      if (pipe) NW_PACK_LAUNCH((nw_pack_kernel<true, true>));
      else NW_PACK_LAUNCH((nw_pack_kernel<true, false>));
    } else {
      NW_LAUNCH((nw_wave_kernel<false, 16>), gridw, ldsw, (const int32_t*)nullptr, e0,
                &ctr->queue_head, w_cap);
      if (pipe) NW_PACK_LAUNCH((nw_pack_kernel<false, true>));
      else NW_PACK_LAUNCH((nw_pack_kernel<false, false>));
    }
#undef NW_PACK_LAUNCH
#undef NW_LAUNCH
  }
  if (scored)
    hipLaunchKernelGGL(nw_plain_kernel<true>, dim3(plain_grid), dim3(256), 0, st,
                       w->plain_list.as<int32_t>(), &ctr->plain_len, d_pairs, d_flat, d_offsets,
                       d_lengths, band_width, threshold, raw, dsub, w->plain_ws.as<int4>(), pw,
                       d_out_sims, d_out_scores, d_out_mask, &ctr->err);
  else
    hipLaunchKernelGGL(nw_plain_kernel<false>, dim3(plain_grid), dim3(256), 0, st,
                       w->plain_list.as<int32_t>(), &ctr->plain_len, d_pairs, d_flat, d_offsets,
                       d_lengths, band_width, threshold, raw, dsub, w->plain_ws.as<int4>(), pw,
                       d_out_sims, d_out_scores, d_out_mask, &ctr->err);
  CK_HIP(hipGetLastError());
  g_stats_ctr = ctr;
  g_ev_valid = false;
  return CK_OK;
}

int ck_nw_align_host(const int32_t* h_pairs, int64_t M, const uint8_t* h_flat, int64_t flat_len,
                     const int64_t* h_offsets, const int32_t* h_lengths, int64_t n_seqs,
                     int32_t band_width, const int8_t* h_sub_matrix, float threshold,
                     uint32_t flags, float* h_out_sims, int32_t* h_out_scores,
                     uint8_t* h_out_mask) {
  if (M < 0 || n_seqs < 0 || flat_len < 0) return fail(CK_ERR_INVALID, "negative size");
  if (M == 0) return CK_OK;
  if (!h_pairs || !h_offsets || !h_lengths || !h_out_sims || !h_out_mask)
    return fail(CK_ERR_INVALID, "null pointer argument");
  if (flat_len > 0 && !h_flat) return fail(CK_ERR_INVALID, "null flat buffer");
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
  std::lock_guard<std::mutex> host_lk(c->host_mu);
  {
    std::lock_guard<std::mutex> lk(c->mu);
    if ((rc = c->h_pairs.ensure(2 * M * sizeof(int32_t))) ||
        (rc = c->h_flat.ensure(std::max<int64_t>(flat_len, 1))) ||
        (rc = c->h_off.ensure(n_seqs * sizeof(int64_t))) ||
        (rc = c->h_len.ensure(n_seqs * sizeof(int32_t))) ||
        (rc = c->o_sims.ensure(M * sizeof(float))) ||
        (rc = c->o_scores.ensure(M * sizeof(int32_t))) || (rc = c->o_mask.ensure(M)))
      return rc;
    CK_HIP(hipMemcpy(c->h_pairs.ptr, h_pairs, 2 * M * sizeof(int32_t), hipMemcpyHostToDevice));
    if (flat_len) CK_HIP(hipMemcpy(c->h_flat.ptr, h_flat, flat_len, hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(c->h_off.ptr, h_offsets, n_seqs * sizeof(int64_t), hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(c->h_len.ptr, h_lengths, n_seqs * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  rc = ck_nw_align_device(c->h_pairs.as<int32_t>(), M, c->h_flat.as<uint8_t>(),
                          c->h_off.as<int64_t>(), c->h_len.as<int32_t>(), max_len, band_width,
                          h_sub_matrix, threshold, flags, c->o_sims.as<float>(),
                          h_out_scores ? c->o_scores.as<int32_t>() : nullptr,
                          c->o_mask.as<uint8_t>(), nullptr);
  if (rc) return rc;
  CK_HIP(hipDeviceSynchronize());
  std::lock_guard<std::mutex> lk(c->mu);
  Counters h;
  CK_HIP(hipMemcpy(&h, g_stats_ctr, sizeof h, hipMemcpyDeviceToHost));
  if (h.err) return fail(CK_ERR_INVALID, "band row wider than the workspace (max_len too small)");
  CK_HIP(hipMemcpy(h_out_sims, c->o_sims.ptr, M * sizeof(float), hipMemcpyDeviceToHost));
  if (h_out_scores)
    CK_HIP(hipMemcpy(h_out_scores, c->o_scores.ptr, M * sizeof(int32_t), hipMemcpyDeviceToHost));
  CK_HIP(hipMemcpy(h_out_mask, c->o_mask.ptr, M, hipMemcpyDeviceToHost));
  return CK_OK;
}

int ck_graph_csr_device(const int32_t* d_pairs, int64_t K, const float* d_sims, int64_t N,
                        int32_t* d_indptr, int32_t* d_indices, float* d_data, int64_t* d_nnz,
                        void* stream) {
  if (K < 0 || N < 0) return fail(CK_ERR_INVALID, "negative size");
  if (N >= INT32_MAX) return fail(CK_ERR_INVALID, "more than 2^31-2 sequences");
  if (2 * K >= INT32_MAX) return fail(CK_ERR_INVALID, "more than 2^30 edges");
  if (!d_indptr || !d_nnz || (K > 0 && (!d_pairs || !d_sims || !d_indices || !d_data)))
    return fail(CK_ERR_INVALID, "null pointer argument");
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  StreamWs* w;
  if ((rc = c->ws_for(st, &w))) return rc;
  if (!w->graph_ws) w->graph_ws = ck::graph_ws_new();
  CK_HIP(ck::graph_csr(w->graph_ws, d_pairs, K, d_sims, N, d_indptr, d_indices, d_data, d_nnz,
                       st));
  return CK_OK;
}

int ck_graph_csr_host(const int32_t* h_pairs, int64_t K, const float* h_sims, int64_t N,
                      int32_t* h_indptr, int32_t* h_indices, float* h_data, int64_t* h_nnz) {
  if (K < 0 || N < 0) return fail(CK_ERR_INVALID, "negative size");
  if (!h_indptr || !h_nnz || (K > 0 && (!h_pairs || !h_sims || !h_indices || !h_data)))
    return fail(CK_ERR_INVALID, "null pointer argument");
  for (int64_t p = 0; p < 2 * K; ++p)
    if (h_pairs[p] < 0 || h_pairs[p] >= N)
      return fail(CK_ERR_INVALID, "pair index %d out of range", h_pairs[p]);
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  std::lock_guard<std::mutex> host_lk(c->host_mu);
  // Staging in the context's grow-only buffers (pairs, sims, indptr,
  // indices, data, nnz).
  DevBuf &bp = c->h_pairs, &bs = c->o_sims, &bi = c->h_len, &bx = c->h_hint, &bd = c->j_out,
         &bn = c->o_scores;
  {
    std::lock_guard<std::mutex> lk(c->mu);
    if ((rc = bp.ensure(std::max<int64_t>(1, 2 * K) * sizeof(int32_t))) ||
        (rc = bs.ensure(std::max<int64_t>(1, K) * sizeof(float))) ||
        (rc = bi.ensure((N + 1) * sizeof(int32_t))) ||
        (rc = bx.ensure(std::max<int64_t>(1, 2 * K) * sizeof(int32_t))) ||
        (rc = bd.ensure(std::max<int64_t>(1, 2 * K) * sizeof(float))) ||
        (rc = bn.ensure(sizeof(int64_t))))
      return rc;
    if (K) {
      CK_HIP(hipMemcpy(bp.ptr, h_pairs, 2 * K * sizeof(int32_t), hipMemcpyHostToDevice));
      CK_HIP(hipMemcpy(bs.ptr, h_sims, K * sizeof(float), hipMemcpyHostToDevice));
    }
  }
  rc = ck_graph_csr_device(bp.as<int32_t>(), K, bs.as<float>(), N, bi.as<int32_t>(),
                           bx.as<int32_t>(), bd.as<float>(), bn.as<int64_t>(), nullptr);
  if (rc) return rc;
  CK_HIP(hipDeviceSynchronize());
  std::lock_guard<std::mutex> lk(c->mu);
  int64_t nnz = 0;
  CK_HIP(hipMemcpy(&nnz, bn.ptr, sizeof nnz, hipMemcpyDeviceToHost));
  CK_HIP(hipMemcpy(h_indptr, bi.ptr, (N + 1) * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (nnz) {
    CK_HIP(hipMemcpy(h_indices, bx.ptr, nnz * sizeof(int32_t), hipMemcpyDeviceToHost));
    CK_HIP(hipMemcpy(h_data, bd.ptr, nnz * sizeof(float), hipMemcpyDeviceToHost));
  }
  *h_nnz = nnz;
  return CK_OK;
}

static size_t g_kmer_budget = size_t(8) << 30;   // bytes of per-chunk scoring workspace

int ck_kmer_set_workspace_budget(int64_t bytes) {
  if (bytes < (int64_t)(64 << 20)) return fail(CK_ERR_INVALID, "budget below 64 MiB");
  g_kmer_budget = (size_t)bytes;
  return CK_OK;
}

static int64_t km_query_bytes(int32_t num_db, int32_t max_db_len, int32_t max_qlen, int32_t k,
                              int32_t diag_bin_width, int32_t max_cands, int32_t phase_a_topk,
                              bool with_hist) {
  if (num_db < 0 || diag_bin_width <= 0) return -1;
  const int64_t ka = std::max<int64_t>(1, std::min<int64_t>(phase_a_topk, num_db));
  const int64_t nb = std::max<int64_t>(1, (int64_t)(max_db_len - k + max_qlen) / diag_bin_width + 1);
  return (int64_t)num_db * 4 + ka * 4 * 4 + 4 + (with_hist ? ka * nb * 4 : 0) +
         (int64_t)max_cands * 4 + 16;
}

int64_t ck_kmer_query_bytes(int32_t num_db, int32_t max_db_len, int32_t max_qlen, int32_t k,
                            int32_t diag_bin_width, int32_t max_cands, int32_t phase_a_topk) {
  return km_query_bytes(num_db, max_db_len, max_qlen, k, diag_bin_width, max_cands, phase_a_topk,
                        true);
}

// sim: BLOSUM-similar Phase A (k = 5, alphabet 20, threshold sim_T) in place
// of exact k-mer counting; the caller passes min_diag_hits <= 1 with it.
static int kmer_score_impl(const uint8_t* d_qflat, int64_t qflat_len, const int64_t* d_qoff,
                           const int32_t* d_qlen, int32_t nq, int32_t max_qlen, int32_t k,
                           int32_t alpha, const int64_t* d_kmer_offsets,
                           const int32_t* d_post_tid, const int32_t* d_post_pos,
                           const int32_t* d_kmer_freqs, const int64_t* d_tile_off,
                           const int32_t* d_db_lengths, const uint8_t* d_db_flat,
                           int64_t db_flat_len, const int64_t* d_db_offsets,
                           const int16_t* d_db_codes, const int64_t* d_db_code_off,
                           int32_t num_db, int32_t max_db_len, int32_t freq_thresh,
                           int32_t min_total_hits, int32_t min_diag_hits, int32_t diag_bin_width,
                           int32_t max_cands, int32_t phase_a_topk, int32_t* d_out_targets,
                           int32_t* d_out_counts, int32_t* d_out_scores, int32_t* d_out_diags,
                           void* stream, const KmSub400* sim, int32_t sim_T) {
  if (nq < 0 || num_db < 0 || qflat_len < 0) return fail(CK_ERR_INVALID, "negative size");
  if (k < 1 || k > 12 || alpha < 1) return fail(CK_ERR_INVALID, "bad k / alphabet");
  if (diag_bin_width < 1) return fail(CK_ERR_INVALID, "diag_bin_width must be >= 1");
  if (max_cands < 1 || max_cands > 16384)
    return fail(CK_ERR_UNSUPPORTED, "max_cands must be in [1, 16384]");
  if (phase_a_topk < 1) return fail(CK_ERR_INVALID, "phase_a_topk must be >= 1");
  if (nq == 0) return CK_OK;
  if (!d_qflat || !d_qoff || !d_qlen || !d_kmer_offsets || !d_post_tid || !d_post_pos ||
      !d_kmer_freqs || !d_db_lengths || !d_out_targets || !d_out_counts || !d_out_scores ||
      !d_out_diags)
    return fail(CK_ERR_INVALID, "null pointer argument");
  if ((d_db_flat == nullptr) != (d_db_offsets == nullptr) || db_flat_len < 0)
    return fail(CK_ERR_INVALID, "d_db_flat and d_db_offsets go together");
  if ((d_db_codes == nullptr) != (d_db_code_off == nullptr))
    return fail(CK_ERR_INVALID, "d_db_codes and d_db_code_off go together");
  if (d_db_codes && (reinterpret_cast<uintptr_t>(d_db_codes) & 15))
    return fail(CK_ERR_INVALID, "d_db_codes must be 16-byte aligned");
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  StreamWs* w;
  if ((rc = c->ws_for(st, &w))) return rc;
  // Target-major Phase B: database residues on the device, a code space that
  // fits the LDS table, windows that fit two dwords, positions in uint16.
  int64_t ncodes = 1;
  for (int i = 0; i < k && ncodes <= KM_TGT_CODES; ++i) ncodes *= alpha;
  const bool codes = !sim && d_db_codes && k <= 5 && ncodes <= KM_TGT_CODES;
  const bool tgt = !sim && (codes || (d_db_flat && alpha <= 128)) && k <= 5 &&
                   ncodes <= KM_TGT_CODES && max_qlen - k + 1 <= KM_TGT_MAX_KMERS;
  const int64_t per_q = km_query_bytes(num_db, max_db_len, max_qlen, k, diag_bin_width,
                                       max_cands, phase_a_topk, !tgt);
  int32_t qc = (int32_t)std::max<int64_t>(1, std::min<int64_t>(4096, (int64_t)g_kmer_budget / per_q));
  qc = std::min(qc, nq);
  const int32_t ka = std::max(1, std::min(phase_a_topk, std::max(num_db, 1)));
  const int64_t nb = std::max<int64_t>(1, (int64_t)(max_db_len - k + max_qlen) / diag_bin_width + 1);
  const int64_t hcap = tgt ? 1 : (int64_t)ka * nb;
  const int32_t mc = max_cands;
  if ((rc = w->km_qcode.ensure(std::max<int64_t>(qflat_len, 1) * 4)) ||
      (rc = w->km_counts.ensure((size_t)qc * std::max(num_db, 1) * 4)) ||
      (rc = w->km_surv.ensure((size_t)qc * ka * 4)) || (rc = w->km_surv_n.ensure(qc * 4)) ||
      (rc = w->km_surv_tr.ensure(qc * 4)) || (rc = w->km_hoff.ensure((size_t)qc * (ka + 1) * 4)) ||
      (rc = w->km_best.ensure((size_t)qc * ka * 4)) || (rc = w->km_bbin.ensure((size_t)qc * ka * 4)) ||
      (rc = w->km_sel.ensure((size_t)qc * mc * 4)) || (rc = w->km_sel_n.ensure(qc * 4)) ||
      (rc = w->km_sel_tr.ensure(qc * 4)))
    return rc;
  constexpr int SEL_MAX_SLICES = 256;
  const size_t ghist_bytes = (size_t)qc * 32768 * 4;
  if ((rc = w->km_plan.ensure((size_t)qc * sizeof(int4))) ||
      (rc = w->km_bcnt.ensure((size_t)qc * SEL_MAX_SLICES * 2 * 4)))
    return rc;
  if (ghist_bytes > w->km_ghist.bytes || w->km_ghist_zeroed < ghist_bytes) {
    if ((rc = w->km_ghist.ensure(ghist_bytes))) return rc;
    CK_HIP(hipMemsetAsync(w->km_ghist.ptr, 0, w->km_ghist.bytes, st));
    w->km_ghist_zeroed = w->km_ghist.bytes;
  }
  // Exact top-K over n_max entries per query: histogram, cut, per-slice
  // counts, ordered emission (and Phase-B offsets), B slices per query.
  auto select = [&](SelectArgs sa, int32_t nqc, int64_t n_max) {
    const int B = (int)std::max<int64_t>(1, std::min<int64_t>(SEL_MAX_SLICES, (n_max + 8191) / 8192));
    sa.ghist = w->km_ghist.as<int32_t>();
    sa.plan = w->km_plan.as<int4>();
    sa.bcnt = w->km_bcnt.as<int32_t>();
    hipLaunchKernelGGL(km_sel_hist_kernel, dim3(B, nqc), dim3(256), 0, st, sa);
    hipLaunchKernelGGL(km_sel_cut_kernel, dim3(nqc), dim3(1024), 0, st, sa);
    hipLaunchKernelGGL(km_sel_count_kernel, dim3(B, nqc), dim3(256), 0, st, sa);
    hipLaunchKernelGGL(km_sel_emit_kernel, dim3(B, nqc), dim3(256), 0, st, sa);
    if (sa.out_hoff) hipLaunchKernelGGL(km_sel_hoff_kernel, dim3(nqc), dim3(1024), 0, st, sa);
  };
  const size_t hist_bytes = (size_t)qc * hcap * 4;
  if (hist_bytes > w->km_hist.bytes || w->km_hist_zeroed < hist_bytes) {
    if ((rc = w->km_hist.ensure(hist_bytes))) return rc;
    CK_HIP(hipMemsetAsync(w->km_hist.ptr, 0, w->km_hist.bytes, st));
    w->km_hist_zeroed = w->km_hist.bytes;
  }
  const size_t fin_lds = [&] { size_t p = 1; while ((int64_t)p < mc) p <<= 1; return p * 8; }();
  if (fin_lds > 64 * 1024)
    CK_HIP(hipFuncSetAttribute((const void*)km_finalize_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)fin_lds));
  const int32_t max_nk = std::max(0, max_qlen - k + 1);
  // table, query positions, 16 wavefronts' bins (rounded to 8 bytes), then
  // the survivor metadata: offsets, lengths, cost prefixes (256 each) and
  // 33 scan words
  const size_t tgt_lds =
      ((((size_t)ncodes + 2 + (size_t)(max_nk + 1) / 2 + 16 * (size_t)KM_TGT_BINS_HOST) & ~(size_t)1) +
       2 * 256 + 256 + 256 + 33) * 4;
  using TgtFn = void (*)(const int32_t*, const int64_t*, const int32_t*, int32_t, int32_t,
                        const uint8_t*, int64_t, const int16_t*, const int64_t*, const int32_t*,
                        const int32_t*, const int32_t*, int32_t, int32_t, int32_t, int32_t,
                        uint32_t*, int32_t*);
  const TgtFn tgt_fns[2][5] = {
      {km_phase_b_target_kernel<1, false, false, false>, km_phase_b_target_kernel<2, false, false, false>,
       km_phase_b_target_kernel<3, false, false, false>, km_phase_b_target_kernel<4, false, false, false>,
       km_phase_b_target_kernel<5, false, false, false>},
      {km_phase_b_target_kernel<1, true, false, false>, km_phase_b_target_kernel<2, true, false, false>,
       km_phase_b_target_kernel<3, true, false, false>, km_phase_b_target_kernel<4, true, false, false>,
       km_phase_b_target_kernel<5, true, false, false>}};
  const TgtFn tgt_dot[2] = {km_phase_b_target_kernel<3, false, true, false>,
                            km_phase_b_target_kernel<3, true, true, false>};
  const TgtFn tgt_codes[2][5] = {
      {km_phase_b_target_kernel<1, false, false, true>, km_phase_b_target_kernel<2, false, false, true>,
       km_phase_b_target_kernel<3, false, false, true>, km_phase_b_target_kernel<4, false, false, true>,
       km_phase_b_target_kernel<5, false, false, true>},
      {km_phase_b_target_kernel<1, true, false, true>, km_phase_b_target_kernel<2, true, false, true>,
       km_phase_b_target_kernel<3, true, false, true>, km_phase_b_target_kernel<4, true, false, true>,
       km_phase_b_target_kernel<5, true, false, true>}};
  // Diagonals n = tpos - qpos + qlen stay below max_db_len + max_qlen: the
  // 24-bit multiply-high bin division is exact below 2^17 for bin widths 3 to
  // 256.  k = 3
  // codes come from byte dot products when alpha^2 <= 510.  With the
  // database's window codes on the device, Phase B reads those instead.
  const bool fdiv = (int64_t)max_db_len + max_qlen <= (1 << 17) && diag_bin_width >= 3 &&
                    diag_bin_width <= 256;
  const bool dot3 = k == 3 && alpha * alpha <= 510;
  const TgtFn tgt_fn = !tgt  ? nullptr
                       : codes ? tgt_codes[fdiv ? 1 : 0][k - 1]
                       : dot3  ? tgt_dot[fdiv ? 1 : 0]
                               : tgt_fns[fdiv ? 1 : 0][k - 1];
  if (tgt && tgt_lds > 64 * 1024)
    CK_HIP(hipFuncSetAttribute((const void*)tgt_fn, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)tgt_lds));
  if (tgt && codes) {
    // the window-code kernels address their code table as LDS address 0,
    // which holds only while the kernel has no static LDS
    hipFuncAttributes fa;
    CK_HIP(hipFuncGetAttributes(&fa, (const void*)tgt_fn));
    if (fa.sharedSizeBytes != 0)
      return fail(CK_ERR_UNSUPPORTED, "Phase B window-code kernel has static LDS");
  }
  const int32_t ntiles = (num_db + KM_TILE_HOST - 1) / KM_TILE_HOST;
  int32_t* qcode = w->km_qcode.as<int32_t>();
  uint32_t* counts = w->km_counts.as<uint32_t>();

  if (num_db == 0) {
    hipLaunchKernelGGL(km_fill_kernel, dim3((nq + 255) / 256), dim3(256), 0, st,
                       (uint32_t*)d_out_counts, (int64_t)nq, 0u);
    CK_HIP(hipGetLastError());
    return CK_OK;
  }
  if (!sim)
    hipLaunchKernelGGL(km_qcodes_kernel,
                       dim3(std::max(1, std::min(64, (max_qlen + 255) / 256)),
                            std::min(nq, 65535)),
                       dim3(256), 0, st, d_qflat, d_qoff, d_qlen, nq, k, alpha, d_kmer_freqs,
                       freq_thresh, qcode);
  for (int32_t q0 = 0; q0 < nq; q0 += qc) {
    const int32_t n = std::min(qc, nq - q0);
    const int64_t* qoff = d_qoff + q0;
    const int32_t* qlen = d_qlen + q0;
    // Small batches: split each query's k-mers over up to 16 workgroups per
    // tile so the grid still fills the device.
    const int64_t wg = (int64_t)ntiles * n;
    const int zs = (int)std::max<int64_t>(
        1, std::min<int64_t>({16, (4 * (int64_t)c->num_cu + wg - 1) / wg,
                              (int64_t)(max_qlen - k + 1 + 31) / 32}));
    if (sim) {
      constexpr int32_t PPB = 4;   // query positions per workgroup
      hipLaunchKernelGGL(km_fill_kernel, dim3(1024), dim3(256), 0, st, counts,
                         (int64_t)n * num_db, 0u);
      hipLaunchKernelGGL(km_similar_kernel, dim3(std::max(1, (max_qlen - 4 + PPB - 1) / PPB), n),
                         dim3(256), 0, st, d_qflat, qoff, qlen, d_kmer_offsets, d_post_tid,
                         d_kmer_freqs, num_db, *sim, sim_T, freq_thresh, PPB, counts);
    } else if (d_tile_off) {
      if (zs > 1)
        hipLaunchKernelGGL(km_fill_kernel, dim3(1024), dim3(256), 0, st, counts,
                           (int64_t)n * num_db, 0u);
      hipLaunchKernelGGL(km_phase_a_tiled_kernel, dim3(ntiles, n, zs), dim3(256), 0, st, qcode,
                         qoff, qlen, k, d_tile_off, ntiles, d_post_tid, num_db, counts);
    } else {
      hipLaunchKernelGGL(km_fill_kernel, dim3(1024), dim3(256), 0, st, counts,
                         (int64_t)n * num_db, 0u);
      hipLaunchKernelGGL(km_phase_a_global_kernel,
                         dim3(std::max(1, std::min(64, (max_qlen + 3) / 4)), n), dim3(256), 0,
                         st, qcode, qoff, qlen, k, d_kmer_offsets, d_post_tid, num_db, counts);
    }
    SelectArgs sa{};
    sa.scores = counts;
    sa.stride = num_db;
    sa.n_of_q = nullptr;
    sa.n_fixed = num_db;
    sa.min_score = min_total_hits;
    sa.qlen = qlen;
    sa.k = k;
    sa.w = diag_bin_width;
    int32_t* out_t = d_out_targets + (int64_t)q0 * mc;
    int32_t* out_s = d_out_scores + (int64_t)q0 * mc;
    int32_t* out_d = d_out_diags + (int64_t)q0 * mc;
    if (min_diag_hits <= 1) {
      sa.K = mc;
      sa.out_ids = w->km_sel.as<int32_t>();
      sa.out_n = w->km_sel_n.as<int32_t>();
      sa.out_trunc = w->km_sel_tr.as<int32_t>();
      sa.mark = 0;
      sa.out_hoff = nullptr;
      sa.id_len = nullptr;
      select(sa, n, num_db);
      hipLaunchKernelGGL(km_finalize_kernel, dim3(n), dim3(1024), fin_lds, st,
                         w->km_sel.as<int32_t>(), w->km_sel_n.as<int32_t>(),
                         w->km_sel_tr.as<int32_t>(), mc, counts, (int64_t)num_db,
                         (const int32_t*)nullptr, 0, (const int32_t*)nullptr, qlen,
                         diag_bin_width, mc, out_t, d_out_counts + q0, out_s, out_d);
      continue;
    }
    sa.K = ka;
    sa.out_ids = w->km_surv.as<int32_t>();
    sa.out_n = w->km_surv_n.as<int32_t>();
    sa.out_trunc = w->km_surv_tr.as<int32_t>();
    sa.mark = 1;
    sa.out_hoff = tgt ? nullptr : w->km_hoff.as<int32_t>();
    sa.id_len = d_db_lengths;
    select(sa, n, num_db);
    if (tgt) {
      // survivors per workgroup: at most 256, fewer when the batch is small
      // so that the grid still spans every CU twice
      const int64_t want = ((int64_t)ka * n + 2 * c->num_cu - 1) / (2 * c->num_cu);
      const int32_t per_wg = (int32_t)std::max<int64_t>(16, std::min<int64_t>(256, (want + 15) / 16 * 16));
      hipLaunchKernelGGL(tgt_fn, dim3((ka + per_wg - 1) / per_wg, n), dim3(1024), tgt_lds, st,
                         qcode, qoff, qlen, alpha, (int32_t)ncodes, d_db_flat, db_flat_len,
                         codes ? d_db_codes : (const int16_t*)nullptr,
                         codes ? d_db_code_off : d_db_offsets, d_db_lengths,
                         w->km_surv.as<int32_t>(),
                         w->km_surv_n.as<int32_t>(), ka, diag_bin_width, per_wg, max_nk,
                         w->km_best.as<uint32_t>(), w->km_bbin.as<int32_t>());
    } else if (d_tile_off && ka < 0xFFFF) {
      hipLaunchKernelGGL(km_phase_b_tiled_kernel<true>, dim3(ntiles, n, zs), dim3(256), 0, st, qcode,
                         qoff, qlen, k, d_tile_off, ntiles, d_post_tid, d_post_pos, num_db,
                         counts, w->km_hoff.as<int32_t>(), ka, diag_bin_width,
                         w->km_hist.as<int32_t>(), hcap);
    } else if (d_tile_off) {
      hipLaunchKernelGGL(km_phase_b_tiled_kernel<false>, dim3(ntiles, n, zs), dim3(256), 0, st,
                         qcode, qoff, qlen, k, d_tile_off, ntiles, d_post_tid, d_post_pos,
                         num_db, counts, w->km_hoff.as<int32_t>(), ka, diag_bin_width,
                         w->km_hist.as<int32_t>(), hcap);
    } else {
      hipLaunchKernelGGL(km_phase_b_global_kernel,
                         dim3(std::max(1, std::min(64, (max_qlen + 3) / 4)), n), dim3(256), 0,
                         st, qcode, qoff, qlen, k, d_kmer_offsets, d_post_tid, d_post_pos,
                         num_db, counts, w->km_hoff.as<int32_t>(), ka, diag_bin_width,
                         w->km_hist.as<int32_t>(), hcap);
    }
    if (!tgt)
      hipLaunchKernelGGL(km_phase_b_best_kernel, dim3((ka + 15) / 16, n), dim3(256), 0, st,
                         w->km_surv_n.as<int32_t>(), w->km_hoff.as<int32_t>(), ka,
                         w->km_hist.as<int32_t>(), hcap, w->km_best.as<uint32_t>(),
                         w->km_bbin.as<int32_t>());
    SelectArgs sb{};
    sb.scores = w->km_best.as<uint32_t>();
    sb.stride = ka;
    sb.n_of_q = w->km_surv_n.as<int32_t>();
    sb.n_fixed = 0;
    sb.min_score = std::max(min_diag_hits, 1);
    sb.K = mc;
    sb.out_ids = w->km_sel.as<int32_t>();
    sb.out_n = w->km_sel_n.as<int32_t>();
    sb.out_trunc = w->km_sel_tr.as<int32_t>();
    sb.mark = 0;
    sb.out_hoff = nullptr;
    sb.id_len = nullptr;
    sb.qlen = nullptr;
    sb.k = k;
    sb.w = diag_bin_width;
    select(sb, n, ka);
    hipLaunchKernelGGL(km_finalize_kernel, dim3(n), dim3(1024), fin_lds, st,
                       w->km_sel.as<int32_t>(), w->km_sel_n.as<int32_t>(),
                       w->km_sel_tr.as<int32_t>(), mc, w->km_best.as<uint32_t>(), (int64_t)ka,
                       w->km_surv.as<int32_t>(), ka, w->km_bbin.as<int32_t>(), qlen,
                       diag_bin_width, mc, out_t, d_out_counts + q0, out_s, out_d);
  }
  CK_HIP(hipGetLastError());
  return CK_OK;
}


int ck_kmer_score_device(const uint8_t* d_qflat, int64_t qflat_len, const int64_t* d_qoff,
                         const int32_t* d_qlen, int32_t nq, int32_t max_qlen, int32_t k,
                         int32_t alpha, const int64_t* d_kmer_offsets,
                         const int32_t* d_post_tid, const int32_t* d_post_pos,
                         const int32_t* d_kmer_freqs, const int64_t* d_tile_off,
                         const int32_t* d_db_lengths, const uint8_t* d_db_flat,
                         int64_t db_flat_len, const int64_t* d_db_offsets,
                         const int16_t* d_db_codes, const int64_t* d_db_code_off, int32_t num_db,
                         int32_t max_db_len, int32_t freq_thresh, int32_t min_total_hits,
                         int32_t min_diag_hits, int32_t diag_bin_width, int32_t max_cands,
                         int32_t phase_a_topk, int32_t* d_out_targets, int32_t* d_out_counts,
                         int32_t* d_out_scores, int32_t* d_out_diags, void* stream) {
  return kmer_score_impl(d_qflat, qflat_len, d_qoff, d_qlen, nq, max_qlen, k, alpha,
                         d_kmer_offsets, d_post_tid, d_post_pos, d_kmer_freqs, d_tile_off,
                         d_db_lengths, d_db_flat, db_flat_len, d_db_offsets, d_db_codes,
                         d_db_code_off, num_db, max_db_len, freq_thresh, min_total_hits,
                         min_diag_hits, diag_bin_width, max_cands, phase_a_topk, d_out_targets,
                         d_out_counts, d_out_scores, d_out_diags, stream, nullptr, 0);
}

int ck_kmer_score_similar_device(const uint8_t* d_qflat, int64_t qflat_len,
                                 const int64_t* d_qoff, const int32_t* d_qlen, int32_t nq,
                                 int32_t max_qlen, const int64_t* d_kmer_offsets,
                                 const int32_t* d_post_tid, const int32_t* d_kmer_freqs,
                                 const int32_t* d_db_lengths, int32_t num_db,
                                 int32_t max_db_len, const int8_t* h_sub_matrix,
                                 int32_t kmer_score_thresh, int32_t freq_thresh,
                                 int32_t min_total_hits, int32_t max_cands,
                                 int32_t* d_out_targets, int32_t* d_out_counts,
                                 int32_t* d_out_scores, int32_t* d_out_diags, void* stream) {
  if (!h_sub_matrix) return fail(CK_ERR_INVALID, "similar k-mers: null substitution table");
  if (kmer_score_thresh < 1)
    return fail(CK_ERR_INVALID, "similar k-mers: kmer_score_thresh must be >= 1");
  KmSub400 sub;
  std::memcpy(sub.v, h_sub_matrix, 400);
  return kmer_score_impl(d_qflat, qflat_len, d_qoff, d_qlen, nq, max_qlen, 5, 20,
                         d_kmer_offsets, d_post_tid, d_post_tid, d_kmer_freqs, nullptr,
                         d_db_lengths, nullptr, 0, nullptr, nullptr, nullptr, num_db,
                         max_db_len, freq_thresh, min_total_hits, 1, 1 << 20, max_cands, max_cands,
                         d_out_targets, d_out_counts, d_out_scores, d_out_diags, stream, &sub,
                         kmer_score_thresh);
}

int ck_search_pairs_device(const int32_t* d_targets, const int32_t* d_counts, int32_t nq,
                           int32_t max_cands, int32_t q_index_base, int32_t empty_index,
                           int32_t* d_pairs, void* stream) {
  if (nq < 0 || max_cands < 1) return fail(CK_ERR_INVALID, "bad sizes");
  if (nq == 0) return CK_OK;
  if (!d_targets || !d_counts || !d_pairs) return fail(CK_ERR_INVALID, "null pointer argument");
  const int64_t n = (int64_t)nq * max_cands;
  hipLaunchKernelGGL(ck_search_pairs_kernel, dim3((unsigned)std::min<int64_t>((n + 255) / 256, 4096)),
                     dim3(256), 0, (hipStream_t)stream, d_targets, d_counts, nq, max_cands,
                     q_index_base, empty_index, d_pairs);
  CK_HIP(hipGetLastError());
  return CK_OK;
}

int ck_kmer_tile_offsets_host(const int64_t* kmer_offsets, const int32_t* post_tid, int64_t n_codes,
                              int32_t num_db, int64_t* out) {
  if (n_codes < 0 || num_db < 0) return fail(CK_ERR_INVALID, "negative size");
  const int32_t nt = (num_db + KM_TILE_HOST - 1) / KM_TILE_HOST;
  for (int64_t cidx = 0; cidx < n_codes; ++cidx) {
    int64_t h = kmer_offsets[cidx];
    const int64_t e = kmer_offsets[cidx + 1];
    for (int32_t j = 0; j <= nt; ++j) {
      const int64_t bound = (int64_t)j * KM_TILE_HOST;
      while (h < e && post_tid[h] < bound) ++h;
      out[cidx * (nt + 1) + j] = (j == nt) ? e : h;
    }
  }
  return CK_OK;
}

int ck_kmer_index_build_device(const uint8_t* d_flat, const int64_t* d_offsets,
                               const int32_t* d_lengths, int32_t n_seqs, int32_t k, int32_t alpha,
                               int64_t cap, int64_t* d_kmer_offsets, int32_t* d_kmer_freqs,
                               int32_t* d_post_tid, int32_t* d_post_pos, int64_t* out_total,
                               void* stream) {
  if (n_seqs < 0 || k < 1 || alpha < 2 || cap < 0 || !out_total)
    return fail(CK_ERR_INVALID, "k-mer index build: bad sizes (n=%d k=%d alpha=%d cap=%lld)",
                n_seqs, k, alpha, (long long)cap);
  int64_t ncodes = 1;
  for (int j = 0; j < k; ++j) {
    ncodes *= alpha;
    if (ncodes > (int64_t(1) << 31))
      return fail(CK_ERR_UNSUPPORTED, "alpha^k = %d^%d codes exceed 2^31", alpha, k);
  }
  if ((n_seqs > 0 && (!d_flat || !d_offsets || !d_lengths)) || !d_kmer_offsets ||
      !d_kmer_freqs || (cap > 0 && (!d_post_tid || !d_post_pos)))
    return fail(CK_ERR_INVALID, "k-mer index build: null pointer");
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  int status = 0;
  hipError_t e = ck::kmer_index_build(d_flat, d_offsets, d_lengths, n_seqs, k, alpha, cap,
                                      d_kmer_offsets, d_kmer_freqs, d_post_tid, d_post_pos,
                                      out_total, &status, (hipStream_t)stream);
  if (e == hipErrorOutOfMemory)
    return fail(CK_ERR_NOMEM, "k-mer index build: device allocation failed");
  if (e != hipSuccess)
    return fail(CK_ERR_HIP, "k-mer index build: %s", hipGetErrorString(e));
  if (status == 1)
    return fail(CK_ERR_INVALID, "k-mer index build: capacity %lld < %lld windows",
                (long long)cap, (long long)*out_total);
  return CK_OK;
}

int ck_kmer_db_codes_device(const uint8_t* d_flat, const int64_t* d_offsets,
                            const int32_t* d_lengths, int32_t n_seqs, int32_t k, int32_t alpha,
                            const int64_t* d_code_off, int16_t* d_codes, void* stream) {
  if (n_seqs < 0 || k < 1 || alpha < 1) return fail(CK_ERR_INVALID, "db codes: bad sizes");
  int64_t ncodes = 1;
  for (int i = 0; i < k && ncodes <= KM_TGT_CODES; ++i) ncodes *= alpha;
  if (k > 5 || ncodes > KM_TGT_CODES)
    return fail(CK_ERR_UNSUPPORTED, "db codes: alpha^k above %d or k above 5", KM_TGT_CODES);
  if (n_seqs == 0) return CK_OK;
  if (!d_flat || !d_offsets || !d_lengths || !d_code_off || !d_codes)
    return fail(CK_ERR_INVALID, "db codes: null pointer");
  if (reinterpret_cast<uintptr_t>(d_codes) & 15)
    return fail(CK_ERR_INVALID, "db codes: d_codes must be 16-byte aligned");
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  CK_HIP(ck::kmer_db_codes(d_flat, d_offsets, d_lengths, n_seqs, k, alpha, d_code_off, d_codes,
                           (hipStream_t)stream));
  return CK_OK;
}

int ck_kmer_tile_offsets_device(const int64_t* d_kmer_offsets, const int32_t* d_post_tid,
                                int64_t n_codes, int32_t num_db, int64_t* d_out, void* stream) {
  if (n_codes < 0 || num_db < 0 || !d_kmer_offsets || !d_out)
    return fail(CK_ERR_INVALID, "tile offsets: bad arguments");
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  CK_HIP(ck::kmer_tile_offsets(d_kmer_offsets, d_post_tid, n_codes, num_db, d_out,
                               (hipStream_t)stream));
  return CK_OK;
}

int ck_sketch_device(const uint8_t* d_flat, const int64_t* d_offsets, const int32_t* d_lengths,
                     int64_t n_seqs, int32_t k, int32_t sketch_size, int32_t alphabet_size,
                     uint64_t seed, uint64_t* d_out_sketches, void* stream) {
  if (n_seqs < 0 || k < 1 || alphabet_size < 1 || sketch_size < 1)
    return fail(CK_ERR_INVALID, "sketch: bad sizes (n=%lld k=%d s=%d alpha=%d)",
                (long long)n_seqs, k, sketch_size, alphabet_size);
  if (sketch_size > ck::sketch_max_size())
    return fail(CK_ERR_UNSUPPORTED, "sketch size %d above %d", sketch_size,
                ck::sketch_max_size());
  if (n_seqs == 0) return CK_OK;
  if (!d_flat || !d_offsets || !d_lengths || !d_out_sketches)
    return fail(CK_ERR_INVALID, "sketch: null pointer");
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  CK_HIP(ck::sketch_build(d_flat, d_offsets, d_lengths, n_seqs, k, sketch_size, alphabet_size,
                          seed, d_out_sketches, (hipStream_t)stream));
  return CK_OK;
}

int ck_lsh_candidates_device(const uint64_t* d_sketches, int64_t n_seqs, int32_t sketch_size,
                             const int32_t* h_band_indices, const int64_t* h_seeds,
                             int32_t num_tables, int32_t num_bands, int32_t max_bucket,
                             int64_t* out_n_pairs, void* stream) {
  if (!out_n_pairs) return fail(CK_ERR_INVALID, "lsh: null count pointer");
  *out_n_pairs = 0;
  if (n_seqs < 0 || n_seqs > INT32_MAX || sketch_size < 1 || num_tables < 0 || num_bands < 1 ||
      max_bucket < 1)
    return fail(CK_ERR_INVALID, "lsh: bad sizes (n=%lld s=%d tables=%d bands=%d cap=%d)",
                (long long)n_seqs, sketch_size, num_tables, num_bands, max_bucket);
  if (num_tables > 0 && (!h_band_indices || !h_seeds))
    return fail(CK_ERR_INVALID, "lsh: null band or seed array");
  for (int64_t x = 0; x < (int64_t)num_tables * num_bands; ++x)
    if (h_band_indices[x] < 0 || h_band_indices[x] >= sketch_size)
      return fail(CK_ERR_INVALID, "lsh: band index %d outside the sketch", h_band_indices[x]);
  if (n_seqs >= 2 && num_tables > 0 && !d_sketches)
    return fail(CK_ERR_INVALID, "lsh: null sketches");
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  StreamWs* w;
  if ((rc = c->ws_for(st, &w))) return rc;
  w->lsh_m = 0;
  w->lsh_n = n_seqs;
  if (n_seqs < 2 || num_tables == 0) return CK_OK;
  if (!w->lsh_ws) w->lsh_ws = ck::lsh_ws_new();
  int64_t m = 0;
  // Tables are processed in groups of at most group_elems (table, sequence)
  // entries (~48 bytes of workspace each); CK_LSH_GROUP_ELEMS overrides.
  int64_t group_elems = int64_t(1) << 25;
  if (const char* env = std::getenv("CK_LSH_GROUP_ELEMS")) {
    const long long v = std::atoll(env);
    if (v > 0) group_elems = v;
  }
  const hipError_t e = ck::lsh_build(w->lsh_ws, d_sketches, n_seqs, sketch_size, h_band_indices,
                                     h_seeds, num_tables, num_bands, max_bucket, group_elems, &m,
                                     st);
  if (e == hipErrorOutOfMemory) return fail(CK_ERR_NOMEM, "lsh: device allocation failed");
  if (e != hipSuccess) return fail(CK_ERR_HIP, "lsh: %s", hipGetErrorString(e));
  w->lsh_m = m;
  *out_n_pairs = m;
  return CK_OK;
}

int ck_lsh_fetch_pairs_device(int32_t* d_out_pairs, int64_t n_pairs, void* stream) {
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  std::lock_guard<std::mutex> lk(c->mu);
  hipStream_t st = (hipStream_t)stream;
  StreamWs* w;
  if ((rc = c->ws_for(st, &w))) return rc;
  if (n_pairs != w->lsh_m)
    return fail(CK_ERR_INVALID, "lsh fetch: %lld pairs asked, the last call on this stream "
                "found %lld", (long long)n_pairs, (long long)w->lsh_m);
  if (n_pairs == 0) return CK_OK;
  if (!d_out_pairs) return fail(CK_ERR_INVALID, "lsh fetch: null output");
  CK_HIP(ck::lsh_unpack(w->lsh_ws, w->lsh_n, d_out_pairs, st));
  return CK_OK;
}

int ck_search_topk_device(const int32_t* d_targets, const int32_t* d_counts,
                          const float* d_sims, const int32_t* d_scores, const uint8_t* d_mask,
                          const int32_t* d_qlens, const int32_t* d_tlens, const float* d_logtab,
                          int64_t logtab_len, int32_t nq, int32_t max_cands, int32_t top_k,
                          int32_t rank_mode, int32_t* d_out_target, float* d_out_identity,
                          float* d_out_rank, int32_t* d_out_n, int32_t* d_status, void* stream) {
  if (rank_mode < CK_RANK_IDENTITY || rank_mode > CK_RANK_SCORE)
    return fail(CK_ERR_INVALID, "top-k: rank mode %d", rank_mode);
  if (nq < 0 || max_cands < 1 || top_k < 1 || logtab_len < 0)
    return fail(CK_ERR_INVALID, "top-k: bad sizes (nq=%d max_cands=%d top_k=%d)", nq,
                max_cands, top_k);
  if (max_cands > ck::topk_max_candidates())
    return fail(CK_ERR_UNSUPPORTED, "top-k: %d candidates per query above %d", max_cands,
                ck::topk_max_candidates());
  if (nq == 0) return CK_OK;
  if (!d_targets || !d_counts || !d_sims || !d_mask || !d_out_target || !d_out_identity ||
      !d_out_rank || !d_out_n || !d_status)
    return fail(CK_ERR_INVALID, "top-k: null pointer");
  if (rank_mode == CK_RANK_EVALUE &&
      (!d_scores || !d_qlens || !d_tlens || !d_logtab || logtab_len < 2))
    return fail(CK_ERR_INVALID, "top-k: rank score needs scores, lengths and the log table");
  if (rank_mode == CK_RANK_SCORE && !d_scores)
    return fail(CK_ERR_INVALID, "top-k: rank by score needs scores");
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  CK_HIP(ck::search_topk(d_targets, d_counts, d_sims, d_scores, d_mask, d_qlens, d_tlens,
                         d_logtab, logtab_len, nq, max_cands, top_k, rank_mode,
                         d_out_target, d_out_identity, d_out_rank, d_out_n, d_status,
                         (hipStream_t)stream));
  return CK_OK;
}

int ck_batch_score_queries(const uint8_t* q_flat, const int64_t* q_offsets, const int32_t* q_lengths,
                           int32_t nq, int32_t k, int32_t alpha_size, const int64_t* kmer_offsets,
                           const int64_t* kmer_entries, const int32_t* compact_entries,
                           const int32_t* kmer_freqs, int32_t freq_thresh, int32_t num_db,
                           int32_t min_total_hits, int32_t min_diag_hits, int32_t diag_bin_width,
                           int32_t max_cands, int32_t phase_a_topk, int32_t* out_targets,
                           int32_t* out_counts, int32_t* out_scores, int32_t* out_diags) {
  (void)compact_entries;   // the target ids are the high halves of kmer_entries
  if (nq < 0 || num_db < 0 || k < 1 || k > 12 || alpha_size < 1)
    return fail(CK_ERR_INVALID, "bad sizes");
  if (nq == 0) return CK_OK;
  if (!q_offsets || !q_lengths || !kmer_offsets || !kmer_freqs || !out_targets || !out_counts)
    return fail(CK_ERR_INVALID, "null pointer argument");
  int64_t ncodes = 1;
  for (int i = 0; i < k; ++i) ncodes *= alpha_size;
  const int64_t E = kmer_offsets[ncodes];
  if (E > 0 && !kmer_entries) return fail(CK_ERR_INVALID, "null kmer_entries");
  int64_t qflat_len = 1;
  int32_t max_qlen = 0;
  for (int32_t q = 0; q < nq; ++q) {
    if (q_lengths[q] < 0) return fail(CK_ERR_INVALID, "negative query length");
    qflat_len = std::max<int64_t>(qflat_len, q_offsets[q] + q_lengths[q]);
    max_qlen = std::max(max_qlen, q_lengths[q]);
  }
  std::vector<int32_t> tid(std::max<int64_t>(E, 1)), pos(std::max<int64_t>(E, 1));
  std::vector<int32_t> lenb(std::max(num_db, 1), 0);
  int32_t max_len = 0;
  for (int64_t h = 0; h < E; ++h) {
    const int32_t t = (int32_t)(kmer_entries[h] >> 32);
    const int32_t p = (int32_t)(kmer_entries[h] & 0xFFFFFFFF);
    if (t < 0 || t >= num_db) return fail(CK_ERR_INVALID, "posting target out of range");
    tid[h] = t;
    pos[h] = p;
    lenb[t] = std::max(lenb[t], p + k);
    max_len = std::max(max_len, p + k);
  }
  const int32_t nt = (num_db + KM_TILE_HOST - 1) / KM_TILE_HOST;
  const bool tiled = num_db > 0 && ncodes * (int64_t)(nt + 1) <= (int64_t)32 << 20 &&
                     E >= ncodes * 64;   // dense postings: walk tiles in LDS
  std::vector<int64_t> toff;
  if (tiled) {
    toff.resize(ncodes * (nt + 1));
    int rc0 = ck_kmer_tile_offsets_host(kmer_offsets, tid.data(), ncodes, num_db, toff.data());
    if (rc0) return rc0;
  }
  DevCtx* c;
  int rc = current_ctx(&c);
  if (rc) return rc;
  std::lock_guard<std::mutex> host_lk(c->host_mu);
  {
    std::lock_guard<std::mutex> lk(c->mu);
    if ((rc = c->kh_qflat.ensure(qflat_len)) || (rc = c->kh_qoff.ensure(nq * 8)) ||
        (rc = c->kh_qlen.ensure(nq * 4)) || (rc = c->kh_koff.ensure((ncodes + 1) * 8)) ||
        (rc = c->kh_tid.ensure(tid.size() * 4)) || (rc = c->kh_pos.ensure(pos.size() * 4)) ||
        (rc = c->kh_freq.ensure(ncodes * 4)) || (rc = c->kh_len.ensure(lenb.size() * 4)) ||
        (rc = c->kh_tg.ensure((size_t)nq * max_cands * 4)) || (rc = c->kh_cn.ensure(nq * 4)) ||
        (rc = c->kh_sc.ensure((size_t)nq * max_cands * 4)) ||
        (rc = c->kh_dg.ensure((size_t)nq * max_cands * 4)))
      return rc;
    if (tiled && (rc = c->kh_tile.ensure(toff.size() * 8))) return rc;
    if (q_flat) CK_HIP(hipMemcpy(c->kh_qflat.ptr, q_flat, qflat_len, hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(c->kh_qoff.ptr, q_offsets, nq * 8, hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(c->kh_qlen.ptr, q_lengths, nq * 4, hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(c->kh_koff.ptr, kmer_offsets, (ncodes + 1) * 8, hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(c->kh_tid.ptr, tid.data(), tid.size() * 4, hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(c->kh_pos.ptr, pos.data(), pos.size() * 4, hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(c->kh_freq.ptr, kmer_freqs, ncodes * 4, hipMemcpyHostToDevice));
    CK_HIP(hipMemcpy(c->kh_len.ptr, lenb.data(), lenb.size() * 4, hipMemcpyHostToDevice));
    if (tiled) CK_HIP(hipMemcpy(c->kh_tile.ptr, toff.data(), toff.size() * 8, hipMemcpyHostToDevice));
  }
  rc = ck_kmer_score_device(
      c->kh_qflat.as<uint8_t>(), qflat_len, c->kh_qoff.as<int64_t>(), c->kh_qlen.as<int32_t>(),
      nq, max_qlen, k, alpha_size, c->kh_koff.as<int64_t>(), c->kh_tid.as<int32_t>(),
      c->kh_pos.as<int32_t>(), c->kh_freq.as<int32_t>(), tiled ? c->kh_tile.as<int64_t>() : nullptr,
      c->kh_len.as<int32_t>(), nullptr, 0, nullptr, nullptr, nullptr, num_db, max_len, freq_thresh,
      min_total_hits, min_diag_hits,
      diag_bin_width, max_cands, phase_a_topk, c->kh_tg.as<int32_t>(), c->kh_cn.as<int32_t>(),
      c->kh_sc.as<int32_t>(), c->kh_dg.as<int32_t>(), nullptr);
  if (rc) return rc;
  CK_HIP(hipDeviceSynchronize());
  std::lock_guard<std::mutex> lk(c->mu);
  CK_HIP(hipMemcpy(out_targets, c->kh_tg.ptr, (size_t)nq * max_cands * 4, hipMemcpyDeviceToHost));
  CK_HIP(hipMemcpy(out_counts, c->kh_cn.ptr, nq * 4, hipMemcpyDeviceToHost));
  if (out_scores)
    CK_HIP(hipMemcpy(out_scores, c->kh_sc.ptr, (size_t)nq * max_cands * 4, hipMemcpyDeviceToHost));
  if (out_diags)
    CK_HIP(hipMemcpy(out_diags, c->kh_dg.ptr, (size_t)nq * max_cands * 4, hipMemcpyDeviceToHost));
  return CK_OK;
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
  std::lock_guard<std::mutex> host_lk(c->host_mu);
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
