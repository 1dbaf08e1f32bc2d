// Similarity graph in CSR form on the device (clustkit/graph.py:7-39,
// build_similarity_graph).
//
// The reference stacks (i, j, w) and (j, i, w) for every passing pair into a
// COO matrix and converts it with scipy's tocsr(): rows in order, column
// indices sorted within each row, duplicate (row, column) entries summed,
// explicit zeros kept.  Here: the 2K entries get a key row << b | column
// (b = bits of N), a stable radix sort orders them, a head flag and an
// exclusive scan place each distinct key, its run is summed in entry order
// (first half of the stack before the second, each in pair order), and the
// row pointer of row r is the lower bound of r among the distinct rows.
//
// Duplicates: a pair listed twice, in either orientation, or a self-pair
// gives two entries per (row, column), and two float32 terms sum the same in
// any order.  Three or more copies of one entry are summed here in entry
// order; scipy sums them in the order of an unstable sort, so those sums may
// differ in the last bit.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>

#include "ck_internal.h"

namespace ck {

namespace {

unsigned gr_grid(int64_t work) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, 65536));
}

struct GBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t need) {
    if (need <= bytes) return hipSuccess;
    const size_t want = std::max<size_t>(need + need / 4, 256);
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, want);
    if (e != hipSuccess) return e;
    if (p) (void)hipFree(p);
    p = q;
    bytes = want;
    return hipSuccess;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Entry e < K is (pairs[e].i, pairs[e].j); entry K + e is (j, i).
__global__ void gr_expand_kernel(const int32_t* __restrict__ pairs, int64_t K,
                                 const float* __restrict__ sims, int32_t b,
                                 uint64_t* __restrict__ keys, float* __restrict__ vals) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < 2 * K;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t s = e < K ? e : e - K;
    const uint32_t x = (uint32_t)pairs[2 * s], y = (uint32_t)pairs[2 * s + 1];
    const uint32_t r = e < K ? x : y, c = e < K ? y : x;
    keys[e] = ((uint64_t)r << b) | c;
    vals[e] = sims[s];
  }
}

__global__ void gr_head_kernel(const uint64_t* __restrict__ keys, int64_t n,
                               int32_t* __restrict__ head) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    head[e] = (e == 0 || keys[e] != keys[e - 1]) ? 1 : 0;
}

// One thread per run head: sum the run in entry order, write column, value
// and row of the distinct entry; the last head also writes nnz.
__global__ void gr_emit_kernel(const uint64_t* __restrict__ keys, const float* __restrict__ vals,
                               const int32_t* __restrict__ head, const int32_t* __restrict__ pos,
                               int64_t n, int32_t b, int32_t* __restrict__ indices,
                               float* __restrict__ data, int32_t* __restrict__ urow,
                               int64_t* __restrict__ nnz) {
  const uint64_t cmask = (b >= 64) ? ~0ull : ((1ull << b) - 1);
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (e == n - 1) *nnz = (int64_t)pos[e] + head[e];
    if (!head[e]) continue;
    const uint64_t k = keys[e];
    float sum = vals[e];
    for (int64_t x = e + 1; x < n && keys[x] == k; ++x) sum = sum + vals[x];
    const int32_t o = pos[e];
    indices[o] = (int32_t)(k & cmask);
    data[o] = sum;
    urow[o] = (int32_t)(k >> b);
  }
}

// indptr[r] = number of distinct entries with row < r (r = 0..N).
__global__ void gr_indptr_kernel(const int32_t* __restrict__ urow, const int64_t* __restrict__ nnz,
                                 int64_t N, int32_t* __restrict__ indptr) {
  const int64_t m = *nnz;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r <= N;
       r += (int64_t)gridDim.x * blockDim.x) {
    int64_t lo = 0, hi = m;
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (urow[mid] < r) lo = mid + 1;
      else hi = mid;
    }
    indptr[r] = (int32_t)lo;
  }
}

}  // namespace

struct GraphWs {
  GBuf k0, k1, v0, v1, head, pos, urow, tmp;
};

GraphWs* graph_ws_new() { return new GraphWs(); }
void graph_ws_free(GraphWs* w) {
  if (!w) return;
  for (GBuf* b : {&w->k0, &w->k1, &w->v0, &w->v1, &w->head, &w->pos, &w->urow, &w->tmp})
    if (b->p) (void)hipFree(b->p);
  delete w;
}

// CSR of the symmetric graph: indptr[N+1], indices/data with room for 2K
// entries, *d_nnz = entries written.  Enqueued on st.
hipError_t graph_csr(GraphWs* w, const int32_t* pairs, int64_t K, const float* sims, int64_t N,
                     int32_t* indptr, int32_t* indices, float* data, int64_t* d_nnz,
                     hipStream_t st) {
  int32_t b = 1;
  while (b < 32 && (int64_t(1) << b) < N) ++b;
  const int64_t n = 2 * K;
  if (n == 0) {
    hipError_t e = hipMemsetAsync(d_nnz, 0, sizeof(int64_t), st);
    if (e != hipSuccess) return e;
    return hipMemsetAsync(indptr, 0, (size_t)(N + 1) * sizeof(int32_t), st);
  }
  hipError_t e;
  if ((e = w->k0.ensure(n * 8)) || (e = w->k1.ensure(n * 8)) || (e = w->v0.ensure(n * 4)) ||
      (e = w->v1.ensure(n * 4)) || (e = w->head.ensure(n * 4)) || (e = w->pos.ensure(n * 4)) ||
      (e = w->urow.ensure(n * 4)))
    return e;
  gr_expand_kernel<<<gr_grid(n), 256, 0, st>>>(pairs, K, sims, b, w->k0.as<uint64_t>(),
                                               w->v0.as<float>());
  rocprim::double_buffer<uint64_t> kb(w->k0.as<uint64_t>(), w->k1.as<uint64_t>());
  rocprim::double_buffer<float> vb(w->v0.as<float>(), w->v1.as<float>());
  size_t b_sort = 0, b_scan = 0;
  if ((e = rocprim::radix_sort_pairs(nullptr, b_sort, kb, vb, (size_t)n, 0, 2 * b, st)) ||
      (e = rocprim::exclusive_scan(nullptr, b_scan, w->head.as<int32_t>(), w->pos.as<int32_t>(),
                                   (int32_t)0, (size_t)n, rocprim::plus<int32_t>(), st)))
    return e;
  if ((e = w->tmp.ensure(std::max(b_sort, b_scan)))) return e;
  if ((e = rocprim::radix_sort_pairs(w->tmp.p, b_sort, kb, vb, (size_t)n, 0, 2 * b, st)))
    return e;
  gr_head_kernel<<<gr_grid(n), 256, 0, st>>>(kb.current(), n, w->head.as<int32_t>());
  if ((e = rocprim::exclusive_scan(w->tmp.p, b_scan, w->head.as<int32_t>(), w->pos.as<int32_t>(),
                                   (int32_t)0, (size_t)n, rocprim::plus<int32_t>(), st)))
    return e;
  gr_emit_kernel<<<gr_grid(n), 256, 0, st>>>(kb.current(), vb.current(), w->head.as<int32_t>(),
                                             w->pos.as<int32_t>(), n, b, indices, data,
                                             w->urow.as<int32_t>(), d_nnz);
  gr_indptr_kernel<<<gr_grid(N + 1), 256, 0, st>>>(w->urow.as<int32_t>(), d_nnz, N, indptr);
  return hipGetLastError();
}

}  // namespace ck
