// Device build of the CSR k-mer inverted index of a sequence table that is
// already in HBM.
//
// Layout (clustkit/kmer_index.py:117-162, database format of
// clustkit/database.py:127-136): for each k-mer code c (code = sum of
// residues in base alpha, first residue most significant), postings
// [kmer_offsets[c], kmer_offsets[c+1]) hold (sequence, position) in
// ascending (sequence, position) order; a window enters only when all of its
// k residues are below alpha.  The device form splits each posting into
// int32 target id and int32 position.
//
// Method: one wavefront per sequence enumerates its windows in position
// order (count pass, exclusive scan over sequences, emit pass with a ballot
// prefix), so the emitted list is in (sequence, position) order; a stable
// LSD radix sort by code (rocPRIM) then yields exactly the reference order.
// Offsets come from a lower-bound search per code over the sorted codes.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "ck_internal.h"

namespace ck {

namespace {

__device__ __forceinline__ bool window_at(const uint8_t* seq, int32_t p, int32_t k, int32_t alpha,
                                          uint32_t* code) {
  uint32_t c = 0;
  bool ok = true;
  for (int32_t j = 0; j < k; ++j) {
    const uint32_t r = seq[p + j];
    ok &= r < (uint32_t)alpha;
    c = c * (uint32_t)alpha + r;
  }
  *code = c;
  return ok;
}

__global__ __launch_bounds__(256) void kb_count_kernel(const uint8_t* __restrict__ flat,
                                                       const int64_t* __restrict__ offsets,
                                                       const int32_t* __restrict__ lengths,
                                                       int32_t n, int32_t k, int32_t alpha,
                                                       int64_t* __restrict__ count) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / WAVE;
  for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; s < n; s += nwaves) {
    const int32_t nw = lengths[s] - k + 1;
    const uint8_t* seq = flat + offsets[s];
    int64_t c = 0;
    for (int32_t p0 = 0; p0 < nw; p0 += WAVE) {
      const int32_t p = p0 + lane;
      uint32_t code;
      const bool v = p < nw && window_at(seq, p, k, alpha, &code);
      c += __popcll(__ballot(v));
    }
    if (lane == 0) count[s] = c;
  }
}

__global__ __launch_bounds__(256) void kb_emit_kernel(const uint8_t* __restrict__ flat,
                                                      const int64_t* __restrict__ offsets,
                                                      const int32_t* __restrict__ lengths,
                                                      int32_t n, int32_t k, int32_t alpha,
                                                      const int64_t* __restrict__ start,
                                                      uint32_t* __restrict__ keys,
                                                      uint64_t* __restrict__ vals) {
  const int lane = threadIdx.x & (WAVE - 1);
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (WAVE - lane));
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / WAVE;
  for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; s < n; s += nwaves) {
    const int32_t nw = lengths[s] - k + 1;
    const uint8_t* seq = flat + offsets[s];
    int64_t base = start[s];
    for (int32_t p0 = 0; p0 < nw; p0 += WAVE) {
      const int32_t p = p0 + lane;
      uint32_t code = 0;
      const bool v = p < nw && window_at(seq, p, k, alpha, &code);
      const uint64_t m = __ballot(v);
      if (v) {
        const int64_t r = base + __popcll(m & below);
        keys[r] = code;
        vals[r] = ((uint64_t)s << 32) | (uint32_t)p;
      }
      base += __popcll(m);
    }
  }
}

// kmer_offsets[c] = first index whose code is >= c (c = 0 .. ncodes).
__global__ void kb_offsets_kernel(const uint32_t* __restrict__ keys, int64_t total,
                                  int64_t ncodes, int64_t* __restrict__ kmer_offsets) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (c > ncodes) return;
  int64_t lo = 0, hi = total;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if ((int64_t)keys[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  kmer_offsets[c] = lo;
}

__global__ void kb_freqs_split_kernel(const int64_t* __restrict__ kmer_offsets, int64_t ncodes,
                                      const uint64_t* __restrict__ vals, int64_t total,
                                      int32_t* __restrict__ freqs, int32_t* __restrict__ tid,
                                      int32_t* __restrict__ pos) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total || i < ncodes;
       i += stride) {
    if (i < ncodes) freqs[i] = (int32_t)(kmer_offsets[i + 1] - kmer_offsets[i]);
    if (i < total) {
      const uint64_t v = vals[i];
      tid[i] = (int32_t)(v >> 32);
      pos[i] = (int32_t)(uint32_t)v;
    }
  }
}

// Per-(k-mer, 8192-target tile) offsets: first posting of code c whose
// target is >= j * 8192 (j = 0 .. nt-1), and kmer_offsets[c+1] at j = nt.
__global__ void kb_tile_offsets_kernel(const int64_t* __restrict__ kmer_offsets,
                                       const int32_t* __restrict__ tid, int64_t ncodes,
                                       int32_t nt, int64_t* __restrict__ out) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= ncodes * (nt + 1)) return;
  const int64_t c = x / (nt + 1);
  const int32_t j = (int32_t)(x - c * (nt + 1));
  if (j == nt) {
    out[x] = kmer_offsets[c + 1];
    return;
  }
  const int32_t bound = j * 8192;
  int64_t lo = kmer_offsets[c], hi = kmer_offsets[c + 1];
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (tid[mid] < bound) lo = mid + 1;
    else hi = mid;
  }
  out[x] = lo;
}

struct Tmp {
  void* p = nullptr;
  ~Tmp() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

// Returns hipSuccess or the first HIP error; *status = 1 when `cap` is
// smaller than the number of windows (nothing is written then).
hipError_t kmer_index_build(const uint8_t* flat, const int64_t* offsets, const int32_t* lengths,
                            int32_t n, int32_t k, int32_t alpha, int64_t cap,
                            int64_t* kmer_offsets, int32_t* kmer_freqs, int32_t* post_tid,
                            int32_t* post_pos, int64_t* out_total, int* status,
                            hipStream_t st) {
  *status = 0;
  *out_total = 0;
  int64_t ncodes = 1;
  for (int j = 0; j < k; ++j) ncodes *= alpha;
  unsigned end_bit = 0;
  while (end_bit < 32 && ((int64_t)1 << end_bit) < ncodes) ++end_bit;
  if (end_bit == 0) end_bit = 1;
  hipError_t e;
  Tmp cnt, start;
  if ((e = hipMalloc(&cnt.p, sizeof(int64_t) * (n + 1)))) return e;
  if ((e = hipMalloc(&start.p, sizeof(int64_t) * (n + 1)))) return e;
  int64_t* d_cnt = static_cast<int64_t*>(cnt.p);
  int64_t* d_start = static_cast<int64_t*>(start.p);
  const int blocks = 4096;
  if (n > 0) {
    kb_count_kernel<<<blocks, 256, 0, st>>>(flat, offsets, lengths, n, k, alpha, d_cnt);
    if ((e = hipGetLastError())) return e;
  }
  if ((e = hipMemsetAsync(d_cnt + n, 0, sizeof(int64_t), st))) return e;
  size_t tbytes = 0;
  if ((e = rocprim::exclusive_scan(nullptr, tbytes, d_cnt, d_start, (int64_t)0, (size_t)n + 1,
                                   rocprim::plus<int64_t>(), st)))
    return e;
  Tmp ttmp;
  if ((e = hipMalloc(&ttmp.p, std::max<size_t>(tbytes, 16)))) return e;
  if ((e = rocprim::exclusive_scan(ttmp.p, tbytes, d_cnt, d_start, (int64_t)0, (size_t)n + 1,
                                   rocprim::plus<int64_t>(), st)))
    return e;
  int64_t total = 0;
  if ((e = hipMemcpyAsync(&total, d_start + n, sizeof(int64_t), hipMemcpyDeviceToHost, st)))
    return e;
  if ((e = hipStreamSynchronize(st))) return e;
  if (total > cap) {
    *status = 1;
    *out_total = total;
    return hipSuccess;
  }
  Tmp k0, k1, v0, v1;
  const size_t nt = (size_t)std::max<int64_t>(total, 1);
  if ((e = hipMalloc(&k0.p, nt * 4)) || (e = hipMalloc(&k1.p, nt * 4)) ||
      (e = hipMalloc(&v0.p, nt * 8)) || (e = hipMalloc(&v1.p, nt * 8)))
    return e;
  if (n > 0) {
    kb_emit_kernel<<<blocks, 256, 0, st>>>(flat, offsets, lengths, n, k, alpha, d_start,
                                           static_cast<uint32_t*>(k0.p),
                                           static_cast<uint64_t*>(v0.p));
    if ((e = hipGetLastError())) return e;
  }
  rocprim::double_buffer<uint32_t> kb(static_cast<uint32_t*>(k0.p), static_cast<uint32_t*>(k1.p));
  rocprim::double_buffer<uint64_t> vb(static_cast<uint64_t*>(v0.p), static_cast<uint64_t*>(v1.p));
  size_t sbytes = 0;
  if ((e = rocprim::radix_sort_pairs(nullptr, sbytes, kb, vb, (size_t)total, 0, end_bit, st)))
    return e;
  Tmp stmp;
  if ((e = hipMalloc(&stmp.p, std::max<size_t>(sbytes, 16)))) return e;
  if (total > 0 &&
      (e = rocprim::radix_sort_pairs(stmp.p, sbytes, kb, vb, (size_t)total, 0, end_bit, st)))
    return e;
  kb_offsets_kernel<<<(unsigned)((ncodes + 1 + 255) / 256), 256, 0, st>>>(kb.current(), total,
                                                                          ncodes, kmer_offsets);
  if ((e = hipGetLastError())) return e;
  const int64_t work = std::max<int64_t>(total, ncodes);
  const unsigned sblocks = (unsigned)std::min<int64_t>((work + 255) / 256, 65536);
  kb_freqs_split_kernel<<<std::max(sblocks, 1u), 256, 0, st>>>(kmer_offsets, ncodes,
                                                                vb.current(), total, kmer_freqs,
                                                                post_tid, post_pos);
  if ((e = hipGetLastError())) return e;
  if ((e = hipStreamSynchronize(st))) return e;
  *out_total = total;
  return hipSuccess;
}

namespace {

// Window codes of every sequence (one wavefront per sequence): the code of
// the window at position p of sequence s, plus one, goes to
// codes[code_off[s] + p]; a window holding a residue >= alpha, and the
// padding from the last window to code_off[s + 1] (a multiple of 8), read 0.
__global__ __launch_bounds__(256) void kb_codes_kernel(const uint8_t* __restrict__ flat,
                                                       const int64_t* __restrict__ offsets,
                                                       const int32_t* __restrict__ lengths,
                                                       int32_t n, int32_t k, int32_t alpha,
                                                       const int64_t* __restrict__ code_off,
                                                       int16_t* __restrict__ codes) {
  const int lane = threadIdx.x & (WAVE - 1);
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) / WAVE;
  for (int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / WAVE; s < n; s += nwaves) {
    const int32_t nw = lengths[s] - k + 1;
    const uint8_t* seq = flat + offsets[s];
    const int64_t c0 = code_off[s];
    const int32_t npad = (int32_t)(code_off[s + 1] - c0);
    for (int32_t p = lane; p < npad; p += WAVE) {
      uint32_t code = 0;
      const bool v = p < nw && window_at(seq, p, k, alpha, &code);
      codes[c0 + p] = v ? (int16_t)(code + 1) : (int16_t)0;
    }
  }
}

}  // namespace

hipError_t kmer_db_codes(const uint8_t* flat, const int64_t* offsets, const int32_t* lengths,
                         int32_t n, int32_t k, int32_t alpha, const int64_t* code_off,
                         int16_t* codes, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<int64_t>(((int64_t)n + 3) / 4, 8192);
  kb_codes_kernel<<<blocks, 256, 0, st>>>(flat, offsets, lengths, n, k, alpha, code_off, codes);
  return hipGetLastError();
}

hipError_t kmer_tile_offsets(const int64_t* kmer_offsets, const int32_t* post_tid, int64_t ncodes,
                             int32_t num_db, int64_t* out, hipStream_t st) {
  const int32_t nt = (num_db + 8191) / 8192;
  const int64_t work = ncodes * (nt + 1);
  if (work == 0) return hipSuccess;
  kb_tile_offsets_kernel<<<(unsigned)((work + 255) / 256), 256, 0, st>>>(kmer_offsets, post_tid,
                                                                        ncodes, nt, out);
  return hipGetLastError();
}

}  // namespace ck
