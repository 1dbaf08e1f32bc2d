// MinHash sketches and LSH candidate pairs, on the device.
//
// Sketch (clustkit/sketch.py:17-80, 137-183; GPU twin sketch.py:193-281).
// For a sequence of length L >= k: the sketch_size smallest values, with
// multiplicity, of fmix64(code ^ seed) over its L-k+1 windows, ascending,
// padded with 2^64-1 when there are fewer windows; code = sum of the window's
// residue codes in base alpha, first residue most significant, mod 2^64 (every
// code enters, also codes >= alpha, as in the reference's rolling code).  A
// sequence shorter than k gets an all-zero sketch.
// Method: one workgroup per sequence.  The bottom-s found so far sits sorted
// at the head of an LDS buffer of SK_SORT words; each chunk of up to
// SK_SORT - s window hashes is appended behind it, the buffer is padded with
// 2^64-1 to a power of two and sorted by a bitonic network, and the head is
// again the bottom-s of everything seen.
//
// LSH (clustkit/lsh.py:20-29, 52-90, 93-178).  Table t: bucket id of a
// sequence = fmix tail of the multiply-add chain seed_t, sketch[band_t[0]],
// ...  Sequences are sorted by bucket id, stably, so members of one bucket
// stay in ascending sequence order (the order of the reference's GPU path,
// lsh.py:258-261); the first min(size, max_bucket) members of each bucket
// form all pairs, packed as lo * N + hi; the union over all tables is sorted
// and de-duplicated.
// With one band position per table, tables are run once per distinct
// position and keyed by the raw sketch value (the same grouping; see
// lsh_build).
// Method per group of tables: hash kernel → radix sort of (table, bucket
// id) keys → run starts by a max-scan of head positions, run ends
// scattered to the run start → pairs per member → exclusive scan → emit →
// radix sort + unique of the packed pairs; groups are merged by one more
// sort + unique.
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include <vector>

#include "ck_internal.h"

namespace ck {

namespace {

constexpr int SK_BLOCK = 256;
constexpr int SK_SORT = 2048;  // LDS sort buffer, 64-bit words

__device__ __forceinline__ uint64_t fmix64(uint64_t h) {
  h ^= h >> 33;
  h *= 0xFF51AFD7ED558CCDull;
  h ^= h >> 33;
  h *= 0xC4CEB9FE1A85EC53ull;
  h ^= h >> 33;
  return h;
}

// Ascending bitonic sort of buf[0, P), P a power of two, by the whole block.
__device__ __forceinline__ void bitonic_sort(uint64_t* buf, int P) {
  for (int kk = 2; kk <= P; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int t = threadIdx.x; t < (P >> 1); t += SK_BLOCK) {
        const int i = 2 * t - (t & (j - 1));
        const int l = i + j;
        const uint64_t a = buf[i], b = buf[l];
        const bool up = (i & kk) == 0;
        if ((a > b) == up) {
          buf[i] = b;
          buf[l] = a;
        }
      }
      __syncthreads();
    }
  }
}

__global__ __launch_bounds__(SK_BLOCK) void sk_sketch_kernel(
    const uint8_t* __restrict__ flat, const int64_t* __restrict__ offsets,
    const int32_t* __restrict__ lengths, int64_t n, int32_t k, int32_t s, uint64_t alpha,
    uint64_t seed, uint64_t* __restrict__ out) {
  __shared__ uint64_t buf[SK_SORT];
  const int cap = SK_SORT - s;
  for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
    const int32_t L = lengths[i];
    uint64_t* row = out + i * (int64_t)s;
    if (L < k) {
      for (int x = threadIdx.x; x < s; x += SK_BLOCK) row[x] = 0;
      continue;
    }
    const uint8_t* seq = flat + offsets[i];
    const int32_t nk = L - k + 1;
    for (int x = threadIdx.x; x < s; x += SK_BLOCK) buf[x] = ~0ull;
    for (int32_t c0 = 0; c0 < nk; c0 += cap) {
      const int32_t nc = min(cap, nk - c0);
      int P = 64;
      while (P < s + nc) P <<= 1;
      for (int x = s + threadIdx.x; x < P; x += SK_BLOCK) {
        const int32_t w = x - s;
        uint64_t h = ~0ull;
        if (w < nc) {
          const uint8_t* p = seq + c0 + w;
          uint64_t code = 0;
          for (int32_t j = 0; j < k; ++j) code = code * alpha + (uint64_t)p[j];
          h = fmix64(code ^ seed);
        }
        buf[x] = h;
      }
      __syncthreads();
      bitonic_sort(buf, P);
    }
    for (int x = threadIdx.x; x < s; x += SK_BLOCK) row[x] = buf[x];
    __syncthreads();
  }
}

constexpr uint64_t LSH_M1 = 0x517CC1B727220A95ull;
constexpr uint64_t LSH_C2 = 0xC4CEB9FE1A85EC53ull;

using u128 = rocprim::uint128_t;

// keys[t*n + i] = (t, bucket id of sequence i in table t0 + t) as t << 64 |
// id (raw: the band's one sketch value, which groups sequences the same
// way); vals = i.
__global__ __launch_bounds__(256) void lsh_hash_kernel(
    const uint64_t* __restrict__ sk, int64_t n, int32_t s, const int32_t* __restrict__ bands,
    const uint64_t* __restrict__ seeds, int32_t nb, int32_t t0, int32_t tg, bool raw,
    u128* __restrict__ keys, int32_t* __restrict__ vals) {
  const int64_t total = (int64_t)tg * n;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += stride) {
    const int32_t t = (int32_t)(x / n);
    const int64_t i = x - (int64_t)t * n;
    const int32_t* bi = bands + (int64_t)(t0 + t) * nb;
    const uint64_t* row = sk + i * (int64_t)s;
    uint64_t h;
    if (raw) {
      h = row[bi[0]];
    } else {
      h = seeds[t0 + t];
      for (int32_t b = 0; b < nb; ++b) h = h * LSH_M1 + row[bi[b]];
      h = (h ^ (h >> 33)) * LSH_C2;
      h ^= h >> 33;
    }
    keys[x] = ((u128)(uint64_t)t << 64) | (u128)h;
    vals[x] = (int32_t)i;
  }
}

// head[x] = x at the first member of a bucket run (or of a table), else 0.
__global__ __launch_bounds__(256) void lsh_head_kernel(const u128* __restrict__ keys,
                                                       int64_t total, int64_t n,
                                                       int32_t* __restrict__ head) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += stride) {
    const bool first = (x % n) == 0 || keys[x] != keys[x - 1];
    head[x] = first ? (int32_t)x : 0;
  }
}

// run_end[start[x]] = x + 1 at the last member of each run.
__global__ __launch_bounds__(256) void lsh_end_kernel(const u128* __restrict__ keys,
                                                      int64_t total, int64_t n,
                                                      const int32_t* __restrict__ start,
                                                      int32_t* __restrict__ run_end) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += stride) {
    const bool last = (x % n) == n - 1 || keys[x] != keys[x + 1];
    if (last) run_end[start[x]] = (int32_t)(x + 1);
  }
}

// cnt[x] = pairs member x forms with the members after it inside its run's
// first max_bucket members; cnt[total] = 0.
__global__ __launch_bounds__(256) void lsh_count_kernel(const int32_t* __restrict__ start,
                                                        const int32_t* __restrict__ run_end,
                                                        int64_t total, int32_t max_bucket,
                                                        int64_t* __restrict__ cnt) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x <= total; x += stride) {
    if (x == total) {
      cnt[x] = 0;
      continue;
    }
    const int32_t a = start[x];
    const int32_t m = min(run_end[a] - a, max_bucket);
    const int32_t r = (int32_t)x - a;
    cnt[x] = r < m ? (int64_t)(m - 1 - r) : 0;
  }
}

__global__ __launch_bounds__(256) void lsh_emit_kernel(const int32_t* __restrict__ vals,
                                                       const int64_t* __restrict__ cnt,
                                                       const int64_t* __restrict__ off,
                                                       int64_t total, int64_t n,
                                                       uint64_t* __restrict__ packed) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < total; x += stride) {
    const int64_t c = cnt[x];
    if (c == 0) continue;
    const int64_t o = off[x];
    const int64_t xi = vals[x];
    for (int64_t q = 1; q <= c; ++q) {
      const int64_t y = vals[x + q];
      const int64_t lo = xi < y ? xi : y;
      const int64_t hi = xi < y ? y : xi;
      packed[o + q - 1] = (uint64_t)(lo * n + hi);
    }
  }
}

__global__ __launch_bounds__(256) void lsh_unpack_kernel(const uint64_t* __restrict__ packed,
                                                         int64_t m, int64_t n,
                                                         int32_t* __restrict__ pairs) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; u < m; u += stride) {
    const uint64_t v = packed[u];
    pairs[2 * u] = (int32_t)(v / (uint64_t)n);
    pairs[2 * u + 1] = (int32_t)(v % (uint64_t)n);
  }
}

// Grow-only device buffer of the LSH workspace.
struct Buf {
  void* p = nullptr;
  size_t bytes = 0;
  // Capacity for `need` bytes; with keep, the first `keep` bytes survive.
  hipError_t ensure(size_t need, hipStream_t st, size_t keep = 0) {
    if (need <= bytes) return hipSuccess;
    size_t want = std::max<size_t>(need + need / 4, 256);
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, want);
    if (e != hipSuccess) return e;
    if (keep && p && (e = hipMemcpyAsync(q, p, keep, hipMemcpyDeviceToDevice, st))) {
      (void)hipFree(q);
      return e;
    }
    if (p) {
      if (keep && (e = hipStreamSynchronize(st))) return e;
      (void)hipFree(p);
    }
    p = q;
    bytes = want;
    return hipSuccess;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

unsigned grid_for(int64_t work) {
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>((work + 255) / 256, 65536));
}

}  // namespace

// Workspace of the LSH build (one per stream): buffers grow to the largest
// call and are reused; `res` holds the last call's de-duplicated pairs.
struct LshWs {
  Buf bands, seeds, k0, k1, v0, v1, head, start, rend, cnt, off, tmp, packed, alt, uniq, acc,
      res, ucount;
  int64_t res_m = 0;
  ~LshWs() {
    for (Buf* b : {&bands, &seeds, &k0, &k1, &v0, &v1, &head, &start, &rend, &cnt, &off, &tmp,
                   &packed, &alt, &uniq, &acc, &res, &ucount})
      b->release();
  }
};

LshWs* lsh_ws_new() { return new LshWs(); }
void lsh_ws_free(LshWs* w) { delete w; }

namespace {

// Sort + unique of in[0, m) (end_bit significant bits, `alt` as the second
// sort buffer) into out; *m_out = number of distinct values.
hipError_t sort_unique(LshWs* w, Buf& in, int64_t m, unsigned end_bit, Buf& out,
                       int64_t* m_out, hipStream_t st) {
  *m_out = 0;
  if (m == 0) return hipSuccess;
  hipError_t e;
  if ((e = w->alt.ensure((size_t)m * 8, st)) || (e = out.ensure((size_t)m * 8, st)) ||
      (e = w->ucount.ensure(sizeof(unsigned long long), st)))
    return e;
  rocprim::double_buffer<uint64_t> kb(in.as<uint64_t>(), w->alt.as<uint64_t>());
  size_t sbytes = 0, ubytes = 0;
  if ((e = rocprim::radix_sort_keys(nullptr, sbytes, kb, (size_t)m, 0, end_bit, st))) return e;
  if ((e = rocprim::unique(nullptr, ubytes, kb.current(), out.as<uint64_t>(),
                           w->ucount.as<unsigned long long>(), (size_t)m,
                           rocprim::equal_to<uint64_t>(), st)))
    return e;
  if ((e = w->tmp.ensure(std::max(sbytes, ubytes), st))) return e;
  if ((e = rocprim::radix_sort_keys(w->tmp.p, sbytes, kb, (size_t)m, 0, end_bit, st))) return e;
  if ((e = rocprim::unique(w->tmp.p, ubytes, kb.current(), out.as<uint64_t>(),
                           w->ucount.as<unsigned long long>(), (size_t)m,
                           rocprim::equal_to<uint64_t>(), st)))
    return e;
  unsigned long long u = 0;
  if ((e = hipMemcpyAsync(&u, w->ucount.p, sizeof(u), hipMemcpyDeviceToHost, st))) return e;
  if ((e = hipStreamSynchronize(st))) return e;
  *m_out = (int64_t)u;
  return hipSuccess;
}

}  // namespace

hipError_t sketch_build(const uint8_t* flat, const int64_t* offsets, const int32_t* lengths,
                        int64_t n, int32_t k, int32_t s, int32_t alpha, uint64_t seed,
                        uint64_t* out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  const unsigned grid = (unsigned)std::min<int64_t>(n, 65536);
  sk_sketch_kernel<<<grid, SK_BLOCK, 0, st>>>(flat, offsets, lengths, n, k, s, (uint64_t)alpha,
                                              seed, out);
  return hipGetLastError();
}

int32_t sketch_max_size() { return SK_SORT / 2; }

// Candidate pairs of all tables, packed and de-duplicated, into w->res
// (w->res_m entries).
hipError_t lsh_build(LshWs* w, const uint64_t* sk, int64_t n, int32_t s, const int32_t* h_bands,
                     const int64_t* h_seeds, int32_t T, int32_t nb, int32_t max_bucket,
                     int64_t group_elems, int64_t* m_out, hipStream_t st) {
  *m_out = 0;
  w->res_m = 0;
  hipError_t e;
  // One band position per table: the bucket id is a bijection of that one
  // sketch value (seed * M1 + v, the xor-shifts and the odd multiplier are
  // each invertible mod 2^64), so a table groups sequences exactly by the
  // value at its position, whatever its seed, and tables that draw the same
  // position yield the same pairs.  Such tables are run once per distinct
  // position, keyed by the raw value.
  std::vector<int32_t> cols;
  const bool raw = nb == 1;
  if (raw) {
    std::vector<char> seen(s, 0);
    for (int32_t t = 0; t < T; ++t)
      if (!seen[h_bands[t]]) {
        seen[h_bands[t]] = 1;
        cols.push_back(h_bands[t]);
      }
    h_bands = cols.data();
    T = (int32_t)cols.size();
  }
  if ((e = w->bands.ensure((size_t)T * nb * 4, st)) || (e = w->seeds.ensure((size_t)T * 8, st)))
    return e;
  if ((e = hipMemcpyAsync(w->bands.p, h_bands, (size_t)T * nb * 4, hipMemcpyHostToDevice, st)) ||
      (e = hipMemcpyAsync(w->seeds.p, h_seeds, (size_t)T * 8, hipMemcpyHostToDevice, st)))
    return e;
  unsigned end_bit = 1;
  {
    const unsigned __int128 nn = (unsigned __int128)n * (unsigned __int128)n;
    while (end_bit < 64 && ((unsigned __int128)1 << end_bit) < nn) ++end_bit;
  }
  // Positions inside a group are int32: at most 2^30 entries per group.
  group_elems = std::min<int64_t>(group_elems, int64_t(1) << 30);
  const int32_t tg_max = (int32_t)std::max<int64_t>(1, std::min<int64_t>(T, group_elems / n));
  int64_t acc_n = 0;
  int groups = 0;
  for (int32_t t0 = 0; t0 < T; t0 += tg_max) {
    const int32_t tg = std::min(tg_max, T - t0);
    const int64_t tot = (int64_t)tg * n;
    if ((e = w->k0.ensure(tot * 16, st)) || (e = w->k1.ensure(tot * 16, st)) ||
        (e = w->v0.ensure(tot * 4, st)) || (e = w->v1.ensure(tot * 4, st)) ||
        (e = w->head.ensure(tot * 4, st)) || (e = w->start.ensure(tot * 4, st)) ||
        (e = w->rend.ensure(tot * 4, st)) || (e = w->cnt.ensure((tot + 1) * 8, st)) ||
        (e = w->off.ensure((tot + 1) * 8, st)))
      return e;
    lsh_hash_kernel<<<grid_for(tot), 256, 0, st>>>(sk, n, s, w->bands.as<int32_t>(),
                                                   w->seeds.as<uint64_t>(), nb, t0, tg, raw,
                                                   w->k0.as<u128>(), w->v0.as<int32_t>());
    if ((e = hipGetLastError())) return e;
    // One radix sort over (table, bucket id): stable, so members of a bucket
    // keep ascending sequence order, and table t ends up in [t*n, (t+1)*n).
    rocprim::double_buffer<u128> kb(w->k0.as<u128>(), w->k1.as<u128>());
    rocprim::double_buffer<int32_t> vb(w->v0.as<int32_t>(), w->v1.as<int32_t>());
    unsigned tbits = 1;
    while (tbits < 31 && (int64_t(1) << tbits) < tg) ++tbits;
    const unsigned key_bits = 64 + tbits;
    size_t b_sort = 0, b_inc = 0, b_exc = 0;
    if ((e = rocprim::radix_sort_pairs(nullptr, b_sort, kb, vb, (size_t)tot, 0, key_bits, st)) ||
        (e = rocprim::inclusive_scan(nullptr, b_inc, w->head.as<int32_t>(),
                                     w->start.as<int32_t>(), (size_t)tot,
                                     rocprim::maximum<int32_t>(), st)) ||
        (e = rocprim::exclusive_scan(nullptr, b_exc, w->cnt.as<int64_t>(), w->off.as<int64_t>(),
                                     (int64_t)0, (size_t)tot + 1, rocprim::plus<int64_t>(), st)))
      return e;
    if ((e = w->tmp.ensure(std::max({b_sort, b_inc, b_exc}), st))) return e;
    if ((e = rocprim::radix_sort_pairs(w->tmp.p, b_sort, kb, vb, (size_t)tot, 0, key_bits, st)))
      return e;
    lsh_head_kernel<<<grid_for(tot), 256, 0, st>>>(kb.current(), tot, n, w->head.as<int32_t>());
    if ((e = hipGetLastError())) return e;
    if ((e = rocprim::inclusive_scan(w->tmp.p, b_inc, w->head.as<int32_t>(),
                                     w->start.as<int32_t>(), (size_t)tot,
                                     rocprim::maximum<int32_t>(), st)))
      return e;
    lsh_end_kernel<<<grid_for(tot), 256, 0, st>>>(kb.current(), tot, n, w->start.as<int32_t>(),
                                                  w->rend.as<int32_t>());
    lsh_count_kernel<<<grid_for(tot + 1), 256, 0, st>>>(w->start.as<int32_t>(),
                                                        w->rend.as<int32_t>(), tot, max_bucket,
                                                        w->cnt.as<int64_t>());
    if ((e = hipGetLastError())) return e;
    if ((e = rocprim::exclusive_scan(w->tmp.p, b_exc, w->cnt.as<int64_t>(), w->off.as<int64_t>(),
                                     (int64_t)0, (size_t)tot + 1, rocprim::plus<int64_t>(), st)))
      return e;
    int64_t m = 0;
    if ((e = hipMemcpyAsync(&m, w->off.as<int64_t>() + tot, 8, hipMemcpyDeviceToHost, st))) return e;
    if ((e = hipStreamSynchronize(st))) return e;
    if (m == 0) continue;
    if ((e = w->packed.ensure((size_t)m * 8, st))) return e;
    lsh_emit_kernel<<<grid_for(tot), 256, 0, st>>>(vb.current(), w->cnt.as<int64_t>(),
                                                   w->off.as<int64_t>(), tot, n,
                                                   w->packed.as<uint64_t>());
    if ((e = hipGetLastError())) return e;
    int64_t u = 0;
    if ((e = sort_unique(w, w->packed, m, end_bit, w->uniq, &u, st))) return e;
    // Append this group's distinct pairs to the accumulator.
    if ((e = w->acc.ensure((size_t)(acc_n + u) * 8, st, (size_t)acc_n * 8))) return e;
    if ((e = hipMemcpyAsync(w->acc.as<uint64_t>() + acc_n, w->uniq.p, (size_t)u * 8,
                            hipMemcpyDeviceToDevice, st)))
      return e;
    acc_n += u;
    ++groups;
  }
  if (groups == 0) return hipSuccess;
  int64_t u = acc_n;
  if (groups == 1) {
    std::swap(w->acc, w->res);
  } else if ((e = sort_unique(w, w->acc, acc_n, end_bit, w->res, &u, st))) {
    return e;
  }
  w->res_m = u;
  *m_out = u;
  return hipSuccess;
}

hipError_t lsh_unpack(const LshWs* w, int64_t n, int32_t* pairs, hipStream_t st) {
  if (w->res_m == 0) return hipSuccess;
  lsh_unpack_kernel<<<grid_for(w->res_m), 256, 0, st>>>(w->res.as<uint64_t>(), w->res_m, n,
                                                        pairs);
  return hipGetLastError();
}

}  // namespace ck
