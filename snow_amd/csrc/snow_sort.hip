// snow_sort.hip - stable LSD radix sort of (cell key, slot) pairs for gfx950.
//
// Each pass sorts by one 8-bit digit:
//   k_radix_hist     per-block digit histograms (LDS atomics; counts are order-independent)
//   k_col_*          exclusive offsets from the block-major histogram table (column sums per
//                    group of blocks, one block for group and digit prefixes, column apply)
//   k_radix_scatter  stable scatter: a block walks its 4096 keys in 16 rounds of 256 in input
//                    order; within a wave, keys with equal digits are matched with 8 ballots and
//                    ranked by lane, waves are ranked in order through LDS counters.
// Stability makes the result a deterministic function of the input order. No hipMemset and no
// host synchronisation: the pass sequence can be captured in a HIP graph.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace snow {

namespace {
constexpr int RADIX_BITS = 8;
constexpr int BINS = 1 << RADIX_BITS;
constexpr int SORT_THREADS = 256;
constexpr int SORT_ITEMS = 16;
constexpr int SORT_TILE = SORT_THREADS * SORT_ITEMS;  // keys per block

__global__ void __launch_bounds__(SORT_THREADS) k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n,
                                                             int shift, uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[BINS];
  const int t = threadIdx.x;
  h[t] = 0u;
  __syncthreads();
  const uint32_t base = blockIdx.x * uint32_t(SORT_TILE);
  uint32_t kreg[SORT_ITEMS];
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const uint32_t i = base + uint32_t(r * SORT_THREADS + t);
    kreg[r] = i < n ? keys[i] : 0xffffffffu;
  }
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r)
    if (base + uint32_t(r * SORT_THREADS + t) < n) atomicAdd(&h[(kreg[r] >> shift) & (BINS - 1)], 1u);
  __syncthreads();
  hist[size_t(blockIdx.x) * BINS + t] = h[t];  // block-major: row = block, column = digit
}

// Offsets from the block-major histogram: O[b][d] = sum_{d'<d} total(d') + sum_{b'<b} hist[b'][d].
// Rows are processed in groups of GROUP blocks: column sums per group, one block scanning the
// group sums and the digit totals, then each group rewrites its rows as offsets.
constexpr int GROUP = 32;

__global__ void __launch_bounds__(BINS) k_col_sums(const uint32_t* __restrict__ hist, uint32_t nblocks,
                                                   uint32_t* __restrict__ gsum) {
  const int d = threadIdx.x;
  const uint32_t b0 = blockIdx.x * GROUP;
  uint32_t s = 0u;
#pragma unroll 8
  for (int r = 0; r < GROUP; ++r)
    if (b0 + r < nblocks) s += hist[size_t(b0 + r) * BINS + d];
  gsum[size_t(blockIdx.x) * BINS + d] = s;
}

__global__ void __launch_bounds__(BINS) k_col_scan(uint32_t* __restrict__ gsum, uint32_t ngroups) {
  __shared__ uint32_t s_tot[BINS];
  const int d = threadIdx.x;
  uint32_t s = 0u;
  for (uint32_t g = 0; g < ngroups; ++g) {
    const uint32_t v = gsum[size_t(g) * BINS + d];
    gsum[size_t(g) * BINS + d] = s;
    s += v;
  }
  s_tot[d] = s;
  __syncthreads();
  // inclusive Hillis-Steele scan of the digit totals (one per thread)
  for (int off = 1; off < BINS; off <<= 1) {
    const uint32_t v = d >= off ? s_tot[d - off] : 0u;
    __syncthreads();
    s_tot[d] += v;
    __syncthreads();
  }
  const uint32_t dbase = s_tot[d] - s;
  for (uint32_t g = 0; g < ngroups; ++g) gsum[size_t(g) * BINS + d] += dbase;
}

__global__ void __launch_bounds__(BINS) k_col_apply(uint32_t* __restrict__ hist, uint32_t nblocks,
                                                    const uint32_t* __restrict__ gsum) {
  const int d = threadIdx.x;
  const uint32_t b0 = blockIdx.x * GROUP;
  uint32_t s = gsum[size_t(blockIdx.x) * BINS + d];
  for (int r = 0; r < GROUP; ++r) {
    if (b0 + r >= nblocks) break;
    const size_t idx = size_t(b0 + r) * BINS + d;
    const uint32_t v = hist[idx];
    hist[idx] = s;
    s += v;
  }
}

__global__ void __launch_bounds__(SORT_THREADS) k_radix_scatter(const uint32_t* __restrict__ kin,
                                                                const uint32_t* __restrict__ vin,
                                                                uint32_t* __restrict__ kout,
                                                                uint32_t* __restrict__ vout, uint32_t n, int shift,
                                                                const uint32_t* __restrict__ offsets,
                                                                uint32_t nblocks) {
  constexpr int NW = SORT_THREADS / 64;
  __shared__ uint32_t s_base[BINS];
  __shared__ uint32_t s_wcnt[NW][BINS];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  s_base[t] = offsets[size_t(blockIdx.x) * BINS + t];
#pragma unroll
  for (int w = 0; w < NW; ++w) s_wcnt[w][t] = 0u;
  __syncthreads();
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  const uint32_t base = blockIdx.x * uint32_t(SORT_TILE);
  // all loads of the block's keys and values are issued before the ranking rounds
  uint32_t kreg[SORT_ITEMS], vreg[SORT_ITEMS];
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const uint32_t i = base + uint32_t(r * SORT_THREADS + t);
    kreg[r] = i < n ? kin[i] : 0u;
    vreg[r] = i < n ? vin[i] : 0u;
  }
#pragma unroll
  for (int r = 0; r < SORT_ITEMS; ++r) {
    const uint32_t i = base + uint32_t(r * SORT_THREADS + t);
    const bool valid = i < n;
    const uint32_t k = kreg[r];
    const uint32_t v = vreg[r];
    const uint32_t d = (k >> shift) & (BINS - 1);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < RADIX_BITS; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = __popcll(peers & lt_mask);
    if (valid && rank == 0) s_wcnt[wave][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = s_base[d] + rank;
#pragma unroll
      for (int w = 0; w < NW; ++w)
        if (w < wave) pos += s_wcnt[w][d];
      kout[pos] = k;
      vout[pos] = v;
    }
    __syncthreads();
    {
      uint32_t add = 0u;
#pragma unroll
      for (int w = 0; w < NW; ++w) {
        add += s_wcnt[w][t];
        s_wcnt[w][t] = 0u;
      }
      s_base[t] += add;
    }
    __syncthreads();
  }
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return uint32_t((a + b - 1) / b); }

struct SortLayout {
  uint32_t nblocks, ngroups;
  size_t hist_off, gsum_off, ktmp_off, vtmp_off, total;
};

SortLayout layout(uint32_t n) {
  SortLayout L;
  L.nblocks = div_up(n ? n : 1, SORT_TILE);
  L.ngroups = div_up(L.nblocks, GROUP);
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  L.hist_off = 0;
  L.gsum_off = al(L.hist_off + sizeof(uint32_t) * BINS * L.nblocks);
  L.ktmp_off = al(L.gsum_off + sizeof(uint32_t) * BINS * L.ngroups);
  L.vtmp_off = al(L.ktmp_off + sizeof(uint32_t) * (n ? n : 1));
  L.total = al(L.vtmp_off + sizeof(uint32_t) * (n ? n : 1));
  return L;
}
}  // namespace

// Stable sort of (keys_in, vals_in) by the low end_bit bits of the key into (keys_out,
// vals_out). With temp == nullptr, returns the temporary storage size in *temp_bytes.
hipError_t radix_sort_pairs(void* temp, size_t* temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                            const uint32_t* vals_in, uint32_t* vals_out, uint32_t n, int end_bit,
                            hipStream_t stream) {
  const SortLayout L = layout(n);
  if (!temp) {
    *temp_bytes = L.total;
    return hipSuccess;
  }
  if (*temp_bytes < L.total) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  char* base = static_cast<char*>(temp);
  uint32_t* hist = reinterpret_cast<uint32_t*>(base + L.hist_off);
  uint32_t* gsum = reinterpret_cast<uint32_t*>(base + L.gsum_off);
  uint32_t* ktmp = reinterpret_cast<uint32_t*>(base + L.ktmp_off);
  uint32_t* vtmp = reinterpret_cast<uint32_t*>(base + L.vtmp_off);
  const int passes = end_bit <= 0 ? 1 : (end_bit + RADIX_BITS - 1) / RADIX_BITS;
  // choose the first destination so that the last pass writes keys_out
  const uint32_t* kin = keys_in;
  const uint32_t* vin = vals_in;
  for (int p = 0; p < passes; ++p) {
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    uint32_t* ko = to_out ? keys_out : ktmp;
    uint32_t* vo = to_out ? vals_out : vtmp;
    const int shift = p * RADIX_BITS;
    k_radix_hist<<<L.nblocks, SORT_THREADS, 0, stream>>>(kin, n, shift, hist);
    k_col_sums<<<L.ngroups, BINS, 0, stream>>>(hist, L.nblocks, gsum);
    k_col_scan<<<1, BINS, 0, stream>>>(gsum, L.ngroups);
    k_col_apply<<<L.ngroups, BINS, 0, stream>>>(hist, L.nblocks, gsum);
    k_radix_scatter<<<L.nblocks, SORT_THREADS, 0, stream>>>(kin, vin, ko, vo, n, shift, hist, L.nblocks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    kin = ko;
    vin = vo;
  }
  return hipSuccess;
}

}  // namespace snow
