// snow_sort.hpp - stable LSD radix sort of (cell key, slot) pairs for gfx950.
//
// The key bits are split into the fewest passes of at most 11 bits (22-bit keys: two passes).
// Each pass runs three kernels over tiles of 4096 keys, one workgroup per tile. The ranking is a
// chain of dependent ballots and LDS round trips, so while all tiles fit one workgroup per CU a
// tile gets 16 waves (short chains); with more tiles it gets 4, so that four tiles share a CU:
//   k_radix_hist     per-tile digit histogram, written tile-major (hist[tile][d], one coalesced
//                    row per tile); each thread counts 4 (or 16) consecutive keys and adds runs of
//                    equal digits with one LDS atomic (nearly sorted input makes long runs)
//   k_radix_colscan  64, 32 or 16 digits per workgroup: exclusive scan of each digit's column
//                    over the tiles (16, 32 or 64 groups of tiles per digit, row segments read
//                    coalesced, up to 16 tiles per group kept in registers between the sum and
//                    the scan), and the digit totals
//   k_radix_scatter  each wave owns 256 (or 1024) consecutive keys of the tile and ranks them in
//                    input order (keys with equal digits in a round of 64 are matched with
//                    ballots, running per-wave digit counters in LDS carry the rank across
//                    rounds); the wave bases follow from the per-wave counts, the tile's row
//                    offsets and the exclusive scan of the digit totals, which every tile
//                    recomputes.
// Stability makes the result a deterministic function of the input order. No hipMemset and no
// host synchronisation: the pass sequence can be captured in a HIP graph. The first launch can
// also carry the step's start (step_begin_block in its workgroup 0), which saves the step a launch.
// The first pass's histogram can also come counted beforehand (hist0: the writer of the keys adds
// each run of equal (tile, digit) with one atomic); its scatter then clears each row after reading
// it, so that the writer can count the next keys into zeros.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "snow_kernels.hpp"

namespace snow {

namespace {
constexpr int MAX_BITS = 11;
constexpr int MAX_BINS = 1 << MAX_BITS;
constexpr int SORT_TILE = SORT_TILE_KEYS;  // keys per tile

template <int THREADS>
__global__ void __launch_bounds__(THREADS) k_radix_hist(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                        int bits, uint32_t* __restrict__ hist, StepBegin begin,
                                                        int with_begin) {
  constexpr int ITEMS = SORT_TILE / THREADS;  // consecutive keys per thread
  __shared__ uint32_t h[MAX_BINS];
  const int t = threadIdx.x;
  if (with_begin && blockIdx.x == 0) step_begin_block(begin, t, THREADS);
  const int bins = 1 << bits;
  const uint32_t mask = uint32_t(bins - 1);
  for (int d = t; d < bins; d += THREADS) h[d] = 0u;
  __syncthreads();
  const uint32_t i0 = blockIdx.x * uint32_t(SORT_TILE) + uint32_t(t) * ITEMS;
  uint32_t k[ITEMS];
  if (i0 + ITEMS <= n) {
    const uint4* p = reinterpret_cast<const uint4*>(keys + i0);
#pragma unroll
    for (int q = 0; q < ITEMS / 4; ++q) {
      const uint4 v = p[q];
      k[4 * q] = v.x; k[4 * q + 1] = v.y; k[4 * q + 2] = v.z; k[4 * q + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int r = 0; r < ITEMS; ++r) k[r] = i0 + r < n ? keys[i0 + r] : 0u;
  }
  const int cnt_valid = i0 >= n ? 0 : int(min(uint32_t(ITEMS), n - i0));
  uint32_t cur = (k[0] >> shift) & mask, run = 0u;
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    if (r < cnt_valid) {
      const uint32_t d = (k[r] >> shift) & mask;
      if (d != cur) {
        atomicAdd(&h[cur], run);
        cur = d;
        run = 0u;
      }
      ++run;
    }
  }
  if (run) atomicAdd(&h[cur], run);
  __syncthreads();
  for (int d = t; d < bins; d += THREADS) hist[size_t(blockIdx.x) * bins + d] = h[d];
}

// Exclusive scan of each digit column of the tile-major histogram, in place; column totals to
// totals[d]. Workgroup g (1024 threads): digits [DG g, DG g + DG) (thread t: digit t % DG), 1024/DG
// groups of consecutive tiles (group t / DG). Fewer digits per workgroup for more tiles keeps a
// group's tiles within the COL_CACHE registers that hold them between the sum and the scan.
constexpr int COL_THREADS = 1024;
constexpr int COL_CACHE = 16;
template <int DG>
__global__ void __launch_bounds__(COL_THREADS) k_radix_colscan(uint32_t* __restrict__ hist, uint32_t ntiles,
                                                               int bins, uint32_t* __restrict__ totals,
                                                               StepBegin begin, int with_begin) {
  constexpr int PARTS = COL_THREADS / DG;
  if (with_begin && blockIdx.x == gridDim.x - 1) {  // one workgroup beyond the digits runs the step start
    step_begin_block(begin, threadIdx.x, COL_THREADS);
    return;
  }
  __shared__ uint32_t s_part[PARTS][DG];
  const int lane = threadIdx.x % DG, part = threadIdx.x / DG;
  const int d = int(blockIdx.x) * DG + lane;
  const bool on = d < bins;
  const uint32_t q = (ntiles + PARTS - 1) / PARTS;
  const uint32_t b0 = min(ntiles, uint32_t(part) * q), b1 = min(ntiles, b0 + q);
  uint32_t cache[COL_CACHE];
  const bool cached = b1 - b0 <= uint32_t(COL_CACHE);
  uint32_t sum = 0u;
  if (on && cached) {
#pragma unroll
    for (int j = 0; j < COL_CACHE; ++j) {
      cache[j] = b0 + j < b1 ? hist[size_t(b0 + j) * bins + d] : 0u;
      sum += cache[j];
    }
  } else if (on) {
#pragma unroll 8
    for (uint32_t b = b0; b < b1; ++b) sum += hist[size_t(b) * bins + d];
  }
  s_part[part][lane] = sum;
  __syncthreads();
  uint32_t run = 0u, total = 0u;
#pragma unroll 16
  for (int p = 0; p < PARTS; ++p) {
    const uint32_t v = s_part[p][lane];
    run += p < part ? v : 0u;
    total += v;
  }
  if (on && cached) {
#pragma unroll
    for (int j = 0; j < COL_CACHE; ++j) {
      if (b0 + j < b1) hist[size_t(b0 + j) * bins + d] = run;
      run += cache[j];
    }
    if (part == 0) totals[d] = total;
  } else if (on) {
#pragma unroll 8
    for (uint32_t b = b0; b < b1; ++b) {
      const size_t idx = size_t(b) * bins + d;
      const uint32_t v = hist[idx];
      hist[idx] = run;
      run += v;
    }
    if (part == 0) totals[d] = total;
  }
}

// WAVES waves rank a tile, each WKEYS = SORT_TILE / WAVES consecutive keys. The per-wave digit
// counters take WAVES x bins words of dynamic LDS (plus WAVES words of wave sums): 128 KB for
// 16 waves and 11-bit digits (one workgroup per CU), 32 KB for 4 waves (four per CU).
template <int WAVES>
__global__ void __launch_bounds__(64 * WAVES) k_radix_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, uint32_t n, int shift, int bits, uint32_t* __restrict__ offsets,
    const uint32_t* __restrict__ totals, int clear_rows) {
  constexpr int THREADS = 64 * WAVES;
  constexpr int ITEMS = SORT_TILE / THREADS;      // rounds of 64 keys per wave
  constexpr int WKEYS = SORT_TILE / WAVES;
  constexpr int PER = MAX_BINS / THREADS;          // digits per thread for the digit bases
  static_assert(MAX_BINS % THREADS == 0 && SORT_TILE % THREADS == 0, "whole digits and keys per thread");
  extern __shared__ uint32_t s_dyn[];
  const int t = threadIdx.x, lane = t & 63;
  const int wave = __builtin_amdgcn_readfirstlane(t >> 6);
  const int bins = 1 << bits;
  const uint32_t mask = uint32_t(bins - 1);
  uint32_t* s_wsum = s_dyn + WAVES * bins;
  uint32_t* cnt = s_dyn + wave * bins;
  {  // each wave clears its own counters (in-order LDS per wave: no workgroup barrier needed)
    uint4* c4 = reinterpret_cast<uint4*>(cnt);
    for (int i = lane; i < bins / 4; i += 64) c4[i] = make_uint4(0u, 0u, 0u, 0u);
  }
  const uint32_t base = blockIdx.x * uint32_t(SORT_TILE) + uint32_t(wave) * WKEYS + uint32_t(lane);
  uint32_t kreg[ITEMS], vreg[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const uint32_t i = base + uint32_t(r * 64);
    kreg[r] = i < n ? kin[i] : 0u;
    vreg[r] = vin ? (i < n ? vin[i] : 0u) : i;  // no value array: the value is the position
  }
  // thread t owns digits [PER t, PER t + PER): their totals and this tile's row offsets load while
  // the wave ranks its keys
  const int d0 = t * PER;
  uint32_t tot[PER], row[PER];
  uint32_t part = 0u;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const bool on = d0 + j < bins;
    tot[j] = on ? totals[d0 + j] : 0u;
    row[j] = on ? offsets[size_t(blockIdx.x) * bins + d0 + j] : 0u;
    part += tot[j];
  }
  if (clear_rows) {  // this tile's row is read by this workgroup only, and no longer needed
#pragma unroll
    for (int j = 0; j < PER; ++j)
      if (d0 + j < bins) offsets[size_t(blockIdx.x) * bins + d0 + j] = 0u;
  }
  __builtin_amdgcn_wave_barrier();
  // rank within the wave: running per-digit counters, one round of 64 keys at a time
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint32_t loc[ITEMS];
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    const bool valid = base + uint32_t(r * 64) < n;
    const uint32_t d = (kreg[r] >> shift) & mask;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < bits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = __popcll(peers & lt_mask);
    const uint32_t old = valid ? cnt[d] : 0u;
    // one lane per digit group updates the counter; the wave's LDS accesses complete in order
    if (valid && rank == 0u) cnt[d] = old + uint32_t(__popcll(peers));
    loc[r] = old + rank;
  }
  // digit bases: exclusive scan of the totals (within the wave by shuffles, then over the wave
  // sums), plus this tile's row offsets, plus the counts of earlier waves
  uint32_t incl = part;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(incl, off, 64);
    if (lane >= off) incl += y;
  }
  if (lane == 63) s_wsum[wave] = incl;
  __syncthreads();  // wave sums and every wave's counters complete
  uint32_t run = incl - part;
  for (int w = 0; w < wave; ++w) run += s_wsum[w];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int d = d0 + j;
    if (d < bins) {
      uint32_t acc = run + row[j];
      run += tot[j];
#pragma unroll
      for (int w = 0; w < WAVES; ++w) {
        const uint32_t c = s_dyn[w * bins + d];
        s_dyn[w * bins + d] = acc;
        acc += c;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < ITEMS; ++r) {
    if (base + uint32_t(r * 64) < n) {
      const uint32_t d = (kreg[r] >> shift) & mask;
      const uint32_t pos = cnt[d] + loc[r];
      kout[pos] = kreg[r];
      vout[pos] = vreg[r];
    }
  }
}

inline uint32_t div_up(uint64_t a, uint64_t b) { return uint32_t((a + b - 1) / b); }

// Digit width of every pass: the key bits split into the fewest passes of at most digit_bits.
inline int pass_bits(int end_bit, int digit_bits, int* passes_out = nullptr) {
  const int eb = end_bit < 1 ? 1 : (end_bit > 32 ? 32 : end_bit);
  const int db = digit_bits < 4 ? 4 : (digit_bits > MAX_BITS ? MAX_BITS : digit_bits);
  const int passes = (eb + db - 1) / db;
  if (passes_out) *passes_out = passes;
  return (eb + passes - 1) / passes;
}

struct SortLayout {
  uint32_t ntiles;
  size_t hist_off, tot_off, ktmp_off, vtmp_off, total;
};

SortLayout layout(uint32_t n) {
  SortLayout L;
  L.ntiles = div_up(n ? n : 1, SORT_TILE);
  auto al = [](size_t x) { return (x + 255) & ~size_t(255); };
  L.hist_off = 0;
  L.tot_off = al(L.hist_off + sizeof(uint32_t) * MAX_BINS * size_t(L.ntiles));
  L.ktmp_off = al(L.tot_off + sizeof(uint32_t) * MAX_BINS);
  L.vtmp_off = al(L.ktmp_off + sizeof(uint32_t) * (n ? n : 1));
  L.total = al(L.vtmp_off + sizeof(uint32_t) * (n ? n : 1));
  return L;
}
}  // namespace

// Stable sort of (keys_in, vals_in) by the low end_bit bits of the key into (keys_out,
// vals_out), in passes of at most digit_bits (4..11) bits. vals_in == nullptr stands for the
// values 0..n-1 (vals_out is then the sorting permutation). With temp == nullptr, returns the
// temporary storage size in *temp_bytes. Arrays must be 16-byte aligned. With begin != nullptr
// (and n > 0) the first launch also runs the step start *begin.
inline hipError_t radix_sort_pairs(void* temp, size_t* temp_bytes, const uint32_t* keys_in, uint32_t* keys_out,
                                   const uint32_t* vals_in, uint32_t* vals_out, uint32_t n, int end_bit,
                                   int digit_bits, hipStream_t stream, int num_cu,
                                   const StepBegin* begin = nullptr, uint32_t* hist0 = nullptr) {
  const SortLayout L = layout(n);
  if (!temp) {
    *temp_bytes = L.total;
    return hipSuccess;
  }
  if (*temp_bytes < L.total) return hipErrorInvalidValue;
  if (n == 0) return hipSuccess;
  char* base = static_cast<char*>(temp);
  uint32_t* hist = reinterpret_cast<uint32_t*>(base + L.hist_off);
  uint32_t* totals = reinterpret_cast<uint32_t*>(base + L.tot_off);
  uint32_t* ktmp = reinterpret_cast<uint32_t*>(base + L.ktmp_off);
  uint32_t* vtmp = reinterpret_cast<uint32_t*>(base + L.vtmp_off);
  const int eb = end_bit < 1 ? 1 : (end_bit > 32 ? 32 : end_bit);
  int passes = 1;
  const int dbits = pass_bits(end_bit, digit_bits, &passes);
  // 16 ranking waves per tile while all tiles fit one workgroup per CU; beyond that 4 waves per
  // tile, so that four tiles share a CU and no second round of workgroups is needed
  const bool wide = L.ntiles <= uint32_t(num_cu > 0 ? num_cu : 1);
  // choose the first destination so that the last pass writes keys_out
  const uint32_t* kin = keys_in;
  const uint32_t* vin = vals_in;
  for (int p = 0; p < passes; ++p) {
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    uint32_t* ko = to_out ? keys_out : ktmp;
    uint32_t* vo = to_out ? vals_out : vtmp;
    const int shift = p * dbits;
    const int bits = eb - shift < dbits ? eb - shift : dbits;
    const int bins = 1 << bits;
    const bool counted = p == 0 && hist0;  // the first pass's histogram is already in hist0
    uint32_t* h = counted ? hist0 : hist;
    const int begin_here = (p == 0 && begin) ? 1 : 0;
    const StepBegin sb = begin_here ? *begin : StepBegin{};
    // the step start rides on the pass's first launch: workgroup 0 of the histogram, or an extra
    // last workgroup of the column scan
    const int begin_hist = counted ? 0 : begin_here, begin_scan = counted ? begin_here : 0;
    if (!counted) {
      if (wide)
        k_radix_hist<1024><<<L.ntiles, 1024, 0, stream>>>(kin, n, shift, bits, hist, sb, begin_hist);
      else
        k_radix_hist<256><<<L.ntiles, 256, 0, stream>>>(kin, n, shift, bits, hist, sb, begin_hist);
    }
    if (L.ntiles <= uint32_t(16 * COL_CACHE))
      k_radix_colscan<64><<<div_up(bins, 64) + begin_scan, COL_THREADS, 0, stream>>>(h, L.ntiles, bins, totals, sb,
                                                                                     begin_scan);
    else if (L.ntiles <= uint32_t(32 * COL_CACHE))
      k_radix_colscan<32><<<div_up(bins, 32) + begin_scan, COL_THREADS, 0, stream>>>(h, L.ntiles, bins, totals, sb,
                                                                                     begin_scan);
    else
      k_radix_colscan<16><<<div_up(bins, 16) + begin_scan, COL_THREADS, 0, stream>>>(h, L.ntiles, bins, totals, sb,
                                                                                     begin_scan);
    const int clear = counted ? 1 : 0;
    if (wide)
      k_radix_scatter<16><<<L.ntiles, 64 * 16, (16 * bins + 16) * sizeof(uint32_t), stream>>>(
          kin, vin, ko, vo, n, shift, bits, h, totals, clear);
    else
      k_radix_scatter<4><<<L.ntiles, 64 * 4, (4 * bins + 4) * sizeof(uint32_t), stream>>>(
          kin, vin, ko, vo, n, shift, bits, h, totals, clear);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    kin = ko;
    vin = vo;
  }
  return hipSuccess;
}

// Words of a first-pass histogram for n keys: one row of up to 2^MAX_BITS digits per tile.
inline size_t hist0_words(uint32_t n) { return size_t(layout(n).ntiles) * MAX_BINS; }

// The first pass's histogram of keys counted into hist0 (rows overwritten), for a sort whose keys
// were not counted by their writer.
inline hipError_t radix_count_first_pass(const uint32_t* keys, uint32_t n, int end_bit, int digit_bits,
                                         uint32_t* hist0, hipStream_t stream, int num_cu) {
  if (n == 0) return hipSuccess;
  const SortLayout L = layout(n);
  const int eb = end_bit < 1 ? 1 : (end_bit > 32 ? 32 : end_bit);
  const int dbits = pass_bits(end_bit, digit_bits);
  const int bits = eb < dbits ? eb : dbits;
  if (L.ntiles <= uint32_t(num_cu > 0 ? num_cu : 1))
    k_radix_hist<1024><<<L.ntiles, 1024, 0, stream>>>(keys, n, 0, bits, hist0, StepBegin{}, 0);
  else
    k_radix_hist<256><<<L.ntiles, 256, 0, stream>>>(keys, n, 0, bits, hist0, StepBegin{}, 0);
  return hipGetLastError();
}

}  // namespace snow
