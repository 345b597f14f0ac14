// axsq_prims.hip -- scan and stable radix sort for gfx950 (wave64).
//
// Radix sort: least-significant-digit first, 8-bit digits; per pass a tile
// histogram, an exclusive scan of the digit-major histogram and a stable
// scatter. With at most 256 tiles the scatter does the scan itself (two
// kernels per pass instead of five), and inputs of at most 262,144 keys use
// tiles of 1,024 keys instead of 4,096. Inside a tile of 256 threads x ITEMS
// keys, keys are ranked in global order: sub-tile k covers
// [base + 256k, base + 256k + 256) and each 64-lane wave finds the lanes that
// share its digit with 8 ballots, so equal keys keep their input order and the
// result equals a stable sort.
#include <stdarg.h>
#include <stdio.h>

#include "axsq_internal.hpp"

namespace axsq {

namespace {

constexpr int kWave = 64;
constexpr int kWavesPerBlock = kBlock / kWave;

__device__ __forceinline__ uint64_t lanemask_lt() {
  const unsigned lane = __lane_id();
  return (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
}

// Inclusive scan across one wave with shuffles.
template <class T>
__device__ __forceinline__ T wave_inclusive(T v) {
  const int lane = __lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    T y = __shfl_up(v, off, kWave);
    if (lane >= off) v += y;
  }
  return v;
}

// Exclusive scan of one value per thread over a 256-thread block.
template <class T>
__device__ __forceinline__ T block_exclusive(T v, T* s_wave, T* total) {
  const int lane = __lane_id();
  const int wid = threadIdx.x / kWave;
  T inc = wave_inclusive(v);
  if (lane == kWave - 1) s_wave[wid] = inc;
  __syncthreads();
  T base = 0, sum = 0;
#pragma unroll
  for (int w = 0; w < kWavesPerBlock; ++w) {
    const T x = s_wave[w];
    if (w < wid) base += x;
    sum += x;
  }
  if (total) *total = sum;
  __syncthreads();
  return base + inc - v;
}

__global__ void k_fill_i32(int32_t* p, int32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---- scan -------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(kBlock) void k_tile_sums(const T* __restrict__ in, int64_t n,
                                                       int64_t* __restrict__ sums) {
  __shared__ int64_t s_wave[kWavesPerBlock];
  const int64_t base = (int64_t)blockIdx.x * kTile;
  int64_t acc = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int64_t i = base + k * kBlock + threadIdx.x;
    if (i < n) acc += (int64_t)in[i];
  }
  int64_t tot;
  block_exclusive<int64_t>(acc, s_wave, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// One block scans the tile sums (in place, exclusive) and writes the total.
__global__ __launch_bounds__(1024) void k_scan_tile_sums(int64_t* __restrict__ sums,
                                                          int64_t ntiles,
                                                          int64_t* __restrict__ total) {
  __shared__ int64_t s_part[1024];
  const int t = threadIdx.x;
  const int64_t per = (ntiles + 1023) / 1024;
  const int64_t b = t * per;
  const int64_t e = (b + per < ntiles) ? b + per : ntiles;
  int64_t acc = 0;
  for (int64_t i = b; i < e; ++i) acc += sums[i];
  s_part[t] = acc;
  __syncthreads();
  // Hillis-Steele over 1024 partials
  for (int off = 1; off < 1024; off <<= 1) {
    int64_t y = (t >= off) ? s_part[t - off] : 0;
    __syncthreads();
    s_part[t] += y;
    __syncthreads();
  }
  int64_t run = s_part[t] - acc;
  for (int64_t i = b; i < e; ++i) {
    const int64_t x = sums[i];
    sums[i] = run;
    run += x;
  }
  if (t == 1023 && total) *total = s_part[1023];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_scan_tiles(const T* __restrict__ in,
                                                        T* __restrict__ out, int64_t n,
                                                        const int64_t* __restrict__ sums) {
  __shared__ T s_tile[kTile];
  __shared__ T s_wave[kWavesPerBlock];
  const int64_t base = (int64_t)blockIdx.x * kTile;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int64_t i = base + k * kBlock + threadIdx.x;
    s_tile[k * kBlock + threadIdx.x] = (i < n) ? in[i] : T(0);
  }
  __syncthreads();
  T v[kItems];
  T acc = 0;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    v[k] = s_tile[threadIdx.x * kItems + k];
    acc += v[k];
  }
  T excl = block_exclusive<T>(acc, s_wave, nullptr);
  const T tile_base = (T)sums[blockIdx.x];
  T run = tile_base + excl;
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    s_tile[threadIdx.x * kItems + k] = run;
    run += v[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kItems; ++k) {
    const int64_t i = base + k * kBlock + threadIdx.x;
    if (i < n) out[i] = s_tile[k * kBlock + threadIdx.x];
  }
}

// ---- radix sort -------------------------------------------------------------
// ITEMS keys per thread make a tile of kBlock * ITEMS keys per workgroup. The
// histogram is digit-major (hist[digit * nblocks + tile], the order the scan
// needs) or, for a scatter that sums it itself, tile-major
// (hist[tile * 256 + digit], so that those sums read it coalesced).
template <int ITEMS, bool TILE_MAJOR>
__global__ __launch_bounds__(kBlock) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                        int64_t n, int shift, int nbits,
                                                        uint32_t* __restrict__ hist,
                                                        int64_t nblocks) {
  __shared__ uint32_t s_h[256];
  s_h[threadIdx.x] = 0;
  __syncthreads();
  const uint32_t mask = (1u << nbits) - 1u;
  const int64_t base = (int64_t)blockIdx.x * (kBlock * ITEMS);
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int64_t i = base + k * kBlock + threadIdx.x;
    if (i < n) atomicAdd(&s_h[(keys[i] >> shift) & mask], 1u);
  }
  __syncthreads();
  if (TILE_MAJOR)
    hist[(int64_t)blockIdx.x * 256 + threadIdx.x] = s_h[threadIdx.x];
  else
    hist[(int64_t)threadIdx.x * nblocks + blockIdx.x] = s_h[threadIdx.x];
}

// Stable scatter of one digit. The output position of digit d's first key in
// this tile is either read from the exclusively scanned digit-major
// histogram (SCANNED) or, when there are few tiles, found here from the raw
// tile-major histogram: the keys of smaller digits everywhere plus the keys of
// digit d in earlier tiles (16 loads in flight per thread). All of the tile's keys are loaded before the ranking loop so
// that their latencies overlap.
template <int ITEMS, bool SCANNED>
__global__ __launch_bounds__(kBlock) void k_radix_scatter(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n, int shift,
    int nbits, const uint32_t* __restrict__ hist, int64_t nblocks,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
    uint32_t* __restrict__ next_hist = nullptr, int next_shift = 0, int next_nbits = 0) {
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_wcnt[kWavesPerBlock][256];
  __shared__ uint32_t s_w[kWavesPerBlock];
  const int t = threadIdx.x;
  const int wid = t / kWave;
  const uint64_t lt = lanemask_lt();
  const uint32_t mask = (1u << nbits) - 1u;
  const int64_t base = (int64_t)blockIdx.x * (kBlock * ITEMS);
  uint32_t kk[ITEMS], vv[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int64_t i = base + k * kBlock + t;
    kk[k] = (i < n) ? keys[i] : 0u;
    vv[k] = (i < n) ? vals[i] : 0u;
  }
  if constexpr (SCANNED) {
    s_base[t] = hist[(int64_t)t * nblocks + blockIdx.x];
  } else {
    const uint32_t* h = hist + t;
    const int64_t me = blockIdx.x;
    uint32_t pre = 0, tot = 0;
    int64_t b = 0;
    for (; b + 16 <= nblocks; b += 16) {
      uint32_t x[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) x[u] = h[(b + u) * 256];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        tot += x[u];
        pre += (b + u < me) ? x[u] : 0u;
      }
    }
    for (; b < nblocks; ++b) {
      const uint32_t x = h[b * 256];
      tot += x;
      pre += (b < me) ? x : 0u;
    }
    s_base[t] = block_exclusive<uint32_t>(tot, s_w, nullptr) + pre;
  }
#pragma unroll
  for (int w = 0; w < kWavesPerBlock; ++w) s_wcnt[w][t] = 0;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < ITEMS; ++k) {
    const int64_t i = base + k * kBlock + t;
    const bool valid = i < n;
    const uint32_t key = kk[k], val = vv[k];
    const uint32_t d = valid ? (key >> shift) & mask : 0u;
    uint64_t peers = __ballot(valid);
    for (int b = 0; b < nbits; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(valid && bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t rank = (uint32_t)__popcll(peers & lt);
    const bool leader = valid && ((peers & lt) == 0);
    if (leader) s_wcnt[wid][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (valid) {
      uint32_t pos = s_base[d] + rank;
      for (int w = 0; w < wid; ++w) pos += s_wcnt[w][d];
      keys_out[pos] = key;
      vals_out[pos] = val;
      if (next_hist)  // the next pass's tile histogram, tiles of kBlock * ITEMS keys
        atomicAdd(&next_hist[(int64_t)(pos / (kBlock * ITEMS)) * 256 +
                             ((key >> next_shift) & ((1u << next_nbits) - 1u))],
                  1u);
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < kWavesPerBlock; ++w) {
      add += s_wcnt[w][t];
      s_wcnt[w][t] = 0;
    }
    s_base[t] += add;
    __syncthreads();
  }
}

// The build's small sort (at most kSmallSortMax keys, tiles of
// kSmallSortTile = 1024 keys, one 1024-thread workgroup each, one key per
// thread): a stable scatter of one digit. hist is the pass's tile-major
// histogram (hist[tile * 256 + d], nb tiles): where digit d starts in this
// tile is the keys of smaller digits everywhere plus those of digit d in
// earlier tiles, summed by four threads per digit over a quarter of the
// tiles each (coalesced across digits, 32 loads in flight per thread). Each
// pass adds the next pass's counts to next_hist as it writes. The last pass
// (LAST) writes the sorted keys to keys_out and, at each sorted position p,
// vals_out[p] = the key's value and the six doubles of row vals_out[p] of
// `rows` to rows_out[p] (the leaves in tree order).
template <bool LAST>
__global__ __launch_bounds__(1024) void k_small_scatter(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n, int shift,
    int nbits, const uint32_t* __restrict__ hist, int32_t nb,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
    uint32_t* __restrict__ next_hist, int next_shift, int next_nbits,
    const double* __restrict__ rows, double* __restrict__ rows_out) {
  static_assert(kSmallSortTile == 1024, "one key per thread");
  constexpr int kW = 1024 / kWave;
  __shared__ uint32_t s_part[2][4][256];  // per quarter of the tiles: all, before this tile
  __shared__ uint32_t s_base[256];
  __shared__ uint32_t s_wsum[4];
  __shared__ uint32_t s_wcnt[kW][256];
  const int t = threadIdx.x;
  const int wid = t / kWave;
  const int64_t i = (int64_t)blockIdx.x * 1024 + t;
  const bool valid = i < n;
  const uint32_t key = valid ? keys[i] : 0u, val = valid ? vals[i] : 0u;
  double2 rw[3];
  if constexpr (LAST) {
    const double2* src = reinterpret_cast<const double2*>(rows + 6 * (int64_t)val);
    for (int e = 0; e < 3; ++e) rw[e] = valid ? src[e] : make_double2(0.0, 0.0);
  }
  {
    const int q = t / 256, d = t % 256;
    const int32_t per = (nb + 3) / 4;  // at most 64: one or two rounds of 32 loads
    const int32_t b0 = q * per, me = (int32_t)blockIdx.x;
    uint32_t tot = 0, pre = 0;
    for (int32_t u0 = 0; u0 < per; u0 += 32) {
      uint32_t x[32];
#pragma unroll
      for (int u = 0; u < 32; ++u) {
        const int32_t b = b0 + u0 + u;
        x[u] = (u0 + u < per && b < nb) ? hist[(int64_t)b * 256 + d] : 0u;
      }
#pragma unroll
      for (int u = 0; u < 32; ++u) {
        tot += x[u];
        pre += (b0 + u0 + u < me) ? x[u] : 0u;
      }
    }
    s_part[0][q][d] = tot;
    s_part[1][q][d] = pre;
  }
  for (int j = t; j < kW * 256; j += 1024) (&s_wcnt[0][0])[j] = 0u;
  __syncthreads();
  uint32_t tot = 0, pre = 0, inc = 0;
  if (t < 256) {
    for (int q = 0; q < 4; ++q) {
      tot += s_part[0][q][t];
      pre += s_part[1][q][t];
    }
    inc = wave_inclusive(tot);
    if (__lane_id() == kWave - 1) s_wsum[wid] = inc;
  }
  __syncthreads();
  if (t < 256) {
    uint32_t b = 0;
    for (int w = 0; w < wid; ++w) b += s_wsum[w];
    s_base[t] = b + inc - tot + pre;
  }
  const uint32_t mask = (1u << nbits) - 1u;
  const uint32_t d = valid ? (key >> shift) & mask : 0u;
  uint64_t peers = __ballot(valid);
  for (int b = 0; b < nbits; ++b) {
    const bool bit = (d >> b) & 1u;
    const uint64_t bb = __ballot(valid && bit);
    peers &= bit ? bb : ~bb;
  }
  const uint64_t lt = lanemask_lt();
  const uint32_t rank = (uint32_t)__popcll(peers & lt);
  if (valid && (peers & lt) == 0) s_wcnt[wid][d] = (uint32_t)__popcll(peers);
  __syncthreads();
  if (valid) {
    uint32_t pos = s_base[d] + rank;
    for (int w = 0; w < wid; ++w) pos += s_wcnt[w][d];
    keys_out[pos] = key;
    vals_out[pos] = val;
    if constexpr (LAST) {
      double2* dst = reinterpret_cast<double2*>(rows_out + 6 * (int64_t)pos);
      for (int e = 0; e < 3; ++e) dst[e] = rw[e];
    } else {
      atomicAdd(&next_hist[(int64_t)(pos / kSmallSortTile) * 256 +
                           ((key >> next_shift) & ((1u << next_nbits) - 1u))],
                1u);
    }
  }
}

}  // namespace

int radix_sort_small_build(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt,
                           uint32_t* vals_alt, int64_t n, int bits, uint32_t* hists,
                           const double* rows, double* rows_out, uint32_t* keys_last,
                           int32_t* vals_last, hipStream_t s) {
  if (n < 1 || n > kSmallSortMax || bits <= 8 || bits > 32) {
    set_error("radix sort: %lld keys of %d bits do not fit the small sort", (long long)n, bits);
    return AXSQ_ERR_INVALID_ARG;
  }
  const int32_t nb = (int32_t)((n + kSmallSortTile - 1) / kSmallSortTile);
  const int32_t hs = small_sort_stride(n);
  uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
  int pass = 0;
  for (int shift = 0; shift < bits; shift += 8, ++pass) {
    const int nbits = (bits - shift < 8) ? bits - shift : 8;
    const uint32_t* h = hists + (int64_t)pass * 256 * hs;
    if (shift + 8 < bits) {
      const int next = shift + 8;
      k_small_scatter<false><<<(unsigned)nb, 1024, 0, s>>>(
          ki, vi, n, shift, nbits, h, nb, ko, vo, hists + (int64_t)(pass + 1) * 256 * hs, next,
          (bits - next < 8) ? bits - next : 8, nullptr, nullptr);
    } else {
      k_small_scatter<true><<<(unsigned)nb, 1024, 0, s>>>(
          ki, vi, n, shift, nbits, h, nb, keys_last, reinterpret_cast<uint32_t*>(vals_last),
          nullptr, 0, 0, rows, rows_out);
    }
    AXSQ_CHECK_HIP(hipGetLastError());
    uint32_t* t = ki;
    ki = ko;
    ko = t;
    t = vi;
    vi = vo;
    vo = t;
  }
  return AXSQ_OK;
}

int fill_i32(int32_t* p, int32_t v, int64_t n, hipStream_t s) {
  if (n <= 0) return AXSQ_OK;
  k_fill_i32<<<grid_for(n), kBlock, 0, s>>>(p, v, n);
  AXSQ_CHECK_HIP(hipGetLastError());
  return AXSQ_OK;
}

int scan_exclusive_i32(const int32_t* in, int32_t* out, int64_t n, int64_t* total_dev,
                       Workspace& ws, hipStream_t s) {
  AXSQ_TRY(ws.tiles.ensure(sizeof(int64_t) * (size_t)scan_tiles(n)));
  return scan_exclusive_i32_into(in, out, n, ws.tiles.as<int64_t>(), total_dev, s);
}

int scan_exclusive_i32_into(const int32_t* in, int32_t* out, int64_t n, int64_t* sums,
                            int64_t* total_dev, hipStream_t s) {
  const int64_t ntiles = (n + kTile - 1) / kTile;
  if (n > 0) {
    k_tile_sums<int32_t><<<(unsigned)ntiles, kBlock, 0, s>>>(in, n, sums);
    AXSQ_CHECK_HIP(hipGetLastError());
  }
  k_scan_tile_sums<<<1, 1024, 0, s>>>(sums, ntiles, total_dev);
  AXSQ_CHECK_HIP(hipGetLastError());
  if (n > 0) {
    k_scan_tiles<int32_t><<<(unsigned)ntiles, kBlock, 0, s>>>(in, out, n, sums);
    AXSQ_CHECK_HIP(hipGetLastError());
  }
  return AXSQ_OK;
}

int radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt,
                     int64_t n, int bits, Workspace& ws, hipStream_t s, uint32_t** keys_out,
                     uint32_t** vals_out) {
  *keys_out = keys;
  *vals_out = vals;
  if (n <= 1) return AXSQ_OK;
  if (n > (int64_t)0x7fffffffll) {
    set_error("radix sort: %lld keys exceed the int32 position range", (long long)n);
    return AXSQ_ERR_INVALID_ARG;
  }
  // Small inputs take 1,024-key tiles, so that more workgroups share the work;
  // with at most kFewTiles tiles the scatter scans the histogram itself.
  constexpr int64_t kFewTiles = 256;
  const bool small = n <= kFewTiles * kBlock * 4;
  const int64_t tile = small ? kBlock * 4 : kTile;
  const int64_t nb = (n + tile - 1) / tile;
  const bool few = nb <= kFewTiles;
  const int64_t nh = 256 * nb;
  AXSQ_TRY(ws.hist.ensure(sizeof(uint32_t) * (size_t)nh));
  uint32_t* hist = ws.hist.as<uint32_t>();
  AXSQ_TRY(ws.misc.ensure(sizeof(int64_t) * 2));
  uint32_t *ki = keys, *vi = vals, *ko = keys_alt, *vo = vals_alt;
  for (int shift = 0; shift < bits; shift += 8) {
    const int nbits = (bits - shift < 8) ? bits - shift : 8;
    if (small) {
      k_radix_hist<4, true><<<(unsigned)nb, kBlock, 0, s>>>(ki, n, shift, nbits, hist, nb);
      AXSQ_CHECK_HIP(hipGetLastError());
      k_radix_scatter<4, false><<<(unsigned)nb, kBlock, 0, s>>>(ki, vi, n, shift, nbits, hist,
                                                                  nb, ko, vo);
    } else {
      if (few)
        k_radix_hist<kItems, true><<<(unsigned)nb, kBlock, 0, s>>>(ki, n, shift, nbits, hist, nb);
      else
        k_radix_hist<kItems, false><<<(unsigned)nb, kBlock, 0, s>>>(ki, n, shift, nbits, hist, nb);
      AXSQ_CHECK_HIP(hipGetLastError());
      if (few) {
        k_radix_scatter<kItems, false><<<(unsigned)nb, kBlock, 0, s>>>(ki, vi, n, shift, nbits,
                                                                       hist, nb, ko, vo);
      } else {
        AXSQ_TRY(scan_exclusive_i32((const int32_t*)hist, (int32_t*)hist, nh,
                                    ws.misc.as<int64_t>(), ws, s));
        k_radix_scatter<kItems, true><<<(unsigned)nb, kBlock, 0, s>>>(ki, vi, n, shift, nbits,
                                                                      hist, nb, ko, vo);
      }
    }
    AXSQ_CHECK_HIP(hipGetLastError());
    uint32_t* t = ki;
    ki = ko;
    ko = t;
    t = vi;
    vi = vo;
    vo = t;
  }
  *keys_out = ki;
  *vals_out = vi;
  return AXSQ_OK;
}

}  // namespace axsq
