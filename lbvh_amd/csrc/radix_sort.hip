// Onesweep-style LSD radix sort of Morton codes (and primitive ids) for gfx950.
//
// The reference orders its curve with an unstable std::sort on the code
// (lbvh.h:1126-1134).  This sort is stable and keyed on the code only, so
// ids stay in ascending order within equal codes: the "(code, id)" order,
// one legal output of the reference's sort.
//
// One kernel per digit pass; the digit histograms of all passes come from
// the encode kernel.  Per workgroup (4 waves, a tile of 256 * ITEMS keys):
//   1. take a virtual tile id from an atomic ticket (tiles are numbered in
//      the order workgroups start, so every tile a workgroup waits on is
//      already running);
//   2. rank keys stably inside each wave with ballot-based digit matching
//      (wave-striped tile order: item j of lane l is key j * 64 + l);
//   3. publish the tile's digit counts, then resolve each digit's global
//      offset by decoupled look-back over earlier tiles' epoch-tagged status
//      words (windowed: several predecessors are read per step);
//   4. stage the tile in LDS in digit order and write it out so that
//      consecutive threads write consecutive addresses.
// Every wait is bounded: past a spin limit the kernel raises an error flag
// and drains, and the host reports a timeout instead of hanging.
//
// A second strategy, reduce-then-scan, has no inter-workgroup waits: an
// upsweep counts each chunk's digits, a scan turns the counts into per-chunk
// digit offsets, and a downsweep ranks and scatters each chunk tile by tile
// with running offsets in LDS.  It reads the keys once more per pass but
// avoids the look-back latency; the tuner picks between the two.
#include "device_common.hpp"
#include "engine.hpp"

namespace lbvh_dev {

constexpr int kSortBlock = 256;
constexpr int kSortWaves = kSortBlock / kWave;
constexpr int kLookbackWindow = 8;
constexpr uint32_t kSpinLimit = 1u << 20;

constexpr uint64_t kFlagAgg = 1;
constexpr uint64_t kFlagInc = 2;

__device__ __forceinline__ uint64_t
pack_status(uint32_t epoch, uint64_t flag, uint32_t count)
{
  return (uint64_t(epoch & 0xffffffu) << 40) | (flag << 32) | uint64_t(count);
}
__device__ __forceinline__ bool
status_ready(uint64_t s, uint32_t epoch)
{
  return uint32_t(s >> 40) == (epoch & 0xffffffu) && ((s >> 32) & 0xff) != 0;
}

// Exclusive scan of one value per thread across the workgroup.
__device__ __forceinline__ uint32_t
block_exclusive_scan(uint32_t x, uint32_t* s_tmp, uint32_t* total)
{
  const int lane = __lane_id();
  const int w = threadIdx.x / kWave;
  uint32_t incl = x;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t y = __shfl_up(incl, off, kWave);
    if (lane >= off)
      incl += y;
  }
  if (lane == kWave - 1)
    s_tmp[w] = incl;
  __syncthreads();
  uint32_t woff = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kSortWaves; i++) {
    uint32_t v = s_tmp[i];
    woff += (i < w) ? v : 0u;
    tot += v;
  }
  __syncthreads();
  *total = tot;
  return incl - x + woff;
}

template<typename KeyT, bool VALS, int RBITS, int ITEMS>
__global__ __launch_bounds__(kSortBlock) void
k_onesweep(const KeyT* __restrict__ kin, KeyT* __restrict__ kout,
           const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout, uint32_t n,
           int shift, const uint32_t* __restrict__ ghist, uint64_t* __restrict__ status,
           uint32_t* __restrict__ ticket, uint32_t epoch, uint32_t* __restrict__ err)
{
  constexpr int RADIX = 1 << RBITS;
  constexpr int TILE = kSortBlock * ITEMS;
  constexpr int DPT = RADIX / kSortBlock; // digits owned per thread
  static_assert(DPT >= 1, "radix must be >= workgroup size");

  __shared__ KeyT s_keys[TILE];
  __shared__ uint32_t s_vals[VALS ? TILE : 1];
  __shared__ uint32_t s_wcnt[kSortWaves][RADIX];
  __shared__ uint32_t s_bstart[RADIX];
  __shared__ uint32_t s_goff[RADIX];
  __shared__ uint32_t s_tmp[kSortWaves];
  __shared__ uint32_t s_bid;

  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int w = tid / kWave;
  const KeyT dmask = KeyT(RADIX - 1);

  if (tid == 0)
    s_bid = atomic_add_agent(ticket, 1u);
  for (int j = tid; j < kSortWaves * RADIX; j += kSortBlock)
    (&s_wcnt[0][0])[j] = 0u;
  __syncthreads();
  const uint32_t bid = s_bid;

  // ---- load (wave-striped) ----
  const uint32_t base = bid * uint32_t(TILE) + uint32_t(w) * (ITEMS * kWave) + uint32_t(lane);
  KeyT key[ITEMS];
  uint32_t val[ITEMS];
  uint32_t rank[ITEMS];
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    uint32_t idx = base + uint32_t(j) * kWave;
    bool valid = idx < n;
    key[j] = valid ? kin[idx] : KeyT(0);
    if (VALS)
      val[j] = valid ? (vin ? vin[idx] : idx) : 0u;
  }

  // ---- stable ranking within the wave ----
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    const uint32_t idx = base + uint32_t(j) * kWave;
    const bool valid = idx < n;
    const uint32_t d = uint32_t((key[j] >> shift) & dmask);
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < RBITS; b++) {
      const bool bit = (d >> b) & 1u;
      const uint64_t m = ballot(valid && bit);
      peers &= bit ? m : ~m;
    }
    uint32_t r = 0;
    if (valid) {
      const uint32_t below = uint32_t(popc64(peers & lt));
      const uint32_t cnt = uint32_t(popc64(peers));
      const uint32_t wbase = s_wcnt[w][d];
      r = wbase + below;
      if (below == 0)
        s_wcnt[w][d] = wbase + cnt;
    }
    rank[j] = r;
  }
  __syncthreads();

  // ---- tile digit counts, wave prefixes, early publication ----
  uint32_t tot[DPT];
  uint32_t my_sum = 0;
#pragma unroll
  for (int dd = 0; dd < DPT; dd++) {
    const int d = tid * DPT + dd;
    uint32_t acc = 0;
#pragma unroll
    for (int ww = 0; ww < kSortWaves; ww++) {
      uint32_t c = s_wcnt[ww][d];
      s_wcnt[ww][d] = acc;
      acc += c;
    }
    tot[dd] = acc;
    my_sum += acc;
    coherent_store_u64(&status[uint64_t(bid) * RADIX + d],
                       pack_status(epoch, bid == 0 ? kFlagInc : kFlagAgg, acc));
  }
  // block-exclusive digit starts
  uint32_t tile_total;
  uint32_t run = block_exclusive_scan(my_sum, s_tmp, &tile_total);
#pragma unroll
  for (int dd = 0; dd < DPT; dd++) {
    s_bstart[tid * DPT + dd] = run;
    run += tot[dd];
  }
  // global digit starts of this pass
  uint32_t gh[DPT];
  uint32_t gsum = 0;
#pragma unroll
  for (int dd = 0; dd < DPT; dd++) {
    gh[dd] = ghist[tid * DPT + dd];
    gsum += gh[dd];
  }
  uint32_t gtotal;
  uint32_t grun = block_exclusive_scan(gsum, s_tmp, &gtotal);
  __syncthreads(); // s_bstart and the wave prefixes are visible

  // ---- stage the tile in LDS in digit order ----
#pragma unroll
  for (int j = 0; j < ITEMS; j++) {
    const uint32_t idx = base + uint32_t(j) * kWave;
    if (idx < n) {
      const uint32_t d = uint32_t((key[j] >> shift) & dmask);
      const uint32_t pos = s_bstart[d] + s_wcnt[w][d] + rank[j];
      s_keys[pos] = key[j];
      if (VALS)
        s_vals[pos] = val[j];
    }
  }

  // ---- decoupled look-back per owned digit ----
#pragma unroll
  for (int dd = 0; dd < DPT; dd++) {
    const int d = tid * DPT + dd;
    uint32_t excl = 0;
    if (bid > 0) {
      int64_t p = int64_t(bid) - 1;
      uint32_t spins = 0;
      bool done = false;
      while (!done) {
        uint64_t sv[kLookbackWindow];
#pragma unroll
        for (int q = 0; q < kLookbackWindow; q++) {
          int64_t pq = p - q;
          sv[q] = (pq >= 0) ? coherent_load_u64(&status[uint64_t(pq) * RADIX + d])
                            : pack_status(epoch, kFlagInc, 0);
        }
        int consumed = 0;
#pragma unroll
        for (int q = 0; q < kLookbackWindow; q++) {
          if (done || consumed != q)
            continue;
          const uint64_t s = sv[q];
          if (!status_ready(s, epoch))
            continue;
          excl += uint32_t(s);
          consumed = q + 1;
          if (((s >> 32) & 0xff) == kFlagInc)
            done = true;
        }
        if (done)
          break;
        p -= consumed;
        if (consumed == 0) {
          if (++spins > kSpinLimit || coherent_load_u32(err) != 0u) {
            atomic_add_agent(err, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
    }
    coherent_store_u64(&status[uint64_t(bid) * RADIX + d],
                       pack_status(epoch, kFlagInc, excl + tot[dd]));
    s_goff[d] = grun + excl - s_bstart[d];
    grun += gh[dd];
  }
  __syncthreads();

  // ---- write the tile out ----
  const uint32_t tile_begin = bid * uint32_t(TILE);
  const uint32_t tile_n = (n - tile_begin < uint32_t(TILE)) ? (n - tile_begin) : uint32_t(TILE);
  for (uint32_t i = tid; i < tile_n; i += kSortBlock) {
    const KeyT k = s_keys[i];
    const uint32_t d = uint32_t((k >> shift) & dmask);
    const uint32_t dst = s_goff[d] + i;
    if (dst < n) { // always true unless a bounded wait expired (then err is set)
      kout[dst] = k;
      if (VALS)
        vout[dst] = s_vals[i];
    }
  }
}

// ---- reduce-then-scan --------------------------------------------------------

constexpr int kUpsweepUnroll = 8;

// Key range [*begin, *end) of chunk g: fixed chunks of chunk_keys keys, or
// the scene-aligned chunk table (begin, end pairs) when chunks != nullptr.
__device__ __forceinline__ void
chunk_range(uint32_t g, uint32_t n, uint64_t chunk_keys, const uint32_t* __restrict__ chunks,
            uint32_t* begin, uint32_t* end)
{
  if (chunks) {
    *begin = chunks[2 * g];
    *end = chunks[2 * g + 1];
    return;
  }
  const uint64_t b64 = uint64_t(g) * chunk_keys;
  const uint64_t e64 = b64 + chunk_keys;
  *begin = b64 < n ? uint32_t(b64) : n;
  *end = e64 < n ? uint32_t(e64) : n;
}

// Number of keys of chunk g whose digit is d: counts[d * groups + g], or
// counts[g * RADIX + d] (chunk-major) with a chunk table.
template<typename KeyT, int RBITS>
__global__ __launch_bounds__(kSortBlock) void
k_rts_upsweep(const KeyT* __restrict__ keys, uint32_t n, int shift, uint32_t chunk_keys,
              uint32_t groups, uint32_t* __restrict__ counts,
              const uint32_t* __restrict__ chunks)
{
  constexpr int RADIX = 1 << RBITS;
  __shared__ uint32_t s_h[kSortWaves][RADIX];
  const int tid = threadIdx.x;
  const int w = tid / kWave;
  for (int j = tid; j < kSortWaves * RADIX; j += kSortBlock)
    (&s_h[0][0])[j] = 0u;
  __syncthreads();
  const uint32_t g = blockIdx.x;
  uint32_t begin, end;
  chunk_range(g, n, chunk_keys, chunks, &begin, &end);
  const KeyT dmask = KeyT(RADIX - 1);
  const uint64_t lt = lanemask_lt();
  for (uint32_t base = begin; base < end; base += kSortBlock * kUpsweepUnroll) {
    KeyT k[kUpsweepUnroll];
#pragma unroll
    for (int u = 0; u < kUpsweepUnroll; u++) {
      uint32_t i = base + u * kSortBlock + tid;
      k[u] = (i < end) ? keys[i] : KeyT(0);
    }
#pragma unroll
    for (int u = 0; u < kUpsweepUnroll; u++) {
      const uint32_t i = base + u * kSortBlock + tid;
      const bool valid = i < end;
      const uint32_t d = uint32_t((k[u] >> shift) & dmask);
      uint64_t peers = ballot(valid);
#pragma unroll
      for (int b = 0; b < RBITS; b++) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = ballot(valid && bit);
        peers &= bit ? m : ~m;
      }
      if (valid && (peers & lt) == 0)
        s_h[w][d] += uint32_t(popc64(peers));
    }
  }
  __syncthreads();
  for (int d = tid; d < RADIX; d += kSortBlock) {
    uint32_t c = 0;
#pragma unroll
    for (int ww = 0; ww < kSortWaves; ww++)
      c += s_h[ww][d];
    counts[chunks ? size_t(g) * RADIX + d : size_t(d) * groups + g] = c;
  }
}

// One workgroup per digit d: counts[d][g] <- (keys with a digit below d) +
// (keys of digit d in chunks before g).
template<int RBITS>
__global__ __launch_bounds__(kSortBlock) void
k_rts_scan(uint32_t* __restrict__ counts, uint32_t groups, const uint32_t* __restrict__ ghist,
           uint32_t* __restrict__ zero_word)
{
  constexpr int RADIX = 1 << RBITS;
  __shared__ uint32_t s_tmp[kSortWaves];
  const int d = blockIdx.x;
  const int tid = threadIdx.x;
  if (zero_word && d == 0 && tid == 0)
    *zero_word = 0u; // a counter of a later kernel of the same sort
  uint32_t below = 0;
  for (int dd = tid; dd < d; dd += kSortBlock)
    below += ghist[dd];
  uint32_t total;
  block_exclusive_scan(below, s_tmp, &total); // total = keys with a smaller digit
  const uint32_t base_d = total;
  uint32_t* row = counts + size_t(d) * groups;
  const uint32_t per = (groups + kSortBlock - 1) / kSortBlock;
  const uint32_t g0 = tid * per;
  uint32_t sum = 0;
  for (uint32_t g = g0; g < g0 + per && g < groups; g++)
    sum += row[g];
  uint32_t all;
  uint32_t run = block_exclusive_scan(sum, s_tmp, &all) + base_d;
  for (uint32_t g = g0; g < g0 + per && g < groups; g++) {
    uint32_t c = row[g];
    row[g] = run;
    run += c;
  }
  (void)RADIX;
}

// The scene-segmented pass: one workgroup per scene s, whose chunks are
// [first[s], first[s + 1]) of the chunk-major counts.  counts[g][d] <-
// off[s] + (keys of scene s with a digit below d) + (keys of digit d in the
// scene's chunks before g): the destinations of a stable sort of each scene
// by the digit, in place.
template<int RBITS>
__global__ __launch_bounds__(kSortBlock) void
k_seg_scan(uint32_t* __restrict__ counts, const uint32_t* __restrict__ first,
           const uint64_t* __restrict__ off, uint32_t* __restrict__ zero_word)
{
  constexpr int RADIX = 1 << RBITS;
  constexpr int DPT = RADIX / kSortBlock; // digits d = tid * DPT + dd
  static_assert(DPT * kSortBlock == RADIX, "digits per thread");
  __shared__ uint32_t s_tmp[kSortWaves];
  const uint32_t sc = blockIdx.x;
  const int tid = threadIdx.x;
  if (zero_word && sc == 0 && tid == 0)
    *zero_word = 0u; // a counter of a later kernel of the same sort
  const uint32_t g0 = first[sc];
  const uint32_t g1 = first[sc + 1];
  uint32_t tot[DPT];
  uint32_t mine = 0;
#pragma unroll
  for (int dd = 0; dd < DPT; dd++) {
    const size_t d = size_t(tid) * DPT + dd;
    uint32_t acc = 0;
    for (uint32_t g = g0; g < g1; g++) {
      const uint32_t c = counts[size_t(g) * RADIX + d];
      counts[size_t(g) * RADIX + d] = acc;
      acc += c;
    }
    tot[dd] = acc;
    mine += acc;
  }
  uint32_t all;
  uint32_t run = block_exclusive_scan(mine, s_tmp, &all) + uint32_t(off[sc]);
#pragma unroll
  for (int dd = 0; dd < DPT; dd++) {
    const size_t d = size_t(tid) * DPT + dd;
    for (uint32_t g = g0; g < g1; g++)
      counts[size_t(g) * RADIX + d] += run;
    run += tot[dd];
  }
}

// Ranks and scatters chunk g, tile by tile, from its per-digit offsets
// (offsets[d * groups + g], or chunk-major offsets[g * RADIX + d] with a
// chunk table).
template<typename KeyT, bool VALS, int RBITS, int ITEMS>
__global__ __launch_bounds__(kSortBlock) void
k_rts_downsweep(const KeyT* __restrict__ kin, KeyT* __restrict__ kout,
                const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout, uint32_t n,
                int shift, uint32_t chunk_tiles, uint32_t groups,
                const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ chunks)
{
  constexpr int RADIX = 1 << RBITS;
  constexpr int TILE = kSortBlock * ITEMS;
  constexpr int DPT = RADIX / kSortBlock;
  __shared__ KeyT s_keys[TILE];
  __shared__ uint32_t s_vals[VALS ? TILE : 1];
  __shared__ uint32_t s_wcnt[kSortWaves][RADIX];
  __shared__ uint32_t s_bstart[RADIX];
  __shared__ uint32_t s_goff[RADIX];
  __shared__ uint32_t s_tmp[kSortWaves];

  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int w = tid / kWave;
  const KeyT dmask = KeyT(RADIX - 1);
  const uint32_t g = blockIdx.x;
  uint32_t running[DPT];
#pragma unroll
  for (int dd = 0; dd < DPT; dd++) {
    const size_t d = size_t(tid) * DPT + dd;
    running[dd] = offsets[chunks ? size_t(g) * RADIX + d : d * groups + g];
  }
  const uint64_t lt = lanemask_lt();
  uint32_t cbegin, cend;
  chunk_range(g, n, uint64_t(chunk_tiles) * TILE, chunks, &cbegin, &cend);

  for (uint64_t tb = cbegin; tb < cend; tb += TILE) {
    const uint32_t tile_begin = uint32_t(tb);
    for (int j = tid; j < kSortWaves * RADIX; j += kSortBlock)
      (&s_wcnt[0][0])[j] = 0u;
    __syncthreads();
    const uint32_t base = tile_begin + uint32_t(w) * (ITEMS * kWave) + uint32_t(lane);
    KeyT key[ITEMS];
    uint32_t val[ITEMS];
    uint32_t rank[ITEMS];
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      uint32_t idx = base + uint32_t(j) * kWave;
      bool valid = idx < cend;
      key[j] = valid ? kin[idx] : KeyT(0);
      if (VALS)
        val[j] = valid ? (vin ? vin[idx] : idx) : 0u;
    }
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      const uint32_t idx = base + uint32_t(j) * kWave;
      const bool valid = idx < cend;
      const uint32_t d = uint32_t((key[j] >> shift) & dmask);
      uint64_t peers = ballot(valid);
#pragma unroll
      for (int b = 0; b < RBITS; b++) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = ballot(valid && bit);
        peers &= bit ? m : ~m;
      }
      uint32_t r = 0;
      if (valid) {
        const uint32_t below = uint32_t(popc64(peers & lt));
        const uint32_t cnt = uint32_t(popc64(peers));
        const uint32_t wbase = s_wcnt[w][d];
        r = wbase + below;
        if (below == 0)
          s_wcnt[w][d] = wbase + cnt;
      }
      rank[j] = r;
    }
    __syncthreads();
    uint32_t tot[DPT];
    uint32_t my_sum = 0;
#pragma unroll
    for (int dd = 0; dd < DPT; dd++) {
      const int d = tid * DPT + dd;
      uint32_t acc = 0;
#pragma unroll
      for (int ww = 0; ww < kSortWaves; ww++) {
        uint32_t c = s_wcnt[ww][d];
        s_wcnt[ww][d] = acc;
        acc += c;
      }
      tot[dd] = acc;
      my_sum += acc;
    }
    uint32_t tile_total;
    uint32_t run = block_exclusive_scan(my_sum, s_tmp, &tile_total);
#pragma unroll
    for (int dd = 0; dd < DPT; dd++) {
      const int d = tid * DPT + dd;
      s_bstart[d] = run;
      s_goff[d] = running[dd] - run;
      run += tot[dd];
      running[dd] += tot[dd];
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < ITEMS; j++) {
      const uint32_t idx = base + uint32_t(j) * kWave;
      if (idx < cend) {
        const uint32_t d = uint32_t((key[j] >> shift) & dmask);
        const uint32_t pos = s_bstart[d] + s_wcnt[w][d] + rank[j];
        s_keys[pos] = key[j];
        if (VALS)
          s_vals[pos] = val[j];
      }
    }
    __syncthreads();
    const uint32_t tile_n =
      (cend - tile_begin < uint32_t(TILE)) ? (cend - tile_begin) : uint32_t(TILE);
    for (uint32_t i = tid; i < tile_n; i += kSortBlock) {
      const KeyT k = s_keys[i];
      const uint32_t d = uint32_t((k >> shift) & dmask);
      const uint32_t dst = s_goff[d] + i;
      if (dst < n) {
        kout[dst] = k;
        if (VALS)
          vout[dst] = s_vals[i];
      }
    }
    __syncthreads(); // LDS is reused by the next tile
  }
}

// ---- hybrid: global passes on the top digits, then LDS sorts of buckets ------
//
// After reduce-then-scan passes on the top digits the keys are grouped into
// buckets of equal top bits, in bucket order, and inside a bucket in input
// order.  Each workgroup then takes the buckets that start in its nominal
// tile of kHybridTileKeys keys (through the end of the last one) and sorts
// that range by the full key in LDS with stable 8-bit passes, skipping
// digits that do not vary in the range.  A range longer than
// kHybridTileCap is listed and sorted by k_big_sort, one workgroup per
// range, with passes through global memory.  Stable everywhere, so equal
// keys keep the input (id) order, as in the plain LSD sort.

constexpr int kTileThreads = 512;
constexpr int kTileWaves = kTileThreads / kWave;
constexpr int kTileItems = int(lbvh_eng::kHybridTileCap) / kTileThreads;
constexpr int kTileRadix = 256;
constexpr uint32_t kBigBlocks = 64;
static_assert(kTileItems * kTileThreads == int(lbvh_eng::kHybridTileCap), "tile shape");
static_assert(lbvh_eng::kHybridTileKeys <= lbvh_eng::kHybridTileCap, "tile shape");

template<int WAVES>
__device__ __forceinline__ uint32_t
block_exclusive_scan_n(uint32_t x, uint32_t* s_tmp, uint32_t* total)
{
  const int lane = __lane_id();
  const int w = threadIdx.x / kWave;
  uint32_t incl = x;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t y = __shfl_up(incl, off, kWave);
    if (lane >= off)
      incl += y;
  }
  if (lane == kWave - 1)
    s_tmp[w] = incl;
  __syncthreads();
  uint32_t woff = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < WAVES; i++) {
    uint32_t v = s_tmp[i];
    woff += (i < w) ? v : 0u;
    tot += v;
  }
  __syncthreads();
  *total = tot;
  return incl - x + woff;
}

template<typename KeyT>
__device__ __forceinline__ KeyT
wave_xor_shfl(KeyT v, int off)
{
  if constexpr (sizeof(KeyT) == 8) {
    uint32_t lo = __shfl_xor(uint32_t(v), off, kWave);
    uint32_t hi = __shfl_xor(uint32_t(uint64_t(v) >> 32), off, kWave);
    return KeyT((uint64_t(hi) << 32) | lo);
  } else {
    return KeyT(__shfl_xor(uint32_t(v), off, kWave));
  }
}

// The key bits that differ between any two of the block's keys (AND ^ OR
// over the block; threads without keys pass ~0 and 0).
template<typename KeyT>
__device__ KeyT
block_varying_bits(KeyT k_and, KeyT k_or, KeyT* s_and, KeyT* s_or)
{
  const int lane = __lane_id();
  const int w = threadIdx.x / kWave;
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    k_and &= wave_xor_shfl(k_and, off);
    k_or |= wave_xor_shfl(k_or, off);
  }
  if (lane == 0) {
    s_and[w] = k_and;
    s_or[w] = k_or;
  }
  __syncthreads();
  KeyT a = ~KeyT(0), o = KeyT(0);
#pragma unroll
  for (int i = 0; i < kTileWaves; i++) {
    a &= s_and[i];
    o |= s_or[i];
  }
  __syncthreads();
  return a ^ o;
}

// First position p of [lo, hi) with (keys[p] >> shift) > h, or hi if there
// is none; keys >> shift must not decrease along the array.  k-ary search:
// each step probes kTileThreads evenly spaced positions of the interval.
// Block-uniform arguments and result.
template<typename KeyT>
__device__ uint32_t
block_upper_bound(const KeyT* __restrict__ keys, uint32_t lo, uint32_t hi, int shift, KeyT h,
                  uint32_t* s_first)
{
  const uint32_t tid = threadIdx.x;
  while (lo < hi) {
    const uint32_t step = (hi - lo + kTileThreads - 1) / kTileThreads;
    if (tid == 0)
      *s_first = kTileThreads;
    __syncthreads();
    const uint64_t p = uint64_t(lo) + uint64_t(tid) * step;
    if (p < hi && (keys[p] >> shift) > h)
      atomicMin(s_first, tid);
    __syncthreads();
    const uint32_t t = *s_first;
    __syncthreads();
    if (t == uint32_t(kTileThreads)) {
      lo += ((hi - lo - 1) / step) * step + 1; // past the last probe
    } else {
      hi = lo + t * step;
      if (t > 0)
        lo += (t - 1) * step + 1;
    }
  }
  return lo;
}

// The range of whole buckets a workgroup sorts: the buckets that start in
// its nominal tile [tile0, tile1), through the end of the last one.  One
// round of probes covers the tile (for the start) and the same number of
// keys after it (for the end); each probe compares a key's top bits with
// its predecessor's.  Only a bucket that runs on past that window needs the
// k-ary search.  Returns false (block-uniform) when no bucket starts in the
// tile.
template<typename KeyT>
__device__ bool
find_range(const KeyT* __restrict__ keys, uint32_t n, uint32_t tile0, uint32_t tile1, int shift,
           uint32_t* s_first, uint32_t* start, uint32_t* end)
{
  constexpr uint32_t W = lbvh_eng::kHybridTileKeys;
  constexpr int PROBES = int(W) / kTileThreads;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) {
    s_first[0] = 0xFFFFFFFFu;
    s_first[1] = 0xFFFFFFFFu;
  }
  __syncthreads();
  uint32_t f0 = 0xFFFFFFFFu, f1 = 0xFFFFFFFFu;
#pragma unroll
  for (int k = 0; k < PROBES; k++) {
    const uint32_t p = tile0 + tid + uint32_t(k) * kTileThreads;
    if (tile0 > 0 && p < tile1 && f0 == 0xFFFFFFFFu && (keys[p] >> shift) != (keys[p - 1] >> shift))
      f0 = p;
    const uint64_t q = uint64_t(tile1) + tid + uint64_t(k) * kTileThreads;
    if (q < n && f1 == 0xFFFFFFFFu && (keys[q] >> shift) != (keys[q - 1] >> shift))
      f1 = uint32_t(q);
  }
  if (f0 != 0xFFFFFFFFu)
    atomicMin(&s_first[0], f0);
  if (f1 != 0xFFFFFFFFu)
    atomicMin(&s_first[1], f1);
  __syncthreads();
  const uint32_t b0 = tile0 == 0 ? 0u : s_first[0];
  const uint32_t b1 = s_first[1];
  __syncthreads(); // s_first is reused by the search below
  if (b0 == 0xFFFFFFFFu)
    return false;
  *start = b0;
  if (tile1 >= n) {
    *end = n;
  } else if (b1 != 0xFFFFFFFFu) {
    *end = b1;
  } else {
    const uint64_t from = uint64_t(tile1) + W;
    *end = from >= n ? n
                     : block_upper_bound(keys, uint32_t(from), n, shift,
                                         KeyT(keys[tile1 - 1] >> shift), s_first);
  }
  return true;
}

// Stable ranks of 8-bit digits inside each wave (wave-striped items: item j
// of lane l of wave w is element w * ipw * 64 + j * 64 + l, j < ipw); digit
// bits outside `dvar` are equal for all items and skipped.  Leaves per-wave
// digit counts in s_wcnt.
__device__ __forceinline__ void
rank_digits(const uint32_t (&dg)[kTileItems], int ipw, uint32_t lbase, uint32_t m, uint32_t dvar,
            uint32_t (&rank)[kTileItems], uint32_t (*s_wcnt)[kTileRadix])
{
  const int w = threadIdx.x / kWave;
  const uint64_t lt = lanemask_lt();
#pragma unroll
  for (int j = 0; j < kTileItems; j++) {
    if (j >= ipw)
      break; // block-uniform
    const bool valid = lbase + uint32_t(j) * kWave < m;
    const uint32_t d = dg[j];
    uint64_t peers = ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      if ((dvar >> b) & 1u) {
        const bool bit = (d >> b) & 1u;
        const uint64_t mm = ballot(valid && bit);
        peers &= bit ? mm : ~mm;
      }
    }
    uint32_t r = 0;
    if (valid) {
      const uint32_t below = uint32_t(popc64(peers & lt));
      const uint32_t cnt = uint32_t(popc64(peers));
      const uint32_t wbase = s_wcnt[w][d];
      r = wbase + below;
      if (below == 0)
        s_wcnt[w][d] = wbase + cnt;
    }
    rank[j] = r;
  }
}

// Turns s_wcnt into exclusive per-wave offsets inside each digit and
// returns (threads below kTileRadix) the digit's total.
__device__ __forceinline__ uint32_t
digit_wave_offsets(uint32_t (*s_wcnt)[kTileRadix])
{
  const int tid = threadIdx.x;
  uint32_t acc = 0;
  if (tid < kTileRadix) {
#pragma unroll
    for (int ww = 0; ww < kTileWaves; ww++) {
      const uint32_t c = s_wcnt[ww][tid];
      s_wcnt[ww][tid] = acc;
      acc += c;
    }
  }
  return acc;
}

// Exclusive count, in element order, of the items whose flag is set (the
// wave-striped layout of rank_digits); *total = the block's count.  Uses
// s_tmp and ends with a barrier.
__device__ __forceinline__ void
flag_prefix(const bool (&flag)[kTileItems], int ipw, uint32_t (&pre)[kTileItems],
            uint32_t* s_tmp, uint32_t* total)
{
  const int lane = __lane_id();
  const int w = threadIdx.x / kWave;
  const uint64_t lt = lanemask_lt();
  uint32_t running = 0;
#pragma unroll
  for (int j = 0; j < kTileItems; j++) {
    if (j >= ipw)
      break;
    const uint64_t mask = ballot(flag[j]);
    pre[j] = running + uint32_t(popc64(mask & lt));
    running += uint32_t(popc64(mask));
  }
  if (lane == 0)
    s_tmp[w] = running;
  __syncthreads();
  uint32_t woff = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kTileWaves; i++) {
    const uint32_t v = s_tmp[i];
    woff += (i < w) ? v : 0u;
    tot += v;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kTileItems; j++) {
    if (j >= ipw)
      break;
    pre[j] += woff;
  }
  *total = tot;
}

// Comparisons per key the counting rank may spend on average before the
// workgroup falls back to LSD passes over every digit.
constexpr uint32_t kCountRankBudget = 64;

// Sorts one range of whole buckets by the full key, stably:
//   1. bucket index b of every key (its top bits are the bucket's, and
//      buckets are already in order);
//   2. stable LDS passes on the top digit of the low bits, then on b: the
//      range is now ordered by (bucket, split digit) in input order within;
//   3. inside each run of equal (bucket, split digit) every key counts the
//      keys below it (ties: earlier positions) and goes to run start + count.
// When the runs are too long for step 3, LDS passes over every low digit
// and then b finish the sort instead.
template<typename KeyT, bool VALS>
__global__ __launch_bounds__(kTileThreads) void
k_tile_sort(const KeyT* __restrict__ kin, KeyT* __restrict__ kout,
            const uint32_t* __restrict__ vin, uint32_t* __restrict__ vout, uint32_t n,
            int hi_shift, int key_bits, uint32_t* __restrict__ big,
            uint32_t* __restrict__ big_count, uint32_t big_cap)
{
  constexpr uint32_t T = lbvh_eng::kHybridTileKeys;
  constexpr uint32_t CAP = lbvh_eng::kHybridTileCap;
  __shared__ KeyT s_keys[CAP];
  __shared__ uint32_t s_vals[VALS ? CAP : 1];
  __shared__ uint16_t s_aux[CAP];                        // bucket index
  __shared__ uint32_t s_wcnt[2][kTileWaves][kTileRadix]; // per pass, alternating
  __shared__ uint32_t s_tmp[kTileWaves];
  __shared__ KeyT s_and[kTileWaves];
  __shared__ KeyT s_or[kTileWaves];
  __shared__ uint32_t s_first[2];
  __shared__ uint32_t s_work;
  (void)key_bits;

  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int w = tid / kWave;
  const uint64_t t0 = uint64_t(blockIdx.x) * T;
  if (t0 >= n)
    return;
  const uint32_t tile0 = uint32_t(t0);
  const uint32_t tile1 = (n - tile0 > T) ? tile0 + T : n;
  uint32_t start = 0, end = 0;
  if (!find_range(kin, n, tile0, tile1, hi_shift, s_first, &start, &end))
    return; // the whole tile continues a bucket that started earlier
  const uint32_t m = end - start;
  if (m > CAP) {
    if (tid == 0) {
      const uint32_t slot = atomicAdd(big_count, 1u);
      if (slot < big_cap) {
        big[2 * slot] = start;
        big[2 * slot + 1] = end;
      }
    }
    return;
  }

  // items per wave: just enough slots for the range
  const int ipw = int((m + kTileThreads - 1) / kTileThreads);
  const uint32_t lbase = uint32_t(w) * uint32_t(ipw * kWave) + uint32_t(lane);
  KeyT key[kTileItems];
  uint32_t val[kTileItems];
  uint32_t aux[kTileItems];
  bool flag[kTileItems];
  KeyT k_and = ~KeyT(0), k_or = KeyT(0);
#pragma unroll
  for (int j = 0; j < kTileItems; j++) {
    if (j >= ipw)
      break;
    const uint32_t li = lbase + uint32_t(j) * kWave;
    const bool valid = li < m;
    key[j] = valid ? kin[start + li] : KeyT(0);
    if (VALS)
      val[j] = valid ? vin[start + li] : 0u;
    flag[j] = valid && (li == 0 || (key[j] >> hi_shift) != (kin[start + li - 1] >> hi_shift));
    if (valid) {
      k_and &= key[j];
      k_or |= key[j];
    }
  }
  for (int j = tid; j < kTileWaves * kTileRadix; j += kTileThreads)
    (&s_wcnt[0][0][0])[j] = 0u;
  uint32_t nb;
  flag_prefix(flag, ipw, aux, s_tmp, &nb); // exclusive count of bucket starts
#pragma unroll
  for (int j = 0; j < kTileItems; j++) {
    if (j >= ipw)
      break;
    aux[j] += flag[j] ? 1u : 0u; // inclusive, minus the first start: the bucket index
    aux[j] -= 1u;
  }
  const KeyT var = block_varying_bits(k_and, k_or, s_and, s_or); // also orders the clear
  const uint32_t bmask = nb > 1 ? (0xFFFFFFFFu >> __clz(nb - 1)) : 0u;
  const int kmax = (hi_shift + 7) / 8; // 8-bit digits below the top bits
  bool staged = false;
  int pb = 0;
  int k = 1;
  for (;;) {
    // stable LDS passes: split digits from the lowest to the highest, then
    // the bucket index (low byte, high byte)
    for (int step = 0; step < k + 2; step++) {
      uint32_t dvar;
      int sh;
      const bool on_aux = step >= k;
      if (!on_aux) {
        sh = hi_shift - 8 * (k - step);
        if (sh < 0)
          sh = 0;
        dvar = uint32_t(var >> sh) & uint32_t(kTileRadix - 1);
      } else {
        sh = 8 * (step - k);
        dvar = (bmask >> sh) & uint32_t(kTileRadix - 1);
      }
      if (dvar == 0)
        continue; // block-uniform
      uint32_t (*cnt)[kTileRadix] = s_wcnt[pb];
      if (staged) {
#pragma unroll
        for (int j = 0; j < kTileItems; j++) {
          if (j >= ipw)
            break;
          const uint32_t li = lbase + uint32_t(j) * kWave;
          if (li < m) {
            key[j] = s_keys[li];
            aux[j] = s_aux[li];
            if (VALS)
              val[j] = s_vals[li];
          }
        }
      }
      uint32_t dg[kTileItems];
#pragma unroll
      for (int j = 0; j < kTileItems; j++) {
        if (j >= ipw)
          break;
        dg[j] = on_aux ? (aux[j] >> sh) & uint32_t(kTileRadix - 1)
                       : uint32_t(key[j] >> sh) & uint32_t(kTileRadix - 1);
      }
      uint32_t rank[kTileItems];
      rank_digits(dg, ipw, lbase, m, dvar, rank, cnt);
      __syncthreads();
      // digit starts: exclusive scan of the digit totals (threads below the
      // radix), folded into the per-wave offsets
      const uint32_t tot = digit_wave_offsets(cnt);
      uint32_t incl = tot;
#pragma unroll
      for (int off = 1; off < kWave; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off, kWave);
        if (lane >= off)
          incl += y;
      }
      if (lane == kWave - 1)
        s_tmp[w] = incl;
      __syncthreads();
      uint32_t woff = 0;
#pragma unroll
      for (int i = 0; i < kTileWaves; i++)
        woff += (i < w) ? s_tmp[i] : 0u;
      if (tid < kTileRadix) {
        const uint32_t dstart = incl - tot + woff;
#pragma unroll
        for (int ww = 0; ww < kTileWaves; ww++)
          cnt[ww][tid] += dstart;
      }
      for (int j = tid; j < kTileWaves * kTileRadix; j += kTileThreads)
        (&s_wcnt[pb ^ 1][0][0])[j] = 0u; // the next pass's counters
      __syncthreads();
#pragma unroll
      for (int j = 0; j < kTileItems; j++) {
        if (j >= ipw)
          break;
        const uint32_t li = lbase + uint32_t(j) * kWave;
        if (li < m) {
          const uint32_t pos = cnt[w][dg[j]] + rank[j];
          s_keys[pos] = key[j];
          s_aux[pos] = uint16_t(aux[j]);
          if (VALS)
            s_vals[pos] = val[j];
        }
      }
      __syncthreads();
      staged = true;
      pb ^= 1;
    }
    if (!staged) {
      // no digit varied: the input order is the order by (bucket, split
      // digit); stage it for the steps below
#pragma unroll
      for (int j = 0; j < kTileItems; j++) {
        if (j >= ipw)
          break;
        const uint32_t li = lbase + uint32_t(j) * kWave;
        if (li < m) {
          s_keys[li] = key[j];
          s_aux[li] = uint16_t(aux[j]);
          if (VALS)
            s_vals[li] = val[j];
        }
      }
      __syncthreads();
      staged = true;
    }
    if (k >= kmax)
      break; // every digit sorted
    // runs of equal (bucket, split digits) in the current order
    int cut = hi_shift - 8 * k;
    if (cut < 0)
      cut = 0;
    uint32_t ss[kTileItems], se[kTileItems];
    uint32_t my_work = 0;
#pragma unroll
    for (int j = 0; j < kTileItems; j++) {
      if (j >= ipw)
        break;
      const uint32_t li = lbase + uint32_t(j) * kWave;
      const bool valid = li < m;
      if (valid) {
        key[j] = s_keys[li];
        if (VALS)
          val[j] = s_vals[li];
      }
      flag[j] = valid && (li == 0 || (key[j] >> cut) != (s_keys[li - 1] >> cut));
    }
    uint32_t ns;
    flag_prefix(flag, ipw, ss, s_tmp, &ns); // ss: exclusive count of run starts
    uint16_t* s_run = reinterpret_cast<uint16_t*>(&s_wcnt[0][0][0]); // CAP entries
    if (tid == 0)
      s_work = 0u;
#pragma unroll
    for (int j = 0; j < kTileItems; j++) {
      if (j >= ipw)
        break;
      const uint32_t li = lbase + uint32_t(j) * kWave;
      if (flag[j])
        s_run[ss[j]] = uint16_t(li);
      if (li < m)
        ss[j] = ss[j] + (flag[j] ? 1u : 0u) - 1u; // run index
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kTileItems; j++) {
      if (j >= ipw)
        break;
      const uint32_t li = lbase + uint32_t(j) * kWave;
      if (li < m) {
        const uint32_t r = ss[j];
        ss[j] = s_run[r];
        se[j] = r + 1 < ns ? uint32_t(s_run[r + 1]) : m;
        my_work += se[j] - ss[j];
      }
    }
    atomicAdd(&s_work, my_work);
    __syncthreads();
    const uint32_t work = s_work;
    if (work <= kCountRankBudget * m) {
      // counting rank inside each run; the output goes straight to HBM
#pragma unroll
      for (int j = 0; j < kTileItems; j++) {
        if (j >= ipw)
          break;
        const uint32_t li = lbase + uint32_t(j) * kWave;
        if (li < m) {
          const KeyT kj = key[j];
          uint32_t r = 0;
          for (uint32_t t = ss[j]; t < se[j]; t++) {
            const KeyT kt = s_keys[t];
            r += (kt < kj || (kt == kj && t < li)) ? 1u : 0u;
          }
          const uint32_t dst = start + ss[j] + r;
          kout[dst] = kj;
          if (VALS)
            vout[dst] = val[j];
        }
      }
      return;
    }
    // long runs: sort by every low digit, then the bucket index
    __syncthreads(); // s_run (in the counters) is read before the passes clear them
    for (int j = tid; j < 2 * kTileWaves * kTileRadix; j += kTileThreads)
      (&s_wcnt[0][0][0])[j] = 0u;
    __syncthreads();
    k = kmax;
  }
  for (uint32_t i = tid; i < m; i += kTileThreads) {
    kout[start + i] = s_keys[i];
    if (VALS)
      vout[start + i] = s_vals[i];
  }
}

// Ranges longer than kHybridTileCap: one workgroup per range, stable 8-bit
// passes (histogram, then tiles in order with running digit offsets) that
// ping-pong between the range's slices of the two buffers; the result ends
// in kout.
template<typename KeyT, bool VALS>
__global__ __launch_bounds__(kTileThreads) void
k_big_sort(KeyT* kin, KeyT* kout, uint32_t* vin, uint32_t* vout, int key_bits,
           const uint32_t* __restrict__ big, const uint32_t* __restrict__ big_count,
           uint32_t big_cap)
{
  constexpr uint32_t CAP = lbvh_eng::kHybridTileCap;
  __shared__ uint32_t s_wcnt[kTileWaves][kTileRadix];
  __shared__ uint32_t s_run[kTileRadix];
  __shared__ uint32_t s_tmp[kTileWaves];
  __shared__ KeyT s_and[kTileWaves];
  __shared__ KeyT s_or[kTileWaves];
  const int tid = threadIdx.x;
  const int lane = __lane_id();
  const int w = tid / kWave;
  const uint32_t count = *big_count < big_cap ? *big_count : big_cap;
  const uint32_t lbase = uint32_t(w) * (kTileItems * kWave) + uint32_t(lane);
  for (uint32_t e = blockIdx.x; e < count; e += gridDim.x) {
    const uint32_t start = big[2 * e];
    const uint32_t m = big[2 * e + 1] - start;
    KeyT* src = kin + start;
    KeyT* dst = kout + start;
    uint32_t* vs = VALS ? vin + start : nullptr;
    uint32_t* vd = VALS ? vout + start : nullptr;
    KeyT k_and = ~KeyT(0), k_or = KeyT(0);
    for (uint32_t i = tid; i < m; i += kTileThreads) {
      const KeyT k = src[i];
      k_and &= k;
      k_or |= k;
    }
    const KeyT var = block_varying_bits(k_and, k_or, s_and, s_or);
    for (int shift = 0; shift < key_bits; shift += 8) {
      const uint32_t dvar = uint32_t(var >> shift) & uint32_t(kTileRadix - 1);
      if (dvar == 0)
        continue;
      if (tid < kTileRadix)
        s_run[tid] = 0u;
      __syncthreads();
      for (uint32_t i = tid; i < m; i += kTileThreads)
        atomicAdd(&s_run[uint32_t(src[i] >> shift) & uint32_t(kTileRadix - 1)], 1u);
      __syncthreads();
      const uint32_t h = tid < kTileRadix ? s_run[tid] : 0u;
      uint32_t all;
      const uint32_t run = block_exclusive_scan_n<kTileWaves>(h, s_tmp, &all);
      if (tid < kTileRadix)
        s_run[tid] = run;
      for (uint32_t tb = 0; tb < m; tb += CAP) {
        for (int j = tid; j < kTileWaves * kTileRadix; j += kTileThreads)
          (&s_wcnt[0][0])[j] = 0u;
        __syncthreads();
        const uint32_t tm = (m - tb < CAP) ? m - tb : CAP;
        KeyT key[kTileItems];
        uint32_t val[kTileItems];
#pragma unroll
        for (int j = 0; j < kTileItems; j++) {
          const uint32_t li = lbase + uint32_t(j) * kWave;
          const bool valid = li < tm;
          key[j] = valid ? src[tb + li] : KeyT(0);
          if (VALS)
            val[j] = valid ? vs[tb + li] : 0u;
        }
        uint32_t dg[kTileItems];
#pragma unroll
        for (int j = 0; j < kTileItems; j++)
          dg[j] = uint32_t(key[j] >> shift) & uint32_t(kTileRadix - 1);
        uint32_t rank[kTileItems];
        rank_digits(dg, kTileItems, lbase, tm, dvar, rank, s_wcnt);
        __syncthreads();
        const uint32_t tot = digit_wave_offsets(s_wcnt);
        __syncthreads();
#pragma unroll
        for (int j = 0; j < kTileItems; j++) {
          const uint32_t li = lbase + uint32_t(j) * kWave;
          if (li < tm) {
            const uint32_t d = uint32_t(key[j] >> shift) & uint32_t(kTileRadix - 1);
            const uint32_t pos = s_run[d] + s_wcnt[w][d] + rank[j];
            dst[pos] = key[j];
            if (VALS)
              vd[pos] = val[j];
          }
        }
        __syncthreads();
        if (tid < kTileRadix)
          s_run[tid] += tot;
      }
      __threadfence();
      __syncthreads();
      KeyT* tk = src;
      src = dst;
      dst = tk;
      uint32_t* tv = vs;
      vs = vd;
      vd = tv;
    }
    if (src != kout + start) {
      for (uint32_t i = tid; i < m; i += kTileThreads) {
        kout[start + i] = src[i];
        if (VALS)
          vout[start + i] = vs[i];
      }
    }
    __threadfence();
    __syncthreads();
  }
}

} // namespace lbvh_dev

namespace lbvh_eng {

using namespace lbvh_dev;

uint64_t
sort_status_words(uint64_t n, const SortConfig& cfg)
{
  const uint64_t tile = uint64_t(kSortBlock) * cfg.items;
  const uint64_t blocks = (n + tile - 1) / tile;
  return blocks * (uint64_t(1) << cfg.digit_bits);
}

template<typename KeyT, bool VALS, int RBITS, int ITEMS>
static hipError_t
run_passes(int key_bits, uint64_t n, KeyT* ka, KeyT* kb, uint32_t* va, uint32_t* vb,
           const SortWorkspace& ws, uint32_t epoch, bool* in_b, hipStream_t s)
{
  constexpr int RADIX = 1 << RBITS;
  constexpr uint32_t TILE = kSortBlock * ITEMS;
  const int passes = sort_passes(key_bits, RBITS);
  const uint32_t blocks = uint32_t((n + TILE - 1) / TILE);
  KeyT* src = ka;
  KeyT* dst = kb;
  uint32_t* vsrc = nullptr; // iota on the first pass
  uint32_t* vdst = vb;
  bool cur_b = false;
  for (int p = 0; p < passes; p++) {
    hipLaunchKernelGGL((k_onesweep<KeyT, VALS, RBITS, ITEMS>), dim3(blocks), dim3(kSortBlock),
                       0, s, src, dst, vsrc, vdst, uint32_t(n), p * RBITS,
                       ws.hist + size_t(p) * RADIX, ws.status, ws.tickets + p, epoch + uint32_t(p),
                       ws.err);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return e;
    cur_b = !cur_b;
    KeyT* t = src;
    src = dst;
    dst = t;
    if (VALS) {
      vsrc = vdst;
      vdst = (vdst == vb) ? va : vb;
    }
  }
  *in_b = cur_b;
  return hipSuccess;
}

uint32_t
rts_groups(uint64_t n, const SortConfig& cfg, uint32_t* chunk_tiles)
{
  const uint64_t tile = uint64_t(kSortBlock) * cfg.items;
  const uint64_t tiles = (n + tile - 1) / tile;
  // default: chunks of about two tiles, 512 to 2,048 of them (more chunks,
  // more workgroups in flight for the memory-bound downsweep; fewer, a
  // smaller count matrix)
  uint64_t target = tiles / 2;
  target = target < 512 ? 512 : (target > 2048 ? 2048 : target);
  if (cfg.groups > 0)
    target = uint64_t(cfg.groups);
  uint64_t ct = (tiles + target - 1) / target;
  if (ct < 1)
    ct = 1;
  if (chunk_tiles)
    *chunk_tiles = uint32_t(ct);
  return uint32_t((tiles + ct - 1) / ct);
}

bool
seg_pass_applies(int key_bits, const SortConfig& cfg)
{
  if (cfg.strategy != 2 || key_bits < 30)
    return false;
  const int passes = sort_passes(key_bits, cfg.digit_bits);
  const int hi_shift = (passes - cfg.global_digits) * cfg.digit_bits;
  return cfg.global_digits >= 1 && hi_shift < 30 && hi_shift + cfg.digit_bits >= 30;
}

uint64_t
seg_chunk_keys(uint64_t n)
{
  // 16,384 keys (measured best on 125M keys in 4,096 scenes against 8,192,
  // 32,768 and 61,440), or about 8,192 chunks over larger arrays, in whole
  // tiles of either size
  constexpr uint64_t kUnit = uint64_t(kSortBlock) * 16;
  uint64_t c = (n + 8191) / 8192;
  if (c < 16384)
    c = 16384;
  return (c + kUnit - 1) / kUnit * kUnit;
}

uint64_t
rts_chunk_keys(uint64_t n, const SortConfig& cfg)
{
  uint32_t ct = 1;
  rts_groups(n, cfg, &ct);
  return uint64_t(ct) * kSortBlock * uint64_t(cfg.items);
}

template<typename KeyT, bool VALS, int RBITS, int ITEMS>
static hipError_t
run_rts(int key_bits, uint64_t n, KeyT* ka, KeyT* kb, uint32_t* va, uint32_t* vb,
        const SortConfig& cfg, const SortWorkspace& ws, bool* in_b, hipStream_t s)
{
  constexpr int RADIX = 1 << RBITS;
  constexpr uint32_t TILE = kSortBlock * ITEMS;
  const int passes = sort_passes(key_bits, RBITS);
  uint32_t chunk_tiles = 1;
  const uint32_t groups = rts_groups(n, cfg, &chunk_tiles);
  if (uint64_t(groups) * RADIX > ws.counts_words)
    return hipErrorInvalidValue;
  KeyT* src = ka;
  KeyT* dst = kb;
  uint32_t* vsrc = nullptr;
  uint32_t* vdst = vb;
  bool cur_b = false;
  for (int p = 0; p < passes; p++) {
    const int shift = p * RBITS;
    if (p > 0 || !ws.first_counts)
      hipLaunchKernelGGL((k_rts_upsweep<KeyT, RBITS>), dim3(groups), dim3(kSortBlock), 0, s, src,
                         uint32_t(n), shift, chunk_tiles * TILE, groups, ws.counts, nullptr);
    hipLaunchKernelGGL((k_rts_scan<RBITS>), dim3(RADIX), dim3(kSortBlock), 0, s, ws.counts,
                       groups, ws.hist + size_t(p) * RADIX, nullptr);
    hipLaunchKernelGGL((k_rts_downsweep<KeyT, VALS, RBITS, ITEMS>), dim3(groups),
                       dim3(kSortBlock), 0, s, src, dst, vsrc, vdst, uint32_t(n), shift,
                       chunk_tiles, groups, ws.counts, nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return e;
    cur_b = !cur_b;
    KeyT* t = src;
    src = dst;
    dst = t;
    if (VALS) {
      vsrc = vdst;
      vdst = (vdst == vb) ? va : vb;
    }
  }
  *in_b = cur_b;
  return hipSuccess;
}

// The hybrid's LDS stage: keys src (grouped into buckets of equal
// key >> hi_shift, in bucket order) sorted into dst; *in_b = dst_is_b.
template<typename KeyT, bool VALS>
static hipError_t
bucket_sorts(int key_bits, uint64_t n, KeyT* src, KeyT* dst, uint32_t* vsrc, uint32_t* vdst,
             int hi_shift, uint32_t* big, uint32_t* big_count, uint64_t big_cap, bool dst_is_b,
             bool* in_b, hipStream_t s)
{
  const uint32_t tiles = uint32_t((n + kHybridTileKeys - 1) / kHybridTileKeys);
  hipLaunchKernelGGL((k_tile_sort<KeyT, VALS>), dim3(tiles), dim3(kTileThreads), 0, s, src, dst,
                     vsrc, vdst, uint32_t(n), hi_shift, key_bits, big, big_count,
                     uint32_t(big_cap));
  hipLaunchKernelGGL((k_big_sort<KeyT, VALS>), dim3(kBigBlocks), dim3(kTileThreads), 0, s, src,
                     dst, vsrc, vdst, key_bits, big, big_count, uint32_t(big_cap));
  *in_b = dst_is_b;
  return hipGetLastError();
}

template<typename KeyT, bool VALS, int RBITS, int ITEMS>
static hipError_t
run_hybrid(int key_bits, uint64_t n, KeyT* ka, KeyT* kb, uint32_t* va, uint32_t* vb,
           const SortConfig& cfg, const SortWorkspace& ws, bool* in_b, hipStream_t s)
{
  constexpr int RADIX = 1 << RBITS;
  constexpr uint32_t TILE = kSortBlock * ITEMS;
  const int passes = sort_passes(key_bits, RBITS);
  const int gd = cfg.global_digits;
  if (gd < 1 || gd >= passes)
    return hipErrorInvalidValue;
  uint32_t chunk_tiles = 1;
  const uint32_t groups = rts_groups(n, cfg, &chunk_tiles);
  const uint64_t big_cap = n / kHybridTileCap + 1; // ranges longer than the cap are disjoint
  if (uint64_t(groups) * RADIX + 2 + 2 * big_cap > ws.counts_words)
    return hipErrorInvalidValue;
  KeyT* src = ka;
  KeyT* dst = kb;
  uint32_t* vsrc = nullptr;
  uint32_t* vdst = vb;
  const int hi_shift = (passes - gd) * RBITS;
  if (ws.seg_chunks && seg_pass_applies(key_bits, cfg)) {
    // keys grouped by scene (bits 30 and up) in ascending order, and the
    // global digits hold one digit of code bits: sort each scene by that
    // digit in one pass over scene-aligned chunks
    const uint32_t sg = ws.seg_groups;
    if (uint64_t(sg) * RADIX + 2 + 2 * big_cap > ws.counts_words)
      return hipErrorInvalidValue;
    uint32_t* big_count = ws.counts + size_t(sg) * RADIX;
    uint32_t* big = big_count + 2;
    if (!ws.first_counts)
      hipLaunchKernelGGL((k_rts_upsweep<KeyT, RBITS>), dim3(sg), dim3(kSortBlock), 0, s, src,
                         uint32_t(n), hi_shift, 0u, sg, ws.counts, ws.seg_chunks);
    hipLaunchKernelGGL((k_seg_scan<RBITS>), dim3(ws.nseg), dim3(kSortBlock), 0, s, ws.counts,
                       ws.seg_first, ws.seg_off, big_count);
    hipLaunchKernelGGL((k_rts_downsweep<KeyT, VALS, RBITS, ITEMS>), dim3(sg), dim3(kSortBlock),
                       0, s, src, dst, vsrc, vdst, uint32_t(n), hi_shift, 0u, sg, ws.counts,
                       ws.seg_chunks);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return e;
    return bucket_sorts<KeyT, VALS>(key_bits, n, dst, src, VALS ? vdst : nullptr,
                                    VALS ? (vdst == vb ? va : vb) : nullptr, hi_shift, big,
                                    big_count, big_cap, src == kb, in_b, s);
  }
  uint32_t* big_count = ws.counts + size_t(groups) * RADIX;
  uint32_t* big = big_count + 2;
  for (int p = passes - gd; p < passes; p++) {
    const int shift = p * RBITS;
    if (p > passes - gd || !ws.first_counts)
      hipLaunchKernelGGL((k_rts_upsweep<KeyT, RBITS>), dim3(groups), dim3(kSortBlock), 0, s, src,
                         uint32_t(n), shift, chunk_tiles * TILE, groups, ws.counts, nullptr);
    hipLaunchKernelGGL((k_rts_scan<RBITS>), dim3(RADIX), dim3(kSortBlock), 0, s, ws.counts,
                       groups, ws.hist + size_t(p) * RADIX,
                       p == passes - gd ? big_count : nullptr);
    hipLaunchKernelGGL((k_rts_downsweep<KeyT, VALS, RBITS, ITEMS>), dim3(groups),
                       dim3(kSortBlock), 0, s, src, dst, vsrc, vdst, uint32_t(n), shift,
                       chunk_tiles, groups, ws.counts, nullptr);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return e;
    KeyT* t = src;
    src = dst;
    dst = t;
    if (VALS) {
      vsrc = vdst;
      vdst = (vdst == vb) ? va : vb;
    }
  }
  return bucket_sorts<KeyT, VALS>(key_bits, n, src, dst, vsrc, vdst, hi_shift, big, big_count,
                                  big_cap, dst == kb, in_b, s);
}

uint64_t
sort_counts_words(uint64_t n, const SortConfig& cfg)
{
  uint64_t w = uint64_t(rts_groups(n, cfg, nullptr)) << cfg.digit_bits;
  if (cfg.strategy == 2)
    w += 2 + 2 * (n / kHybridTileCap + 1);
  return w;
}

template<typename KeyT, bool VALS>
static hipError_t
dispatch(int key_bits, uint64_t n, KeyT* ka, KeyT* kb, uint32_t* va, uint32_t* vb,
         const SortConfig& cfg, const SortWorkspace& ws, uint32_t epoch, bool* in_b,
         hipStream_t s)
{
  if (cfg.strategy == 2) {
    if (cfg.digit_bits == 8 && cfg.items == 8)
      return run_hybrid<KeyT, VALS, 8, 8>(key_bits, n, ka, kb, va, vb, cfg, ws, in_b, s);
    if (cfg.digit_bits == 8 && cfg.items == 16)
      return run_hybrid<KeyT, VALS, 8, 16>(key_bits, n, ka, kb, va, vb, cfg, ws, in_b, s);
    if (cfg.digit_bits == 10 && cfg.items == 8)
      return run_hybrid<KeyT, VALS, 10, 8>(key_bits, n, ka, kb, va, vb, cfg, ws, in_b, s);
    if (cfg.digit_bits == 10 && cfg.items == 16)
      return run_hybrid<KeyT, VALS, 10, 16>(key_bits, n, ka, kb, va, vb, cfg, ws, in_b, s);
    return hipErrorInvalidValue;
  }
  if (cfg.strategy == 1) {
    if (cfg.digit_bits == 8 && cfg.items == 8)
      return run_rts<KeyT, VALS, 8, 8>(key_bits, n, ka, kb, va, vb, cfg, ws, in_b, s);
    if (cfg.digit_bits == 8 && cfg.items == 16)
      return run_rts<KeyT, VALS, 8, 16>(key_bits, n, ka, kb, va, vb, cfg, ws, in_b, s);
    if (cfg.digit_bits == 10 && cfg.items == 8)
      return run_rts<KeyT, VALS, 10, 8>(key_bits, n, ka, kb, va, vb, cfg, ws, in_b, s);
    if (cfg.digit_bits == 10 && cfg.items == 16)
      return run_rts<KeyT, VALS, 10, 16>(key_bits, n, ka, kb, va, vb, cfg, ws, in_b, s);
    return hipErrorInvalidValue;
  }
  if (cfg.digit_bits == 8) {
    if (cfg.items == 8)
      return run_passes<KeyT, VALS, 8, 8>(key_bits, n, ka, kb, va, vb, ws, epoch, in_b, s);
    if (cfg.items == 16)
      return run_passes<KeyT, VALS, 8, 16>(key_bits, n, ka, kb, va, vb, ws, epoch, in_b, s);
    if (cfg.items == 4)
      return run_passes<KeyT, VALS, 8, 4>(key_bits, n, ka, kb, va, vb, ws, epoch, in_b, s);
  } else if (cfg.digit_bits == 10) {
    if (cfg.items == 8)
      return run_passes<KeyT, VALS, 10, 8>(key_bits, n, ka, kb, va, vb, ws, epoch, in_b, s);
    if (cfg.items == 16)
      return run_passes<KeyT, VALS, 10, 16>(key_bits, n, ka, kb, va, vb, ws, epoch, in_b, s);
  }
  return hipErrorInvalidValue;
}

hipError_t
launch_sort(int key_bits, int key_bytes, uint64_t n, void* keys_a, void* keys_b, uint32_t* ids_a,
            uint32_t* ids_b, const SortConfig& cfg, const SortWorkspace& ws, uint32_t epoch,
            bool* result_in_b, hipStream_t s)
{
  if (n == 0 || n >= (uint64_t(1) << 31) || (key_bytes != 4 && key_bytes != 8) ||
      key_bits > 8 * key_bytes)
    return hipErrorInvalidValue;
  const bool vals = ids_a != nullptr;
  if (key_bytes == 4) {
    auto* a = static_cast<uint32_t*>(keys_a);
    auto* b = static_cast<uint32_t*>(keys_b);
    return vals ? dispatch<uint32_t, true>(key_bits, n, a, b, ids_a, ids_b, cfg, ws, epoch,
                                           result_in_b, s)
                : dispatch<uint32_t, false>(key_bits, n, a, b, nullptr, nullptr, cfg, ws, epoch,
                                            result_in_b, s);
  }
  auto* a = static_cast<uint64_t*>(keys_a);
  auto* b = static_cast<uint64_t*>(keys_b);
  return vals ? dispatch<uint64_t, true>(key_bits, n, a, b, ids_a, ids_b, cfg, ws, epoch,
                                         result_in_b, s)
              : dispatch<uint64_t, false>(key_bits, n, a, b, nullptr, nullptr, cfg, ws, epoch,
                                          result_in_b, s);
}

} // namespace lbvh_eng
