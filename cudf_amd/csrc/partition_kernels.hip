// Radix-partitioned hash aggregation for high-cardinality keys on gfx950.
//
// The reference aggregates every row with device-wide atomics into sparse
// result columns (groupby/hash/groupby_kernels.cuh:98-105) and runs a second
// full pass for VARIANCE (multi_pass_kernels.cuh:74-111).  With ~10 rows per
// group and millions of groups that state does not fit in any on-chip cache,
// so every row costs several random HBM lines.  Here rows are first moved, by
// the top bits of their MurmurHash3 row hash, into bins whose groups fit in
// one workgroup's LDS; each bin is then aggregated entirely with LDS atomics
// (including the variance pass, whose re-read of the bin hits L2) and its
// groups are written straight to the dense output columns.
//
//   k_rpa_hist      rows per final bin (LDS histogram per block)
//   k_rpa_scatter1  input columns -> first-level bins (buffer A, or B when
//                   one level suffices); ranks from LDS atomics, one global
//                   range reservation per (tile, bin)
//   k_rpa_scatter2  first-level bins -> final bins (buffer B)
//   k_rpa_aggregate one workgroup per final bin (partition_aggregate.hip)
#include "partition_common.hpp"

namespace mig {

// ---------------------------------------------------------------------------
// Rows per final bin.  Each block counts its slice of the rows in LDS and
// writes its own histogram row; k_rpa_hist_reduce sums the rows.  The LDS
// holds a window of at most RPA_HIST_WINDOW bins: more bins take one launch
// per window, each reading the keys again.
constexpr int RPA_HIST_WINDOW = 1 << 15;

__global__ void __launch_bounds__(RPA_HIST_BLOCK)
  k_rpa_hist(gb_params const* __restrict__ Pp, rpa_params const* __restrict__ Rp, int64_t n, int window)
{
  extern __shared__ uint32_t hist[];
  gb_params const& P = *Pp;
  rpa_params const& R = *Rp;
  int const B    = 1 << R.log_bins;
  int const W    = B < RPA_HIST_WINDOW ? B : RPA_HIST_WINDOW;
  uint32_t const lo = uint32_t(window) * uint32_t(W);
  for (int i = threadIdx.x; i < W; i += RPA_HIST_BLOCK) hist[i] = 0;
  __syncthreads();
  int64_t const per = (n + gridDim.x - 1) / gridDim.x;
  int64_t const r0  = int64_t(blockIdx.x) * per;
  int64_t const r1  = (r0 + per < n) ? r0 + per : n;
  constexpr int U   = 4;  // independent rows in flight per thread
  for (int64_t row0 = r0 + threadIdx.x; row0 < r1; row0 += int64_t(U) * RPA_HIST_BLOCK) {
    uint32_t bin[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      int64_t const row = row0 + int64_t(u) * RPA_HIST_BLOCK;
      bin[u]            = 0xffffffffu;
      uint64_t bits;
      bool key_null;
      uint32_t h;
      if (row < r1 && input_key(P, row, bits, key_null, h)) bin[u] = top_bits(h, R.log_bins);
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (bin[u] != 0xffffffffu && bin[u] - lo < uint32_t(W))
        __hip_atomic_fetch_add(&hist[bin[u] - lo], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  __syncthreads();
  auto out = gp(R.hist_partial + int64_t(blockIdx.x) * B + lo);
  for (int i = threadIdx.x; i < W; i += RPA_HIST_BLOCK) out[i] = hist[i];
}

// Column sums of the per-block histogram rows: blockIdx.y takes one slice of
// the rows and adds its partial sums into the (zeroed) histogram.
constexpr int RPA_HIST_SLICES = 16;

__global__ void __launch_bounds__(RPA_SMALL_BLOCK) k_rpa_hist_reduce(rpa_params const* __restrict__ Rp)
{
  rpa_params const& R = *Rp;
  int const B         = 1 << R.log_bins;
  int const b         = blockIdx.x * RPA_SMALL_BLOCK + threadIdx.x;
  if (b >= B) return;
  int const per = (R.hist_blocks + RPA_HIST_SLICES - 1) / RPA_HIST_SLICES;
  int const k0  = int(blockIdx.y) * per;
  int const k1  = (k0 + per < R.hist_blocks) ? k0 + per : R.hist_blocks;
  uint32_t s    = 0;
  auto const hp = gp(static_cast<uint32_t const*>(R.hist_partial));
  for (int k = k0; k < k1; k++) s += hp[int64_t(k) * B + b];
  if (s) aadd(R.hist + b, s);
}

// Bin layout from the histogram, on the device: bin_start (exclusive prefix
// sums of the rows per final bin, rows < 2^31), the scatter cursors (each
// bin's start) and the first record of each first-level bin.  One
// workgroup; each thread scans a contiguous run of bins.
constexpr int RPA_LAYOUT_BLOCK = 1024;

__global__ void __launch_bounds__(RPA_LAYOUT_BLOCK) k_rpa_layout(rpa_params const* __restrict__ Rp)
{
  __shared__ uint32_t wave_sum[RPA_LAYOUT_BLOCK / 64];
  rpa_params const& R = *Rp;
  int const B         = 1 << R.log_bins;
  int const F1        = 1 << R.log_f1;
  int const lf2       = R.log_bins - R.log_f1;
  int const tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  auto const hist      = gp(static_cast<uint32_t const*>(R.hist));
  auto const bin_start = gp(R.bin_start);
  int const per        = (B + RPA_LAYOUT_BLOCK - 1) / RPA_LAYOUT_BLOCK;
  int const i0         = tid * per < B ? tid * per : B;
  int const i1         = i0 + per < B ? i0 + per : B;
  uint32_t local       = 0;
  for (int i = i0; i < i1; i++) local += hist[i];
  uint32_t incl = local;  // inclusive scan over the block
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t const y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  if (lane == 63) wave_sum[wave] = incl;
  __syncthreads();
  uint32_t before = 0;
  for (int w = 0; w < wave; w++) before += wave_sum[w];
  uint64_t run = uint64_t(before) + incl - local;
  for (int i = i0; i < i1; i++) {
    bin_start[i]       = int64_t(run);
    gp(R.cursor2)[int64_t(i) * RPA_CURSOR_STRIDE] = run;
    run += hist[i];
  }
  if (i1 == B && i0 < B) bin_start[B] = int64_t(run);
  __syncthreads();
  // first-level bins (F1 <= 256): their first records and scatter cursors
  if (tid < F1) {
    int64_t const lo         = bin_start[int64_t(tid) << lf2];
    gp(R.first_start)[tid]   = lo;
    gp(R.cursor1)[tid * RPA_CURSOR_STRIDE] = uint64_t(lo);
  }
}

// Speculative layout: bins at fixed strides of their capacities.
__global__ void __launch_bounds__(RPA_SMALL_BLOCK) k_rpa_spec_layout(rpa_params const* __restrict__ Rp)
{
  rpa_params const& R = *Rp;
  int64_t const B     = int64_t(1) << R.log_bins;
  int64_t const F1    = int64_t(1) << R.log_f1;
  for (int64_t b = int64_t(blockIdx.x) * RPA_SMALL_BLOCK + threadIdx.x; b <= B; b += int64_t(gridDim.x) * RPA_SMALL_BLOCK) {
    int64_t const st = b * R.cap2;
    gp(R.bin_start)[b] = st;
    if (b < B) gp(R.cursor2)[b * RPA_CURSOR_STRIDE] = uint64_t(st);
    if (b < F1) {
      // one level: the first-level bins are the final bins (buffer B, cap2)
      int64_t const fs      = R.two_level ? b * R.cap1 : st;
      gp(R.first_start)[b] = fs;
      gp(R.cursor1)[b * RPA_CURSOR_STRIDE] = uint64_t(fs);
    }
  }
}

void launch_rpa_spec_layout(rpa_params const* dR, rpa_params const& hR, hipStream_t s)
{
  int64_t const B = (int64_t(1) << hR.log_bins) + 1;
  int64_t blocks  = (B + RPA_SMALL_BLOCK - 1) / RPA_SMALL_BLOCK;
  if (blocks > 1024) blocks = 1024;
  hipLaunchKernelGGL(k_rpa_spec_layout, dim3(blocks), dim3(RPA_SMALL_BLOCK), 0, s, dR);
}

// Rows of first-level bin b1 after scatter 1 (clamped to its capacity).
__device__ __forceinline__ int64_t first_rows(rpa_params const& R, int b1)
{
  int64_t const lo = gp(static_cast<int64_t const*>(R.first_start))[b1];
  int64_t hi       = int64_t(gp(static_cast<unsigned long long const*>(R.cursor1))[int64_t(b1) * RPA_CURSOR_STRIDE]);
  if (hi - lo > R.cap1) hi = lo + R.cap1;
  return hi - lo;
}

// Second-pass tiles: thread b1 takes first-level bin b1, a block scan numbers
// the tiles (F1 <= 256).
__global__ void __launch_bounds__(RPA_SMALL_BLOCK) k_rpa_tiles(rpa_params const* __restrict__ Rp)
{
  __shared__ uint32_t wave_sum[RPA_SMALL_BLOCK / 64];
  rpa_params const& R = *Rp;
  int const F1        = 1 << R.log_f1;
  int const tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  uint32_t tiles = 0;
  if (tid < F1) tiles = uint32_t((first_rows(R, tid) + R.tile_rows - 1) / R.tile_rows);
  uint32_t tincl = tiles;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t const y = __shfl_up(tincl, o, 64);
    if (lane >= o) tincl += y;
  }
  if (lane == 63) wave_sum[wave] = tincl;
  __syncthreads();
  uint32_t tbefore = 0;
  for (int w = 0; w < wave; w++) tbefore += wave_sum[w];
  if (tid < F1) gp(R.tile_start1)[tid] = int64_t(tbefore + tincl - tiles);
  if (tid == F1 - 1) gp(R.tile_start1)[F1] = int64_t(tbefore + tincl);
}

void launch_rpa_tiles(rpa_params const* dR, hipStream_t s)
{
  hipLaunchKernelGGL(k_rpa_tiles, dim3(1), dim3(RPA_SMALL_BLOCK), 0, s, dR);
}

void launch_rpa_layout(rpa_params const* dR, hipStream_t s)
{
  hipLaunchKernelGGL(k_rpa_layout, dim3(1), dim3(RPA_LAYOUT_BLOCK), 0, s, dR);
}

void launch_rpa_hist(gb_params const* dP, rpa_params const* dR, rpa_params const& hR, int64_t n,
                     int num_cus, hipStream_t s)
{
  int64_t const B = int64_t(1) << hR.log_bins;
  int64_t const W = std::min<int64_t>(B, RPA_HIST_WINDOW);
  size_t const lds = sizeof(uint32_t) * size_t(W);
  allow_dynamic_lds(reinterpret_cast<void const*>(&k_rpa_hist), int(lds));
  for (int w = 0; w < int(B / W); w++)
    hipLaunchKernelGGL(k_rpa_hist, dim3(hR.hist_blocks), dim3(RPA_HIST_BLOCK), lds, s, dP, dR, n, w);
  hipLaunchKernelGGL(k_rpa_hist_reduce, dim3((B + RPA_SMALL_BLOCK - 1) / RPA_SMALL_BLOCK, RPA_HIST_SLICES),
                     dim3(RPA_SMALL_BLOCK), 0, s, dR);
}

// ---------------------------------------------------------------------------
__device__ __forceinline__ void copy_width(uint8_t* dst, int64_t di, uint8_t const* src, int64_t si, int w)
{
  switch (w) {
    case 1: gp(dst)[di] = src[si]; break;
    case 2: gp(reinterpret_cast<uint16_t*>(dst))[di] = reinterpret_cast<uint16_t const*>(src)[si]; break;
    case 4: gp(reinterpret_cast<uint32_t*>(dst))[di] = reinterpret_cast<uint32_t const*>(src)[si]; break;
    default: gp(reinterpret_cast<uint64_t*>(dst))[di] = reinterpret_cast<uint64_t const*>(src)[si]; break;
  }
}

constexpr int RPA_MAX_RPT    = RPA_TILE / RPA_SCATTER_BLOCK;  // rows per thread per tile (at most)
constexpr unsigned long long NO_BASE = ~0ull;                // a bin whose rows are not written
constexpr int RPA_MAX_FANOUT = 256;

// Exclusive scan of cnt[0, fan) into off[0, fan] by wave 0 (fan <= 256).
__device__ __forceinline__ void scan_counts(uint32_t const* cnt, uint32_t* off, int fan)
{
  if (threadIdx.x >= 64) return;
  int const lane = threadIdx.x;
  uint32_t v[4];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    int const i = lane * 4 + k;
    v[k]        = i < fan ? cnt[i] : 0u;
    sum += v[k];
  }
  uint32_t incl = sum;
  for (int o = 1; o < 64; o <<= 1) {
    uint32_t y = __shfl_up(incl, o, 64);
    if (lane >= o) incl += y;
  }
  uint32_t run = incl - sum;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    int const i = lane * 4 + k;
    if (i < fan) off[i] = run;
    run += v[k];
  }
  if (lane == 63) off[fan] = incl;
}

constexpr int RPA_PREFETCH_COLS = 2;  // value columns loaded together with the keys

// Value column v of the tile's rows (raw, zero-extended) and, in the first
// pass, their validity; rows with bin < 0 are skipped.
template <bool SECOND>
__device__ __forceinline__ void load_vals(gb_params const& P, rpa_params const& R, int v, int64_t r0,
                                          int const (&bin)[RPA_MAX_RPT], uint64_t (&raw)[RPA_MAX_RPT],
                                          bool (&vv)[RPA_MAX_RPT])
{
  int const w    = storage_width(P.vals[v].storage);
  auto const src = gp(static_cast<uint8_t const*>(SECOND ? R.valsA[v] : P.vals[v].data));
  int const fb   = SECOND ? -1 : R.val_flag_bit[v];
#pragma unroll
  for (int k = 0; k < RPA_MAX_RPT; k++) {
    int64_t const row = r0 + int64_t(k) * RPA_SCATTER_BLOCK + threadIdx.x;
    raw[k]            = 0;
    vv[k]             = false;
    if (bin[k] < 0) continue;
    switch (w) {
      case 1: raw[k] = src[row]; break;
      case 2: raw[k] = reinterpret_cast<gmem<uint16_t const>*>(src)[row]; break;
      case 4: raw[k] = reinterpret_cast<gmem<uint32_t const>*>(src)[row]; break;
      default: raw[k] = reinterpret_cast<gmem<uint64_t const>*>(src)[row]; break;
    }
    if (fb >= 0) vv[k] = dcol_valid(P.vals[v], row);
  }
}

// Stage value column v of the tile at the rows' bin-order positions `lp`,
// and fold its validity into the rows' flags.
template <bool SECOND>
__device__ __forceinline__ void stage_vals(gb_params const& P, rpa_params const& R, uint8_t* stage, int v,
                                           int const (&lp)[RPA_MAX_RPT], uint64_t const (&raw)[RPA_MAX_RPT],
                                           bool const (&vv)[RPA_MAX_RPT], uint8_t (&kf)[RPA_MAX_RPT])
{
  int const w        = storage_width(P.vals[v].storage);
  uint8_t* const dst = stage + R.stage_val_off[v];
  int const fb       = SECOND ? -1 : R.val_flag_bit[v];
#pragma unroll
  for (int k = 0; k < RPA_MAX_RPT; k++) {
    if (lp[k] < 0) continue;
    switch (w) {
      case 1: dst[lp[k]] = uint8_t(raw[k]); break;
      case 2: reinterpret_cast<uint16_t*>(dst)[lp[k]] = uint16_t(raw[k]); break;
      case 4: reinterpret_cast<uint32_t*>(dst)[lp[k]] = uint32_t(raw[k]); break;
      default: reinterpret_cast<uint64_t*>(dst)[lp[k]] = raw[k]; break;
    }
    if (!SECOND && vv[k]) kf[k] |= uint8_t(1u << (fb & 7));
  }
}

// One tile of the scatter: rows are ranked per bin with LDS atomics, staged
// in LDS in bin order, and written out so that consecutive threads store
// consecutive rows of a bin (coalesced runs of tile_rows / fanout rows).
//   SECOND = false: input columns -> first-level bins
//   SECOND = true : first-level bin in buffer A -> final bins in buffer B
template <bool SECOND>
__global__ void __launch_bounds__(RPA_SCATTER_BLOCK) __attribute__((amdgpu_waves_per_eu(6)))
  k_rpa_scatter(gb_params const* __restrict__ Pp, rpa_params const* __restrict__ Rp, int64_t n, int64_t t_begin)
{
  extern __shared__ __attribute__((aligned(16))) uint8_t stage[];
  __shared__ uint32_t cnt[RPA_MAX_FANOUT];
  __shared__ uint32_t toff[RPA_MAX_FANOUT + 1];
  __shared__ unsigned long long base[RPA_MAX_FANOUT];
  __shared__ int64_t tstart[SECOND ? RPA_MAX_FANOUT + 1 : 1];
  gb_params const& P = *Pp;
  rpa_params const& R = *Rp;
  int const T     = R.tile_rows;
  int const rpt   = T / RPA_SCATTER_BLOCK;
  int const lf2   = R.log_bins - R.log_f1;
  int const F1    = 1 << R.log_f1;
  int const fan   = SECOND ? (1 << lf2) : F1;
  auto* skeys     = reinterpret_cast<uint64_t*>(stage);
  auto* sflags    = stage + R.stage_flags_off;
  auto* sbins     = reinterpret_cast<uint16_t*>(stage + R.stage_bin_off);
  auto* srows     = reinterpret_cast<uint32_t*>(stage + R.stage_row_off);  // hashed general keys only
  bool const hashed     = R.general_hashed != 0;
  uint64_t* const okeys = (SECOND || !R.two_level) ? R.keysB : R.keysA;
  uint8_t* const oflags = (SECOND || !R.two_level) ? R.flagsB : R.flagsA;
  uint32_t* const orows = (SECOND || !R.two_level) ? R.rowsB : R.rowsA;
  int64_t ntiles;
  if constexpr (SECOND) {
    for (int i = threadIdx.x; i <= F1; i += RPA_SCATTER_BLOCK)
      tstart[i] = gp(static_cast<int64_t const*>(R.tile_start1))[i];
    lds_barrier();
    ntiles = n < 0 ? tstart[F1] : n;  // tiles [t_begin, n) of the first-level bins (n < 0: all)
  } else {
    ntiles = (n + T - 1) / T;
  }
  for (int64_t t = t_begin + blockIdx.x; t < ntiles; t += gridDim.x) {
    // rows [r0, r1) of this tile; for the second pass `b1` is the first-level bin
    int64_t r0, r1;
    int b1 = 0;
    if constexpr (SECOND) {
      int lo = 0, hi = F1 - 1;
      while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (tstart[mid] <= t) lo = mid;
        else hi = mid - 1;
      }
      b1                 = lo;
      int64_t const bs   = gp(static_cast<int64_t const*>(R.first_start))[b1];
      int64_t const be   = bs + first_rows(R, b1);
      r0                 = bs + (t - tstart[b1]) * T;
      r1                 = (r0 + T < be) ? r0 + T : be;
    } else {
      r0 = t * T;
      r1 = (r0 + T < n) ? r0 + T : n;
    }
    for (int i = threadIdx.x; i < fan; i += RPA_SCATTER_BLOCK) cnt[i] = 0;
    lds_barrier();
    // ---- rank rows per bin
    int bin[RPA_MAX_RPT];
    uint32_t rank[RPA_MAX_RPT];
    uint64_t kb[RPA_MAX_RPT];
    uint8_t kf[RPA_MAX_RPT];
    uint32_t rw[RPA_MAX_RPT];  // input row of each record (hashed general keys)
    // all loads of the tile's keys first (validity and key together), then the ranks
#pragma unroll
    for (int k = 0; k < RPA_MAX_RPT; k++) {
      bin[k] = -1;
      if (k >= rpt) continue;
      int64_t const row = r0 + int64_t(k) * RPA_SCATTER_BLOCK + threadIdx.x;
      if (row >= r1) continue;
      uint32_t h;
      if constexpr (SECOND) {
        kf[k] = gp(static_cast<uint8_t const*>(R.flagsA))[row];
        kb[k] = gp(static_cast<uint64_t const*>(R.keysA))[row];
        if (hashed) rw[k] = gp(static_cast<uint32_t const*>(R.rowsA))[row];
        h     = record_hash(P, kb[k], kf[k]);
        bin[k] = int(top_bits(h, R.log_bins) & uint32_t(fan - 1));
      } else if (P.key_mode == KM_GENERAL) {
        if (hashed) {  // (row hash, second row hash); the row travels beside it
          if (P.skip_null_keys && any_key_null(P, row)) continue;
          uint32_t h2;
          row_hash2(P.keys, P.num_keys, row, SECOND_ROW_HASH_SEED, h, h2);
          kb[k] = (uint64_t(h) << 32) | h2;
          rw[k] = uint32_t(row);
        } else {
          bool kn;
          if (!input_key(P, row, kb[k], kn, h)) continue;
        }
        kf[k]  = 0;
        bin[k] = int(top_bits(h, R.log_f1));
      } else {
        dcol const& kc       = P.keys[0];
        bool const kvalid    = dcol_valid(kc, row);
        uint64_t const bits  = load_fixed_bits(kc, row);
        if (!kvalid && P.skip_null_keys) continue;
        kb[k]  = kvalid ? bits : 0;
        kf[k]  = kvalid ? 0 : RPA_KEY_NULL;
        h      = kvalid ? inline_key_hash(kc.storage, bits) : null_key_hash();
        bin[k] = int(top_bits(h, R.log_f1));
      }
    }
    // the first value columns' loads go out with the keys' (their latency
    // overlaps the ranking, the scan and the range reservation)
    int const ncols = R.plan_ncols;
    uint64_t praw[RPA_PREFETCH_COLS][RPA_MAX_RPT];
    bool pvv[RPA_PREFETCH_COLS][RPA_MAX_RPT];
#pragma unroll
    for (int i = 0; i < RPA_PREFETCH_COLS; i++)
      if (i < ncols) load_vals<SECOND>(P, R, R.plan_col[i], r0, bin, praw[i], pvv[i]);
#pragma unroll
    for (int k = 0; k < RPA_MAX_RPT; k++)
      if (bin[k] >= 0)
        rank[k] = __hip_atomic_fetch_add(&cnt[bin[k]], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    lds_barrier();
    scan_counts(cnt, toff, fan);
    for (int i = threadIdx.x; i < fan; i += RPA_SCATTER_BLOCK) {
      if (cnt[i]) {
        int64_t const fb        = (int64_t(b1) << lf2) + i;
        unsigned long long* cur = SECOND ? R.cursor2 + fb * RPA_CURSOR_STRIDE : R.cursor1 + int64_t(i) * RPA_CURSOR_STRIDE;
        unsigned long long const b0 = __hip_atomic_fetch_add(gp(cur), (unsigned long long)cnt[i], __ATOMIC_RELAXED,
                                                             __HIP_MEMORY_SCOPE_AGENT);
        // a bin of a speculative layout may fill up: its rows that do not
        // fit are not written, and the call is redone with an exact layout
        int64_t const start = SECOND ? gp(static_cast<int64_t const*>(R.bin_start))[fb]
                                     : gp(static_cast<int64_t const*>(R.first_start))[i];
        int64_t const cap   = (SECOND || !R.two_level) ? R.cap2 : R.cap1;
        if (int64_t(b0) + int64_t(cnt[i]) > start + cap) {
          astore(R.cap_overflow, 1);
          base[i] = NO_BASE;
        } else {
          base[i] = b0;
        }
      }
    }
    lds_barrier();
    // ---- stage the tile in bin order
    int lp[RPA_MAX_RPT];
#pragma unroll
    for (int k = 0; k < RPA_MAX_RPT; k++) {
      lp[k] = bin[k] < 0 ? -1 : int(toff[bin[k]] + rank[k]);
      if (lp[k] < 0) continue;
      skeys[lp[k]] = kb[k];
      sbins[lp[k]] = uint16_t(bin[k]);
      if (hashed) srows[lp[k]] = rw[k];
    }
    // values: the prefetched columns, then any others column by column
#pragma unroll
    for (int i = 0; i < RPA_PREFETCH_COLS; i++)
      if (i < ncols) stage_vals<SECOND>(P, R, stage, R.plan_col[i], lp, praw[i], pvv[i], kf);
    for (int i = RPA_PREFETCH_COLS; i < ncols; i++) {
      int const v = R.plan_col[i];
      uint64_t raw[RPA_MAX_RPT];
      bool vv[RPA_MAX_RPT];
      load_vals<SECOND>(P, R, v, r0, bin, raw, vv);
      stage_vals<SECOND>(P, R, stage, v, lp, raw, vv, kf);
    }
#pragma unroll
    for (int k = 0; k < RPA_MAX_RPT; k++)
      if (lp[k] >= 0) sflags[lp[k]] = kf[k];
    lds_barrier();
    // ---- write bin runs
    int const total = int(toff[fan]);
    for (int lp = threadIdx.x; lp < total; lp += RPA_SCATTER_BLOCK) {
      int const b       = sbins[lp];
      if (base[b] == NO_BASE) continue;
      int64_t const dst = int64_t(base[b]) + (lp - int(toff[b]));
      gp(okeys)[dst]    = skeys[lp];
      gp(oflags)[dst]   = sflags[lp];
      if (hashed) gp(orows)[dst] = srows[lp];
      for (int i = 0; i < ncols; i++) {
        int const v       = R.plan_col[i];
        uint8_t* const ov = (SECOND || !R.two_level) ? R.valsB[v] : R.valsA[v];
        copy_width(ov, dst, stage + R.stage_val_off[v], lp, storage_width(P.vals[v].storage));
      }
    }
    lds_barrier();
  }
}

// Dynamic LDS above 64 KB must be opted into per kernel (gfx950 has 160 KB).
static void allow_large_lds(void const* fn, int bytes) { allow_dynamic_lds(fn, bytes); }

// Workgroups of a scatter kernel that are resident at once: the tiles are
// strided over the grid, so a grid larger than what fits would leave its last
// workgroups to run after the others.
static int64_t resident_grid(void const* fn, int dyn_lds, int num_cus)
{
  int per_cu = 0;  // registers and LDS both limit it
  MIG_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, RPA_SCATTER_BLOCK, size_t(dyn_lds)));
  if (per_cu < 1) per_cu = 1;
  return int64_t(num_cus) * per_cu;
}

void launch_rpa_scatter1(gb_params const* dP, rpa_params const* dR, rpa_params const& hR, int64_t n,
                         int num_cus, hipStream_t s)
{
  void const* fn = reinterpret_cast<void const*>(&k_rpa_scatter<false>);
  int64_t tiles  = (n + hR.tile_rows - 1) / hR.tile_rows;
  int64_t blocks = resident_grid(fn, hR.stage_bytes, num_cus);
  if (blocks > tiles) blocks = tiles;
  if (blocks < 1) blocks = 1;
  allow_large_lds(fn, hR.stage_bytes);
  hipLaunchKernelGGL(k_rpa_scatter<false>, dim3(blocks), dim3(RPA_SCATTER_BLOCK), hR.stage_bytes, s, dP, dR, n,
                     int64_t(0));
}

void launch_rpa_scatter2(gb_params const* dP, rpa_params const* dR, rpa_params const& hR, int num_cus,
                         int64_t tile_begin, int64_t tile_end, hipStream_t s)
{
  if (tile_end >= 0 && tile_end <= tile_begin) return;
  int64_t blocks = resident_grid(reinterpret_cast<void const*>(&k_rpa_scatter<true>), hR.stage_bytes, num_cus);
  if (tile_end >= 0 && blocks > tile_end - tile_begin) blocks = tile_end - tile_begin;
  allow_large_lds(reinterpret_cast<void const*>(&k_rpa_scatter<true>), hR.stage_bytes);
  hipLaunchKernelGGL(k_rpa_scatter<true>, dim3(blocks), dim3(RPA_SCATTER_BLOCK), hR.stage_bytes, s, dP, dR,
                     tile_end, tile_begin);
}

}  // namespace mig
