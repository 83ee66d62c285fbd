// kmst_kernels.hip -- hand-written CDNA4 (gfx950) kernels of the Boruvka /
// filter-Boruvka MST path.  All work is integer and HBM-bound; no MFMA.
//
// Record layout (structure of arrays, 16 bytes per directed record):
//   src[i], dst[i]  u32 endpoint ids in the current (compacted) numbering
//   val[i]          u64 strict key (weight << 32) | uid, uid = input position
//                   of the (min, max) twin; VAL_NONE marks a dead record.
#include <type_traits>
#include "kmst_kernels.hpp"

namespace kmst {

// ============================================================== fills
__global__ __launch_bounds__(BLOCK) void fill_u64_kernel(u64* p, u64 n, u64 value) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += u64(gridDim.x) * BLOCK) p[i] = value;
}
__global__ __launch_bounds__(BLOCK) void fill_u32_kernel(u32* p, u64 n, u32 value) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += u64(gridDim.x) * BLOCK) p[i] = value;
}

static u32 grid_for(u64 n, u32 cap = 65536) {
  const u64 g = (n + BLOCK - 1) / BLOCK;
  return static_cast<u32>(g == 0 ? 1 : (g > cap ? cap : g));
}

void fill_u64(hipStream_t st, u64* p, u64 n, u64 value) {
  if (n == 0) return;
  KTScope t(st, K_FILL, 8.0 * n);
  fill_u64_kernel<<<grid_for(n), BLOCK, 0, st>>>(p, n, value);
  KMST_HIP(hipGetLastError());
}
void fill_u32(hipStream_t st, u32* p, u64 n, u32 value) {
  if (n == 0) return;
  KTScope t(st, K_FILL, 4.0 * n);
  fill_u32_kernel<<<grid_for(n), BLOCK, 0, st>>>(p, n, value);
  KMST_HIP(hipGetLastError());
}

// Counters (and optionally a block of staged words) written straight into
// coherent pinned host memory, then the sequence word with a system-scope
// release: the host spins on the sequence word instead of a device-to-host
// copy and a stream synchronisation.  One workgroup; every store is a vector
// store.
__global__ __launch_bounds__(BLOCK) void publish_kernel(const u32* __restrict__ ctr, u32 ctr_words,
                                                        const u32* __restrict__ extra,
                                                        u32 extra_words, u32* host,
                                                        u32 extra_at, u32 seq_at, u32 seq) {
  for (u32 i = threadIdx.x; i < ctr_words; i += BLOCK) host[i] = ctr[i];
  for (u32 i = threadIdx.x; i < extra_words; i += BLOCK) host[extra_at + i] = extra[i];
  __threadfence_system();
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(host + seq_at, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void publish_words(hipStream_t st, const u32* ctr, u32 ctr_words, const u32* extra,
                   u32 extra_words, u32* host, u32 extra_at, u32 seq_at, u32 seq) {
  publish_kernel<<<1, BLOCK, 0, st>>>(ctr, ctr_words, extra, extra_words, host, extra_at, seq_at,
                                      seq);
  KMST_HIP(hipGetLastError());
}

// ============================================================== scan
// Three-phase exclusive scan: per-tile reduce, scan of tile sums (recursive),
// per-tile downsweep.  Loads and stores are striped (element k * BLOCK + t of
// a tile for thread t, coalesced); the downsweep turns them into runs of
// SCAN_IPT consecutive elements per thread through LDS, padded by one word
// every SCAN_IPT elements so that both orders are free of bank conflicts.
template <typename T>
__global__ __launch_bounds__(BLOCK) void scan_reduce_kernel(const T* __restrict__ in, u64 n,
                                                            T* __restrict__ partial) {
  __shared__ T lds[WAVES];
  const u64 base = u64(blockIdx.x) * SCAN_TILE + threadIdx.x;
  T sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_IPT; ++k) {
    const u64 i = base + u64(k) * BLOCK;
    sum += (i < n) ? in[i] : T(0);
  }
  T total;
  block_exclusive_sum(sum, total, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

__device__ __forceinline__ u32 scan_pad(u32 i) { return i + i / SCAN_IPT; }

// With flagged_low, element i also gets flagged_low[i] = low 32 bits of its
// exclusive sum when bit 0 of in[i] is set, NONE32 otherwise (the new id of
// a root in the hook state scan, one 4-byte gather later instead of two
// 8-byte ones).  With tag_only only flagged elements write, their low 32 bits
// tagged with ROOT_TAG (into the hook's parent array: the root's new id rides
// on the load that finds the root).
template <typename T>
__global__ __launch_bounds__(BLOCK) void scan_down_kernel(const T* in, T* out, u64 n,
                                                          const T* __restrict__ partial,
                                                          T* total_dev, u32* flagged_low,
                                                          bool tag_only) {
  __shared__ T lds[WAVES];
  __shared__ T tile[SCAN_TILE + SCAN_TILE / SCAN_IPT];
  const u64 base = u64(blockIdx.x) * SCAN_TILE + threadIdx.x;
  T v[SCAN_IPT];
#pragma unroll
  for (int k = 0; k < SCAN_IPT; ++k) {
    const u64 i = base + u64(k) * BLOCK;
    v[k] = (i < n) ? in[i] : T(0);
    tile[scan_pad(k * BLOCK + threadIdx.x)] = v[k];
  }
  __syncthreads();
  T b[SCAN_IPT];
  T sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_IPT; ++k) {
    b[k] = tile[scan_pad(threadIdx.x * SCAN_IPT + k)];
    sum += b[k];
  }
  T total;
  T run = block_exclusive_sum(sum, total, lds);  // also orders the reads above before the writes
  const T tile_off = partial ? partial[blockIdx.x] : T(0);
  run += tile_off;
#pragma unroll
  for (int k = 0; k < SCAN_IPT; ++k) {
    tile[scan_pad(threadIdx.x * SCAN_IPT + k)] = run;
    run += b[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SCAN_IPT; ++k) {
    const u64 i = base + u64(k) * BLOCK;
    if (i < n) {
      const T o = tile[scan_pad(k * BLOCK + threadIdx.x)];
      out[i] = o;
      if (flagged_low) {
        if (!tag_only) flagged_low[i] = (v[k] & 1) ? static_cast<u32>(o) : NONE32;
        else if (v[k] & 1) flagged_low[i] = static_cast<u32>(o) | ROOT_TAG;
      }
    }
  }
  if (total_dev && partial == nullptr && threadIdx.x == 0 && blockIdx.x == gridDim.x - 1)
    *total_dev = tile_off + total;
}

u64 scan_scratch_elems(u64 n) {
  u64 s = 0;
  while (n > SCAN_TILE) {
    n = (n + SCAN_TILE - 1) / SCAN_TILE;
    s += n;
  }
  return s + 1;
}

template <typename T>
void scan_exclusive(hipStream_t st, const T* in, T* out, u64 n, T* scratch, T* total_dev,
                    u32* flagged_low = nullptr, bool tag_only = false) {
  if (n == 0) {
    if (total_dev) {
      if (sizeof(T) == 8) fill_u64(st, reinterpret_cast<u64*>(total_dev), 1, 0);
      else fill_u32(st, reinterpret_cast<u32*>(total_dev), 1, 0);
    }
    return;
  }
  const u32 tiles = ceil_div_u32(n, SCAN_TILE);
  if (tiles == 1) {
    KTScope t(st, K_SCAN_DOWN, 2.0 * sizeof(T) * n);
    scan_down_kernel<T><<<1, BLOCK, 0, st>>>(in, out, n, nullptr, total_dev, flagged_low,
                                             tag_only);
    KMST_HIP(hipGetLastError());
    return;
  }
  T* partial = scratch;
  {
    KTScope t(st, K_SCAN_REDUCE, 1.0 * sizeof(T) * n);
    scan_reduce_kernel<T><<<tiles, BLOCK, 0, st>>>(in, n, partial);
    KMST_HIP(hipGetLastError());
  }
  scan_exclusive<T>(st, partial, partial, tiles, scratch + tiles, total_dev);
  KTScope t(st, K_SCAN_DOWN, 2.0 * sizeof(T) * n);
  scan_down_kernel<T><<<tiles, BLOCK, 0, st>>>(in, out, n, partial, nullptr, flagged_low,
                                               tag_only);
  KMST_HIP(hipGetLastError());
}

void scan_exclusive_u32(hipStream_t st, const u32* in, u32* out, u64 n, u32* scratch,
                        u32* total_dev) {
  scan_exclusive<u32>(st, in, out, n, scratch, total_dev);
}
void scan_exclusive_u64(hipStream_t st, const u64* in, u64* out, u64 n, u64* scratch,
                        u64* total_dev, u32* flagged_low, bool tag_only) {
  if (tag_only && n > ROOT_TAG) throw Error(KMST_EINVAL, "root tags need n <= 2^31");
  scan_exclusive<u64>(st, in, out, n, scratch, total_dev, flagged_low, tag_only);
}

// ============================================================== setup
// Contract check of the reference's assumed input: sorted by (src, dst), no
// self-loops, ids < n (include/util/utils.hpp:391-395).  Also gathers the
// largest weight (to size the packed key), the number of min<max records,
// and hash sums of (min, max, w) over min<max and over max<min records:
// the input is symmetric as a multiset iff both counts agree and the hash
// sums agree, up to a collision of the sums.
// The hash costs two 32 x 32 -> 64 multiply-adds (the passes that take it
// are bound by instruction issue): the first mixes the pair non-linearly, its
// addend (b, a) keeping distinct pairs apart where a factor is zero; the
// second folds the weight in, so that for a fixed pair the hash is injective
// in w (a reverse record of another weight always changes the sums).
__device__ __forceinline__ u64 sym_hash(u32 a, u32 b, u32 w) {
  const u64 p = u64(a ^ 0x9E3779B9u) * u64(b ^ 0x85EBCA77u) + ((u64(b) << 32) | a);
  const u32 c = static_cast<u32>(p) ^ static_cast<u32>(p >> 32) ^ w;
  return u64(c) * 0xC2B2AE3Du + p;
}

__global__ __launch_bounds__(BLOCK) void fingerprint_kernel(const u32* __restrict__ s,
                                                          const u32* __restrict__ d,
                                                          const u32* __restrict__ w, u64 count,
                                                          u64* out) {
  __shared__ u64 lds[WAVES];
  u64 h = 0;
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < count; i += u64(gridDim.x) * BLOCK)
    h += sym_hash(s[i], d[i], w[i]);
  u64 total;
  block_exclusive_sum(h, total, lds);
  if (threadIdx.x == 0 && total)
    atomicAdd(reinterpret_cast<unsigned long long*>(out), static_cast<unsigned long long>(total));
}

void edge_fingerprint(hipStream_t st, const u32* s, const u32* d, const u32* w, u64 count,
                      u64* out) {
  if (count == 0) return;
  KTScope t(st, K_OTHER, 12.0 * count);
  fingerprint_kernel<<<grid_for(count, 1024), BLOCK, 0, st>>>(s, d, w, count, out);
  KMST_HIP(hipGetLastError());
}

// Weight class of w for ascending pivots q[0..K): the number of pivots below
// w (0 = lightest class, K = heavier than every pivot).
__device__ __forceinline__ u32 weight_class(u32 w, const WeightPivots& P) {
  u32 c = 0;
#pragma unroll
  for (u32 j = 0; j < MAX_PIVOTS; ++j) c += (j < P.K && w > P.q[j]) ? 1u : 0u;
  return c;
}

// Sums of one column of a per-wave reduction table (block-wide).
template <int COLS>
__device__ __forceinline__ u64 s_red_sum(const u64 (&t)[WAVES][COLS], int col) {
  u64 x = 0;
#pragma unroll
  for (int w = 0; w < WAVES; ++w) x += t[w][col];
  return x;
}
template <int COLS>
__device__ __forceinline__ u64 s_red_total(const u64 (&t)[WAVES][COLS]) {
  return s_red_sum(t, 3);  // column 3 holds the record counts
}

// Records are taken in units of 256: unit u is chunk 64u .. 64u + 63 of four
// consecutive records, one chunk per lane of a wave.  Validation writes each
// unit's count of the two lightest weight classes, so that the class scatter
// can place every unit without waiting on other units.
constexpr u64 UNIT_RECORDS = 4 * WAVE;
__device__ __forceinline__ u64 record_units_dev(u64 m) {
  return (m + UNIT_RECORDS - 1) / UNIT_RECORDS;
}

template <bool VEC>
__global__ __launch_bounds__(BLOCK) void validate_kernel(const u32* __restrict__ src,
                                                         const u32* __restrict__ dst,
                                                         const u32* __restrict__ wt, u64 m,
                                                         u64 n, u32* err, u32* fwd_count,
                                                         u32* max_w, u64* hsum,
                                                         const WeightPivots P,
                                                         u64* class_counts,
                                                         u32* __restrict__ seg_start,
                                                         u32* __restrict__ seg_end, u32 q0,
                                                         u32 q1, u32* __restrict__ unit_c0,
                                                         u32* __restrict__ unit_c1) {
  __shared__ u64 s_red[WAVES][MAX_PIVOTS + 4];
  const int lane = lane_id();
  u32 bad = 0, fwd = 0, mw = 0, seen = 0;
  u64 hf = 0, hr = 0;
  u32 above[MAX_PIVOTS];  // records with w > q[j]
#pragma unroll
  for (int j = 0; j < MAX_PIVOTS; ++j) above[j] = 0;
  auto one = [&](u64 i, u32 s, u32 d, u32 w, u32 ps, u32 pd, bool has_prev) {
    if (s >= n || d >= n || s == d) bad |= ERR_INVALID;
    if (has_prev && (ps > s || (ps == s && pd > d))) bad |= ERR_UNSORTED;
    if (seg_end && (!has_prev || ps != s)) {
      // record i opens s's range and closes the previous vertex's
      if (s < n) seg_start[s] = static_cast<u32>(i);
      if (has_prev && ps < n) seg_end[ps] = static_cast<u32>(i);
    }
    if (seg_end && i + 1 == m && s < n) seg_end[s] = static_cast<u32>(m);
    mw = w > mw ? w : mw;
    ++seen;
    const bool f = s < d;
    const u64 h = sym_hash(f ? s : d, f ? d : s, w);
    fwd += f ? 1u : 0u;
    hf += f ? h : 0ull;
    hr += f ? 0ull : h;
#pragma unroll
    for (u32 j = 0; j < MAX_PIVOTS; ++j)
      if (j < P.K) above[j] += (w > P.q[j]) ? 1u : 0u;
  };
  auto cls_of = [&](u32 w) {  // class 0 / class 1 as 16-bit fields
    return (w <= q0 ? 1u : 0u) + ((w > q0 && w <= q1) ? 0x10000u : 0u);
  };
  const u64 full = m / 4, chunks = (m + 3) / 4;
  if (VEC) {
    // chunks of 4 consecutive records (16-byte loads); consecutive lanes take
    // consecutive chunks, so the record before a chunk is the previous lane's
    // last one.  Two chunks per thread and iteration are loaded before use.
    // A unit is the 64 chunks of one wave in one iteration.
    constexpr int VU = 2;
    const u64 stride = u64(gridDim.x) * BLOCK;
    for (u64 base = u64(blockIdx.x) * BLOCK; base < chunks; base += VU * stride) {  // uniform
      uint4 vs[VU], vd[VU], vw[VU];
#pragma unroll
      for (int u = 0; u < VU; ++u) {
        const u64 ch = min(base + u * stride + threadIdx.x, full - 1);
        vs[u] = reinterpret_cast<const uint4*>(src)[ch];
        vd[u] = reinterpret_cast<const uint4*>(dst)[ch];
        vw[u] = reinterpret_cast<const uint4*>(wt)[ch];
      }
#pragma unroll
      for (int u = 0; u < VU; ++u) {
        const u64 ch = base + u * stride + threadIdx.x;
        u32 ps = __shfl_up(vs[u].w, 1, WAVE), pd = __shfl_up(vd[u].w, 1, WAVE);
        u32 cls = 0;
        if (ch < full) {
          if (lane == 0 && ch > 0) {
            ps = src[ch * 4 - 1];
            pd = dst[ch * 4 - 1];
          }
          one(ch * 4, vs[u].x, vd[u].x, vw[u].x, ps, pd, ch > 0);
          one(ch * 4 + 1, vs[u].y, vd[u].y, vw[u].y, vs[u].x, vd[u].x, true);
          one(ch * 4 + 2, vs[u].z, vd[u].z, vw[u].z, vs[u].y, vd[u].y, true);
          one(ch * 4 + 3, vs[u].w, vd[u].w, vw[u].w, vs[u].z, vd[u].z, true);
          cls = cls_of(vw[u].x) + cls_of(vw[u].y) + cls_of(vw[u].z) + cls_of(vw[u].w);
        } else if (ch == full && unit_c0) {
          // the last, partial chunk: counted here, checked below
          for (u64 i = full * 4; i < m; ++i) cls += cls_of(wt[i]);
        }
        if (unit_c0) {  // uniform
          // per-lane counts are 0..4: sum them bit by bit with ballots
          const u32 c0 = cls & 0xFFFFu, c1 = cls >> 16;
          u32 t0 = 0, t1 = 0;
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            t0 += static_cast<u32>(__popcll(__ballot((c0 >> b) & 1u))) << b;
            t1 += static_cast<u32>(__popcll(__ballot((c1 >> b) & 1u))) << b;
          }
          const u64 unit_chunk = base + u * stride + u64(wave_id()) * WAVE;
          if (lane == 0 && unit_chunk < chunks) {
            unit_c0[unit_chunk / WAVE] = t0;
            unit_c1[unit_chunk / WAVE] = t1;
          }
        }
      }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      for (u64 i = full * 4; i < m; ++i)
        one(i, src[i], dst[i], wt[i], i ? src[i - 1] : 0u, i ? dst[i - 1] : 0u, i > 0);
    }
  } else {
    // unaligned arrays: one record per thread (no unit counts; see
    // unit_class_counts)
    for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < m; i += u64(gridDim.x) * BLOCK)
      one(i, src[i], dst[i], wt[i], i ? src[i - 1] : 0u, i ? dst[i - 1] : 0u, i > 0);
  }
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    fwd += __shfl_xor(fwd, o, WAVE);
    seen += __shfl_xor(seen, o, WAVE);
    const u32 om = __shfl_xor(mw, o, WAVE);
    mw = om > mw ? om : mw;
    hf += __shfl_xor(hf, o, WAVE);
    hr += __shfl_xor(hr, o, WAVE);
#pragma unroll
    for (u32 j = 0; j < MAX_PIVOTS; ++j)
      if (j < P.K) above[j] += __shfl_xor(above[j], o, WAVE);
  }
  const int w_ = wave_id();
  if (lane == 0) {
    s_red[w_][0] = fwd;
    s_red[w_][1] = hf;
    s_red[w_][2] = hr;
    s_red[w_][3] = seen;
    for (int j = 0; j < MAX_PIVOTS; ++j) s_red[w_][4 + j] = above[j];
    if (mw) atomicMax(max_w, mw);
  }
  if (bad) atomicOr(err, bad);
  __syncthreads();
  if (threadIdx.x < MAX_PIVOTS + 4) {
    u64 t = 0;
    for (int ww = 0; ww < WAVES; ++ww) t += s_red[ww][threadIdx.x];
    const u32 x = threadIdx.x;
    if (x == 0) {
      if (t) atomicAdd(fwd_count, static_cast<u32>(t));
    } else if (x <= 2) {
      if (t) atomicAdd(reinterpret_cast<unsigned long long*>(hsum + x - 1),
                       static_cast<unsigned long long>(t));
    } else if (P.K) {
      // class c counts: total - above[0], above[c-1] - above[c], above[K-1]
      // (wrapping u64 arithmetic; the block's contributions sum correctly)
      const u32 c = x - 3;  // 0..MAX_PIVOTS
      if (c <= P.K) {
        const u64 lo = (c == 0) ? s_red_total(s_red) : s_red_sum(s_red, 4 + c - 1);
        const u64 hi = (c == P.K) ? 0ull : s_red_sum(s_red, 4 + c);
        const u64 v = lo - hi;
        if (v) atomicAdd(reinterpret_cast<unsigned long long*>(class_counts + c),
                         static_cast<unsigned long long>(v));
      }
    }
  }
}

u64 record_units(u64 m) { return (m + UNIT_RECORDS - 1) / UNIT_RECORDS; }

bool validate_records(hipStream_t st, const u32* src, const u32* dst, const u32* w, u64 m, u64 n,
                      u32* err, u32* fwd_count, u32* max_w, u64* hsum, const WeightPivots& P,
                      u64* class_counts, u32* seg_start, u32* seg_end, u32 q0, u32 q1,
                      u32* unit_c0, u32* unit_c1) {
  if (m == 0) return false;
  KTScope t(st, K_VALIDATE, 12.0 * m + (unit_c0 ? 8.0 * record_units(m) : 0.0));
  const bool vec = m >= 4 && ((reinterpret_cast<uintptr_t>(src) |
                               reinterpret_cast<uintptr_t>(dst) |
                               reinterpret_cast<uintptr_t>(w)) & 15) == 0;
  const u32 grid = grid_for((m + 3) / 4, 4096);
  if (vec)
    validate_kernel<true><<<grid, BLOCK, 0, st>>>(src, dst, w, m, n, err, fwd_count, max_w, hsum,
                                                  P, class_counts, seg_start, seg_end, q0, q1,
                                                  unit_c0, unit_c1);
  else
    validate_kernel<false><<<grid, BLOCK, 0, st>>>(src, dst, w, m, n, err, fwd_count, max_w, hsum,
                                                   P, class_counts, seg_start, seg_end, q0, q1,
                                                   nullptr, nullptr);
  KMST_HIP(hipGetLastError());
  return vec && unit_c0;
}

// Class 0 / class 1 counts of every 256-record unit from the weights alone
// (for inputs whose arrays are not 16-byte aligned).
__global__ __launch_bounds__(BLOCK) void unit_class_counts_kernel(const u32* __restrict__ wt,
                                                                  u64 m, u32 q0, u32 q1,
                                                                  u32* __restrict__ unit_c0,
                                                                  u32* __restrict__ unit_c1) {
  const int lane = lane_id();
  const u64 units = record_units_dev(m);
  const u64 wstride = u64(gridDim.x) * WAVES;
  for (u64 unit = u64(blockIdx.x) * WAVES + wave_id(); unit < units; unit += wstride) {
    u32 cls = 0;
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const u64 i = unit * UNIT_RECORDS + u64(e) * WAVE + lane;
      if (i < m) {
        const u32 w = wt[i];
        cls += (w <= q0 ? 1u : 0u) + ((w > q0 && w <= q1) ? 0x10000u : 0u);
      }
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) cls += __shfl_xor(cls, o, WAVE);
    if (lane == 0) {
      unit_c0[unit] = cls & 0xFFFFu;
      unit_c1[unit] = cls >> 16;
    }
  }
}

void unit_class_counts(hipStream_t st, const u32* w, u64 m, u32 q0, u32 q1, u32* unit_c0,
                       u32* unit_c1) {
  if (m == 0) return;
  KTScope t(st, K_OTHER, 4.0 * m + 8.0 * record_units(m));
  const u32 grid = static_cast<u32>(std::min<u64>((record_units(m) + WAVES - 1) / WAVES, 8192));
  unit_class_counts_kernel<<<grid, BLOCK, 0, st>>>(w, m, q0, q1, unit_c0, unit_c1);
  KMST_HIP(hipGetLastError());
}

// Weights-only pass: the largest weight, the record count of every weight
// class and the class 0 / class 1 counts of every 256-record unit.  A wave
// takes WC_UNITS consecutive units per step (lane l: records 4l .. 4l + 3 of
// each, one 16-byte load), loads them all before counting, and writes their
// unit counts with one store.  The pass is bound by instruction issue, not by
// HBM, unless it stays lean: a step whose units are all in range counts
// classes 0 and 1 with one compare and a ballot per 64 records (the counts
// stay wave-uniform), and only the pivots that exist are compared.
constexpr int WC_UNITS = 4;

template <bool VEC>
__global__ __launch_bounds__(BLOCK) void weight_classes_kernel(
    const u32* __restrict__ wt, u64 m, const WeightPivots P, u32 q0, u32 q1,
    u64* class_counts, u32* max_w, u32* __restrict__ unit_c0, u32* __restrict__ unit_c1) {
  __shared__ u64 s_red[WAVES][MAX_PIVOTS + 4];
  const int lane = lane_id();
  const u64 units = record_units_dev(m);
  const u64 steps = (units + WC_UNITS - 1) / WC_UNITS;
  const u64 wstride = u64(gridDim.x) * WAVES;
  // per lane: the largest weight, and (partial steps) records seen and
  // class 0 / class 1 totals; wave-uniform: the same totals of full steps
  u32 mw = 0, seen = 0, tot0 = 0, tot1 = 0;
  u32 useen = 0, utot0 = 0, utot1 = 0;
  // above[j]: records with w > q[j]; for j < 2 they follow from the class 0
  // and class 1 totals
  u32 above[MAX_PIVOTS];
#pragma unroll
  for (int j = 0; j < MAX_PIVOTS; ++j) above[j] = 0;
  for (u64 step = u64(blockIdx.x) * WAVES + wave_id(); step < steps; step += wstride) {
    const u64 u0 = step * WC_UNITS;
    u32 w4[WC_UNITS][4];
    if (VEC && (u0 + WC_UNITS) * UNIT_RECORDS <= m) {  // wave-uniform
#pragma unroll
      for (int u = 0; u < WC_UNITS; ++u) {
        const uint4 a =
            *reinterpret_cast<const uint4*>(wt + (u0 + u) * UNIT_RECORDS + 4 * u64(lane));
        w4[u][0] = a.x; w4[u][1] = a.y; w4[u][2] = a.z; w4[u][3] = a.w;
      }
      u32 mine0 = 0, mine1 = 0;  // lane u < WC_UNITS keeps unit u0 + u's counts
#pragma unroll
      for (int u = 0; u < WC_UNITS; ++u) {
        u32 t0 = 0, t01 = 0;  // wave-uniform
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const u32 w = w4[u][e];
          mw = w > mw ? w : mw;
          t0 += static_cast<u32>(__popcll(__ballot(w <= q0)));
          t01 += static_cast<u32>(__popcll(__ballot(w <= q1)));
        }
        if (lane == u) {
          mine0 = t0;
          mine1 = t01 - t0;
        }
        utot0 += t0;
        utot1 += t01 - t0;
      }
#pragma unroll
      for (u32 j = 2; j < MAX_PIVOTS; ++j) {
        if (j < P.K) {  // uniform
#pragma unroll
          for (int u = 0; u < WC_UNITS; ++u)
#pragma unroll
            for (int e = 0; e < 4; ++e) above[j] += (w4[u][e] > P.q[j]) ? 1u : 0u;
        }
      }
      useen += WC_UNITS * static_cast<u32>(UNIT_RECORDS);
      if (lane < WC_UNITS) {
        unit_c0[u0 + lane] = mine0;
        unit_c1[u0 + lane] = mine1;
      }
      continue;
    }
    // a step past the end of the records (or unaligned weights)
#pragma unroll
    for (int u = 0; u < WC_UNITS; ++u) {
      const u64 r0 = (u0 + u) * UNIT_RECORDS + 4 * u64(lane);
      if (VEC && r0 + 3 < m) {
        const uint4 a = *reinterpret_cast<const uint4*>(wt + r0);
        w4[u][0] = a.x; w4[u][1] = a.y; w4[u][2] = a.z; w4[u][3] = a.w;
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) w4[u][e] = (r0 + e < m) ? wt[r0 + e] : 0u;
      }
    }
    u32 mine0 = 0, mine1 = 0;
#pragma unroll
    for (int u = 0; u < WC_UNITS; ++u) {
      const u64 unit = u0 + u;  // wave-uniform
      if (unit >= units) break;
      const u64 r0 = unit * UNIT_RECORDS + 4 * u64(lane);
      u32 c0 = 0, c1 = 0;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (r0 + e >= m) continue;
        const u32 w = w4[u][e];
        mw = w > mw ? w : mw;
        ++seen;
        c0 += w <= q0 ? 1u : 0u;
        c1 += (w > q0 && w <= q1) ? 1u : 0u;
#pragma unroll
        for (u32 j = 2; j < MAX_PIVOTS; ++j)
          if (j < P.K) above[j] += (w > P.q[j]) ? 1u : 0u;
      }
      tot0 += c0;
      tot1 += c1;
      // per-lane counts are 0..4: sum them bit by bit with ballots
      u32 t0 = 0, t1 = 0;
#pragma unroll
      for (int b = 0; b < 3; ++b) {
        t0 += static_cast<u32>(__popcll(__ballot((c0 >> b) & 1u))) << b;
        t1 += static_cast<u32>(__popcll(__ballot((c1 >> b) & 1u))) << b;
      }
      if (lane == u) {
        mine0 = t0;
        mine1 = t1;
      }
    }
    if (lane < WC_UNITS && u0 + lane < units) {
      unit_c0[u0 + lane] = mine0;
      unit_c1[u0 + lane] = mine1;
    }
  }
  // the wave-uniform totals count once, on lane 0
  if (lane == 0) {
    seen += useen;
    tot0 += utot0;
    tot1 += utot1;
  }
  if (P.K >= 1) above[0] = seen - tot0;
  if (P.K >= 2) above[1] = seen - tot0 - tot1;
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    seen += __shfl_xor(seen, o, WAVE);
    const u32 om = __shfl_xor(mw, o, WAVE);
    mw = om > mw ? om : mw;
#pragma unroll
    for (u32 j = 0; j < MAX_PIVOTS; ++j)
      if (j < P.K) above[j] += __shfl_xor(above[j], o, WAVE);
  }
  const int w_ = wave_id();
  if (lane == 0) {
    s_red[w_][3] = seen;
    for (int j = 0; j < MAX_PIVOTS; ++j) s_red[w_][4 + j] = above[j];
    if (mw) atomicMax(max_w, mw);
  }
  __syncthreads();
  if (threadIdx.x <= P.K) {
    // class c counts: total - above[0], above[c-1] - above[c], above[K-1]
    const u32 c = threadIdx.x;
    const u64 lo = (c == 0) ? s_red_total(s_red) : s_red_sum(s_red, 4 + c - 1);
    const u64 hi = (c == P.K) ? 0ull : s_red_sum(s_red, 4 + c);
    const u64 v = lo - hi;
    if (v) atomicAdd(reinterpret_cast<unsigned long long*>(class_counts + c),
                     static_cast<unsigned long long>(v));
  }
}

void weight_classes(hipStream_t st, const u32* w, u64 m, const WeightPivots& P, u32 q0, u32 q1,
                    u64* class_counts, u32* max_w, u32* unit_c0, u32* unit_c1) {
  if (m == 0) return;
  KTScope t(st, K_VALIDATE, 4.0 * m + 8.0 * record_units(m));
  const u64 waves = (record_units(m) + WC_UNITS - 1) / WC_UNITS;
  // 1024 workgroups (4 per CU) keep the loads in flight; more only add
  // block-final reductions and atomics on the few class counters (GNM 2^24 /
  // 2^28: 0.438 ms with 1024, 0.453 with 2048; RGG2D 2^24: 0.231 and 0.239)
  const u32 grid = static_cast<u32>(std::min<u64>((waves + WAVES - 1) / WAVES, 1024));
  if ((reinterpret_cast<uintptr_t>(w) & 15) == 0)
    weight_classes_kernel<true><<<grid, BLOCK, 0, st>>>(w, m, P, q0, q1, class_counts, max_w,
                                                        unit_c0, unit_c1);
  else
    weight_classes_kernel<false><<<grid, BLOCK, 0, st>>>(w, m, P, q0, q1, class_counts, max_w,
                                                         unit_c0, unit_c1);
  KMST_HIP(hipGetLastError());
}

// Packed strict key: (w << 2b) | (min << b) | max with b = bits(n); orders
// exactly like the reference check's (weight, src, dst) on the min<max copy.
__global__ __launch_bounds__(BLOCK) void keys_packed_kernel(const u32* __restrict__ src,
                                                            const u32* __restrict__ dst,
                                                            const u32* __restrict__ w, u64 m,
                                                            u32 vbits, u64* __restrict__ val) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < m; i += u64(gridDim.x) * BLOCK) {
    const u32 s = src[i], d = dst[i];
    const u64 lo = s < d ? s : d, hi = s < d ? d : s;
    val[i] = (u64(w[i]) << (2 * vbits)) | (lo << vbits) | hi;
  }
}

void build_keys_packed(hipStream_t st, const u32* src, const u32* dst, const u32* w, u64 m,
                       u32 vbits, u64* val) {
  if (m == 0) return;
  KTScope t(st, K_KEYS, 20.0 * m);
  keys_packed_kernel<<<grid_for(m), BLOCK, 0, st>>>(src, dst, w, m, vbits, val);
  KMST_HIP(hipGetLastError());
}

// off[v] = first record of vertex v, off[n] = m, for a validated input (sorted
// by src, so the writes total n + 1).  A thread takes 4 consecutive records
// per step (one 16-byte load when aligned) and opens the vertices from the
// previous record's source + 1 to its own.
__global__ __launch_bounds__(BLOCK) void offsets_kernel(const u32* __restrict__ src, u64 m, u64 n,
                                                        u32* __restrict__ off, bool vec) {
  for (u64 q = u64(blockIdx.x) * BLOCK + threadIdx.x; 4 * q < m; q += u64(gridDim.x) * BLOCK) {
    const u64 i0 = 4 * q;
    u32 r[4];
    if (vec && i0 + 3 < m) {
      const uint4 a = *reinterpret_cast<const uint4*>(src + i0);
      r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) r[e] = src[min(i0 + e, m - 1)];
    }
    u64 first = i0 == 0 ? 0 : u64(src[i0 - 1]) + 1;  // first vertex record i0 may open
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const u64 i = i0 + e;
      if (i >= m) break;
      const u64 s = r[e];
      for (u64 v = first; v <= s && v <= n; ++v) off[v] = static_cast<u32>(i);
      if (i == m - 1)
        for (u64 v = s + 1; v <= n; ++v) off[v] = static_cast<u32>(m);
      first = s + 1;
    }
  }
}

void build_offsets(hipStream_t st, const u32* src, u64 m, u64 n, u32* off) {
  if (m == 0) {
    fill_u32(st, off, n + 1, 0);
    return;
  }
  KTScope t(st, K_OFFSETS, 4.0 * m + 4.0 * n);
  const bool vec = (reinterpret_cast<uintptr_t>(src) & 15) == 0;
  offsets_kernel<<<grid_for((m + 3) / 4, 8192), BLOCK, 0, st>>>(src, m, n, off, vec);
  KMST_HIP(hipGetLastError());
}

// Every vertex's record range [seg_start[v], seg_end[v]) of records sorted by
// src: the first record of a run of equal src writes its start and the end of
// the run before it.  One write per run boundary, so an unchecked input costs
// no more (ids >= n are skipped; on unsorted input the ranges are garbage,
// which the readers clamp).
__global__ __launch_bounds__(BLOCK) void record_ranges_kernel(const u32* __restrict__ src, u64 m,
                                                              u64 n, u32* __restrict__ seg_start,
                                                              u32* __restrict__ seg_end, bool vec) {
  for (u64 q = u64(blockIdx.x) * BLOCK + threadIdx.x; 4 * q < m; q += u64(gridDim.x) * BLOCK) {
    const u64 i0 = 4 * q;
    u32 r[4];
    if (vec && i0 + 3 < m) {
      const uint4 a = *reinterpret_cast<const uint4*>(src + i0);
      r[0] = a.x; r[1] = a.y; r[2] = a.z; r[3] = a.w;
    } else {
#pragma unroll
      for (int e = 0; e < 4; ++e) r[e] = src[min(i0 + e, m - 1)];
    }
    u32 ps = i0 == 0 ? 0u : src[i0 - 1];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const u64 i = i0 + e;
      if (i >= m) break;
      const u32 s = r[e];
      if (i == 0 || ps != s) {
        if (s < n) seg_start[s] = static_cast<u32>(i);
        if (i > 0 && ps < n) seg_end[ps] = static_cast<u32>(i);
      }
      if (i + 1 == m && s < n) seg_end[s] = static_cast<u32>(m);
      ps = s;
    }
  }
}

void record_ranges(hipStream_t st, const u32* src, u64 m, u64 n, u32* seg_start, u32* seg_end) {
  if (m == 0) return;
  KTScope t(st, K_OFFSETS, 4.0 * m + 8.0 * n);
  const bool vec = (reinterpret_cast<uintptr_t>(src) & 15) == 0;
  record_ranges_kernel<<<grid_for((m + 3) / 4, 8192), BLOCK, 0, st>>>(src, m, n, seg_start,
                                                                       seg_end, vec);
  KMST_HIP(hipGetLastError());
}

// Position of the first record of the pair (x, y), x < y, in x's record
// range [lo, hi) of an input sorted by (src, dst) (binary search), NO_TWIN
// when the pair has no record there.  It is the positional key's uid: every
// parallel copy of an edge, in both directions, gets the same position, so
// copies of equal weight tie (they are the same (w, min, max) triple, as
// with packed keys) and copies of different weights differ in the weight
// field; the search is O(log degree) however long a run of copies is.
constexpr u64 NO_TWIN = ~u64{0};

__device__ __forceinline__ u64 pair_position(const u32* __restrict__ dst, u32 y, u64 lo, u64 hi) {
  const u64 end = hi;
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    if (dst[mid] < y) lo = mid + 1;
    else hi = mid;
  }
  return (lo < end && dst[lo] == y) ? lo : NO_TWIN;
}

// Positional strict key, used when the packed key does not fit 64 bits:
// val = (w << 32) | uid, uid = pair_position of (min, max) in min's range
// [lo[min], hi[min]): the record's own position for an s < d record that no
// copy of itself precedes, a binary search otherwise.  The input is validated.  (Listing the s > d records of a wave first and
// searching them one per lane measured slower here, 4.2 against 3.2 ms at
// RGG2D 2^24: this pass has no stage to hide the list behind.)
__global__ __launch_bounds__(BLOCK) void keys_kernel(const u32* __restrict__ src,
                                                     const u32* __restrict__ dst,
                                                     const u32* __restrict__ w, u64 m,
                                                     const u32* __restrict__ row_lo,
                                                     const u32* __restrict__ row_hi,
                                                     u64* __restrict__ val, u32* err) {
  u32 bad = 0;
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < m; i += u64(gridDim.x) * BLOCK) {
    const u32 s = src[i], d = dst[i], wi = w[i];
    if (s < d && (i == 0 || src[i - 1] != s || dst[i - 1] != d)) {
      val[i] = (u64(wi) << 32) | i;
      continue;
    }
    const u32 x = min(s, d), y = max(s, d);
    const u64 j = pair_position(dst, y, row_lo[x], row_hi[x]);
    if (j == NO_TWIN) {
      val[i] = VAL_NONE;
      bad |= ERR_ASYMMETRIC;
    } else {
      val[i] = (u64(wi) << 32) | j;
    }
  }
  if (bad) atomicOr(err, bad);
}

void build_keys(hipStream_t st, const u32* src, const u32* dst, const u32* w, u64 m,
                const u32* row_lo, const u32* row_hi, u64* val, u32* err) {
  if (m == 0) return;
  KTScope t(st, K_KEYS, 20.0 * m);
  keys_kernel<<<grid_for(m, 4096), BLOCK, 0, st>>>(src, dst, w, m, row_lo, row_hi, val, err);
  KMST_HIP(hipGetLastError());
}

// ============================================================== K1 argmin
// Segmented argmin over records grouped by src, record-centric so that hub
// vertices cost no more than their records.  A workgroup owns ARG_TILE
// consecutive records, ARG_IPT per thread.  Runs inside one thread are
// finished locally; runs spanning threads are combined by a segmented min
// scan over the threads' tail pieces; runs spanning tiles go through a u64
// atomicMin plus a fix-up pass that writes best_dst.
// With packed keys and the tile's sources inside [first, first + ARG_STAGE)
// (the tile's first source), the finished minima are staged in LDS by
// source and leave in coalesced rows, best_dst unpacked from the key (the
// end that is not the source), instead of one scattered store per run end.
constexpr u32 ARG_STAGE = ARG_TILE;
struct Piece {
  u64 val;
  u32 idx;
  u32 flag;  // segment starts here
};

__device__ __forceinline__ void seg_combine(Piece& cur, const Piece& prev) {
  // cur <- prev (+) cur
  if (!cur.flag) {
    if (prev.val < cur.val) {
      cur.val = prev.val;
      cur.idx = prev.idx;
    }
    cur.flag = prev.flag;
  }
}

template <bool VEC>
__global__ __launch_bounds__(BLOCK) void argmin_kernel(const u32* __restrict__ src,
                                                       const u32* __restrict__ dst,
                                                       const KeySrc keys, u64 m,
                                                       u64* __restrict__ best_val,
                                                       u32* __restrict__ best_dst,
                                                       ArgPiece* __restrict__ pieces,
                                                       u32* piece_count) {
  __shared__ u32 s_tail_src[BLOCK];
  __shared__ u32 s_first_src[BLOCK];
  __shared__ u64 s_scan_val[BLOCK];
  __shared__ u32 s_scan_idx[BLOCK];
  __shared__ u64 s_wave_val[WAVES];
  __shared__ u32 s_wave_idx[WAVES];
  __shared__ u32 s_wave_flag[WAVES];
  __shared__ u64 s_best[ARG_STAGE];

  const int t = threadIdx.x;
  const u64 tile_base = u64(blockIdx.x) * ARG_TILE;
  const u64 tile_end = (tile_base + ARG_TILE < m) ? tile_base + ARG_TILE : m;
  const u64 my_base = tile_base + u64(t) * ARG_IPT;

  // tile context: does the first run continue from the previous tile, does the
  // last run continue into the next one?
  const u32 tile_first_src = src[tile_base];
  const u32 tile_last_src = src[tile_end - 1];
  const bool cross_prev = tile_base > 0 && src[tile_base - 1] == tile_first_src;
  const bool cross_next = tile_end < m && src[tile_end] == tile_last_src;

  bool has_items = my_base < tile_end;
  u32 head_src = NONE32, cur_src = NONE32;
  u64 head_val = VAL_NONE, cur_val = VAL_NONE;
  u32 head_idx = 0, cur_idx = 0, cur_dst = 0;
  bool head_done = false;  // first run ended inside this thread
  // the thread's records, loaded before use (16-byte loads when aligned)
  u32 rs[ARG_IPT], rd[ARG_IPT];
  u64 rv[ARG_IPT];
  if (VEC && my_base + ARG_IPT <= m) {
    const uint4* s4 = reinterpret_cast<const uint4*>(src + my_base);
    const uint4* d4 = reinterpret_cast<const uint4*>(dst + my_base);
#pragma unroll
    for (int q = 0; q < ARG_IPT / 4; ++q) {
      const uint4 a = s4[q], b = d4[q];
      rs[4 * q] = a.x; rs[4 * q + 1] = a.y; rs[4 * q + 2] = a.z; rs[4 * q + 3] = a.w;
      rd[4 * q] = b.x; rd[4 * q + 1] = b.y; rd[4 * q + 2] = b.z; rd[4 * q + 3] = b.w;
    }
    if (keys.v) {
      const uint4* v4 = reinterpret_cast<const uint4*>(keys.v + my_base);
#pragma unroll
      for (int q = 0; q < ARG_IPT / 2; ++q) {
        const uint4 a = v4[q];
        rv[2 * q] = (u64(a.y) << 32) | a.x;
        rv[2 * q + 1] = (u64(a.w) << 32) | a.z;
      }
    } else {
      const uint4* w4 = reinterpret_cast<const uint4*>(keys.w + my_base);
#pragma unroll
      for (int q = 0; q < ARG_IPT / 4; ++q) {
        const uint4 a = w4[q];
        rv[4 * q] = pack_key(a.x, rs[4 * q], rd[4 * q], keys.vbits);
        rv[4 * q + 1] = pack_key(a.y, rs[4 * q + 1], rd[4 * q + 1], keys.vbits);
        rv[4 * q + 2] = pack_key(a.z, rs[4 * q + 2], rd[4 * q + 2], keys.vbits);
        rv[4 * q + 3] = pack_key(a.w, rs[4 * q + 3], rd[4 * q + 3], keys.vbits);
      }
    }
  } else {
#pragma unroll
    for (int k = 0; k < ARG_IPT; ++k) {
      const u64 i = min(my_base + k, m - 1);
      rs[k] = src[i];
      rd[k] = dst[i];
      rv[k] = load_key(keys, i, rs[k], rd[k]);
    }
  }
  // staged: every record of the tile has its source within ARG_STAGE of the
  // tile's first (the ones past m - 1 repeat the last record)
  bool near = true;
#pragma unroll
  for (int k = 0; k < ARG_IPT; ++k) near = near && rs[k] - tile_first_src < ARG_STAGE;
  for (u32 i = t; i < ARG_STAGE; i += BLOCK) s_best[i] = VAL_NONE;
  const bool staged = keys.v == nullptr && __syncthreads_and(near);
  if (has_items) {
    const u32 cnt = static_cast<u32>(min<u64>(ARG_IPT, tile_end - my_base));
    cur_src = rs[0];
    cur_val = rv[0];
    cur_idx = static_cast<u32>(my_base);
    cur_dst = rd[0];
#pragma unroll
    for (u32 k = 1; k < ARG_IPT; ++k) {
      if (k >= cnt) break;
      const u32 s = rs[k];
      const u64 v = rv[k];
      if (s != cur_src) {
        if (!head_done) {
          head_src = cur_src;
          head_val = cur_val;
          head_idx = cur_idx;
          head_done = true;
        } else if (staged) {  // interior run: complete inside this thread
          s_best[cur_src - tile_first_src] = cur_val;
        } else {
          best_val[cur_src] = cur_val;
          best_dst[cur_src] = cur_dst;
        }
        cur_src = s;
        cur_val = v;
        cur_idx = static_cast<u32>(my_base + k);
        cur_dst = rd[k];
      } else if (v < cur_val) {
        cur_val = v;
        cur_idx = static_cast<u32>(my_base + k);
        cur_dst = rd[k];
      }
    }
  }
  const u32 first_src = head_done ? head_src : cur_src;
  s_tail_src[t] = cur_src;
  s_first_src[t] = first_src;
  __syncthreads();

  // segmented inclusive min-scan over tail pieces
  const u32 prev_tail = (t > 0) ? s_tail_src[t - 1] : NONE32;
  Piece p;
  p.val = cur_val;
  p.idx = cur_idx;
  p.flag = (!has_items || head_done || t == 0 || prev_tail != cur_src) ? 1u : 0u;
  const int lane = lane_id(), wid = wave_id();
#pragma unroll
  for (int dd = 1; dd < WAVE; dd <<= 1) {
    Piece q;
    q.val = __shfl_up(p.val, dd, WAVE);
    q.idx = __shfl_up(p.idx, dd, WAVE);
    q.flag = __shfl_up(p.flag, dd, WAVE);
    if (lane >= dd) seg_combine(p, q);
  }
  if (lane == WAVE - 1) {
    s_wave_val[wid] = p.val;
    s_wave_idx[wid] = p.idx;
    s_wave_flag[wid] = p.flag;
  }
  __syncthreads();
  if (wid > 0 && !p.flag) {
    Piece carry{s_wave_val[0], s_wave_idx[0], s_wave_flag[0]};
    for (int w = 1; w < wid; ++w) {
      Piece nx{s_wave_val[w], s_wave_idx[w], s_wave_flag[w]};
      seg_combine(nx, carry);
      carry = nx;
    }
    seg_combine(p, carry);
  }
  s_scan_val[t] = p.val;
  s_scan_idx[t] = p.idx;
  __syncthreads();

  if (!has_items && !staged) return;

  // emit one finished run: staged or direct store, or atomic + fix-up when
  // it crosses a tile boundary
  auto emit = [&](u32 s, u64 v, u32 idx) {
    const bool crosses = (cross_prev && s == tile_first_src) || (cross_next && s == tile_last_src);
    if (!crosses && staged) {
      s_best[s - tile_first_src] = v;
    } else if (!crosses) {
      best_val[s] = v;
      best_dst[s] = dst[idx];
    } else {
      atomicMin(reinterpret_cast<unsigned long long*>(best_val + s),
                static_cast<unsigned long long>(v));
      const u32 slot = atomicAdd(piece_count, 1u);
      pieces[slot] = ArgPiece{v, s, idx};
    }
  };

  // head run (thread has more than one run)
  if (has_items && head_done) {
    if (t > 0 && prev_tail == head_src) {
      // run started in an earlier thread and ends here
      const u64 pv = s_scan_val[t - 1];
      if (pv < head_val) emit(head_src, pv, s_scan_idx[t - 1]);
      else emit(head_src, head_val, head_idx);
    } else {
      emit(head_src, head_val, head_idx);
    }
  }
  // tail run: this thread finishes it when the next piece has another src
  const bool last_thread = (t == BLOCK - 1) || (my_base + ARG_IPT >= tile_end);
  const u32 next_first = last_thread ? NONE32 : s_first_src[t + 1];
  if (has_items && (last_thread || next_first != cur_src)) emit(cur_src, p.val, p.idx);
  if (!staged) return;
  __syncthreads();
  const u64 id_mask = (u64{1} << keys.vbits) - 1;
  for (u32 i = t; i < ARG_STAGE; i += BLOCK) {
    const u64 v = s_best[i];
    if (v == VAL_NONE) continue;
    const u32 s = tile_first_src + i;
    const u32 lo = static_cast<u32>((v >> keys.vbits) & id_mask);
    const u32 hi = static_cast<u32>(v & id_mask);
    best_val[s] = v;
    best_dst[s] = lo == s ? hi : lo;
  }
}

__global__ __launch_bounds__(BLOCK) void argmin_fixup_kernel(const ArgPiece* __restrict__ pieces,
                                                             const u32* piece_count,
                                                             const u32* __restrict__ dst,
                                                             const u64* __restrict__ best_val,
                                                             u32* __restrict__ best_dst) {
  const u32 cnt = *piece_count;
  for (u32 i = blockIdx.x * BLOCK + threadIdx.x; i < cnt; i += gridDim.x * BLOCK) {
    const ArgPiece pc = pieces[i];
    if (best_val[pc.src] == pc.val) best_dst[pc.src] = dst[pc.idx];
  }
}

__global__ __launch_bounds__(BLOCK) void argmin_init_kernel(u64* __restrict__ best_val, u64 n,
                                                           u32* __restrict__ piece_count) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *piece_count = 0;
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += u64(gridDim.x) * BLOCK)
    best_val[i] = VAL_NONE;
}

void segmented_argmin(hipStream_t st, const u32* src, const u32* dst, const KeySrc& keys, u64 m,
                      u64 n, u64* best_val, u32* best_dst, ArgPiece* pieces, u32* piece_count) {
  {
    KTScope t(st, K_FILL, 8.0 * n);
    argmin_init_kernel<<<grid_for(n), BLOCK, 0, st>>>(best_val, n, piece_count);
    KMST_HIP(hipGetLastError());
  }
  if (m == 0) return;
  const u32 tiles = ceil_div_u32(m, ARG_TILE);
  {
    KTScope t(st, K_ARGMIN, 12.0 * m + 16.0 * n);
    const bool vec = ((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst) |
                       reinterpret_cast<uintptr_t>(keys.v ? static_cast<const void*>(keys.v)
                                                          : static_cast<const void*>(keys.w))) &
                      15) == 0;
    if (vec)
      argmin_kernel<true><<<tiles, BLOCK, 0, st>>>(src, dst, keys, m, best_val, best_dst, pieces,
                                                   piece_count);
    else
      argmin_kernel<false><<<tiles, BLOCK, 0, st>>>(src, dst, keys, m, best_val, best_dst, pieces,
                                                    piece_count);
    KMST_HIP(hipGetLastError());
  }
  KTScope t(st, K_ARGMIN_FIXUP, 32.0 * tiles);
  argmin_fixup_kernel<<<grid_for(2ull * tiles, 1024), BLOCK, 0, st>>>(pieces, piece_count, dst,
                                                                      best_val, best_dst);
  KMST_HIP(hipGetLastError());
}

// ============================================================== K2 hook
// Each vertex hooks onto the other endpoint of its lightest edge; of a mutual
// pair the lower endpoint stays root (gbbs_reimplementation.hpp:117-127).
// state[v] = (root with an edge ? 1 : 0) | (hooked, i.e. emits its edge to
// the MST ? 1 << 32 : 0); one 64-bit exclusive scan of state then gives both
// the consecutive new id of every root and the MST slot of every emitter,
// without atomics.
constexpr u64 ST_ROOT = 1ull;
constexpr u64 ST_EMIT = 1ull << 32;

// K1 without grouping: records in any order.  The offer pass folds the keys
// of adjacent lanes with equal sources (order-keeping compactions keep the
// records of one original vertex adjacent, so long runs survive the rounds)
// and takes one 64-bit atomic minimum per run; the bind pass names the other
// endpoint of every minimum.
constexpr int UA_ROWS = 4;  // rows of 64 records a wave loads before it offers
__global__ __launch_bounds__(BLOCK) void argmin_offer_kernel(const u32* __restrict__ src,
                                                             const u32* __restrict__ dst,
                                                             const KeySrc keys, u64 m,
                                                             u64* __restrict__ best) {
  constexpr u64 CH = u64(WAVE) * UA_ROWS;
  const u64 gw = (u64(blockIdx.x) * BLOCK + threadIdx.x) / WAVE;
  const u64 nw = u64(gridDim.x) * WAVES;
  const int lane = lane_id();
  for (u64 base = gw * CH; base < m; base += nw * CH) {
    u32 ss[UA_ROWS];
    u64 kk[UA_ROWS];
#pragma unroll
    for (int u = 0; u < UA_ROWS; ++u) {
      const u64 i = base + u64(u) * WAVE + lane;
      ss[u] = i < m ? src[i] : NONE32;
    }
#pragma unroll
    for (int u = 0; u < UA_ROWS; ++u) {
      const u64 i = base + u64(u) * WAVE + lane;
      kk[u] = i < m ? load_key(keys, i, ss[u], keys.v ? 0u : dst[i]) : VAL_NONE;
    }
#pragma unroll
    for (int u = 0; u < UA_ROWS; ++u) offer_min(best, ss[u], kk[u]);
  }
}

__global__ __launch_bounds__(BLOCK) void argmin_bind_kernel(const u32* __restrict__ src,
                                                            const u32* __restrict__ dst,
                                                            const KeySrc keys, u64 m,
                                                            const u64* __restrict__ best,
                                                            u32* __restrict__ best_dst) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < m; i += u64(gridDim.x) * BLOCK) {
    const u32 s = src[i], d = dst[i];
    const u64 k = load_key(keys, i, s, d);
    if (k != VAL_NONE && k == best[s]) best_dst[s] = d;
  }
}

void unordered_argmin(hipStream_t st, const u32* src, const u32* dst, const KeySrc& keys, u64 m,
                      u64 n, u64* best_val, u32* best_dst) {
  fill_u64(st, best_val, n, VAL_NONE);
  if (m == 0) return;
  {
    KTScope t(st, K_ARGMIN, 12.0 * double(m));
    argmin_offer_kernel<<<grid_for((m + WAVE * UA_ROWS - 1) / (WAVE * UA_ROWS) * WAVE, 16384),
                          BLOCK, 0, st>>>(src, dst, keys, m, best_val);
    KMST_HIP(hipGetLastError());
  }
  {
    KTScope t(st, K_ARGMIN, 24.0 * double(m) + 4.0 * double(n));
    argmin_bind_kernel<<<grid_for(m, 16384), BLOCK, 0, st>>>(src, dst, keys, m, best_val,
                                                              best_dst);
    KMST_HIP(hipGetLastError());
  }
}

__global__ __launch_bounds__(BLOCK) void hook_kernel(u64 n, const u64* __restrict__ best_val,
                                                     const u32* __restrict__ best_dst,
                                                     u32* __restrict__ parent,
                                                     u64* __restrict__ state) {
  const u64 v = u64(blockIdx.x) * BLOCK + threadIdx.x;
  if (v >= n) return;
  const u64 bv = best_val[v];
  if (bv == VAL_NONE) {
    parent[v] = static_cast<u32>(v);
    state[v] = 0;
    return;
  }
  const u32 u = best_dst[v];
  // only the lower end of a mutual pair becomes a root: a vertex above its
  // target hooks either way and skips the gather
  bool root = false;
  if (v < u) root = best_val[u] == bv;
  if (root) {
    parent[v] = static_cast<u32>(v);
    state[v] = ST_ROOT;
  } else {
    parent[v] = u;
    state[v] = ST_EMIT;
  }
}

void hook(hipStream_t st, u64 n, const u64* best_val, const u32* best_dst, u32* parent,
          u64* state) {
  if (n == 0) return;
  KTScope t(st, K_HOOK, 32.0 * n);
  hook_kernel<<<ceil_div_u32(n, BLOCK), BLOCK, 0, st>>>(n, best_val, best_dst, parent, state);
  KMST_HIP(hipGetLastError());
}

// ============================================================== K3 jump
// Label doubling to the fixpoint in one launch: every vertex keeps replacing
// its parent by its grandparent until it points at a root; the shortcuts it
// writes are ancestors, so concurrent chases only speed each other up.  The
// hook graph is a forest (the only 2-cycles are broken by the root rule), so
// every chase ends; a loop bound turns a corrupted forest into an error.
__global__ __launch_bounds__(BLOCK) void jump_kernel(u64 n, u32* parent, u32* err) {
  const u64 v = u64(blockIdx.x) * BLOCK + threadIdx.x;
  if (v >= n) return;
  u32 x = parent[v];
  u32 y = parent[x];
  u32 guard = 0;
  // every step stores its progress, which shortens the walks of vertices
  // that pass through v; a vertex already pointing at its root stores nothing
  while (x != y) {
    parent[v] = y;
    x = y;
    y = parent[x];
    if (++guard > (1u << 26)) {
      atomicOr(err, ERR_SPIN);
      return;
    }
  }
}

// jump_kernel that also counts the roots with records (parent[v] == v and
// seg_end[v] != 0) into *roots: grid-strided over at most 2048 workgroups,
// one atomic each (the block-final atomics of a one-shot grid of n / 256
// workgroups on one counter would serialise).
__global__ __launch_bounds__(BLOCK) void jump_count_kernel(u64 n, u32* parent,
                                                           const u32* __restrict__ seg_end,
                                                           u32* roots, u32* err) {
  __shared__ u32 s_c[WAVES];
  u32 c = 0;
  bool spun = false;
  for (u64 v = u64(blockIdx.x) * BLOCK + threadIdx.x; v < n && !spun;
       v += u64(gridDim.x) * BLOCK) {
    u32 x = parent[v];
    if (x == v) {  // a root stays one: only non-roots store
      c += seg_end[v] ? 1u : 0u;
      continue;
    }
    u32 y = parent[x];
    u32 guard = 0;
    while (x != y) {
      parent[v] = y;
      x = y;
      y = parent[x];
      if (++guard > (1u << 26)) {
        atomicOr(err, ERR_SPIN);
        spun = true;
        break;
      }
    }
  }
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, WAVE);
  if (lane_id() == 0) s_c[wave_id()] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 t = 0;
    for (int w = 0; w < WAVES; ++w) t += s_c[w];
    if (t) atomicAdd(roots, t);
  }
}

void pointer_jump(hipStream_t st, u64 n, u32* parent, u32* err, const u32* seg_end, u32* roots) {
  if (roots) fill_u32(st, roots, 1, 0);
  if (n == 0) return;
  KTScope t(st, K_JUMP, 12.0 * n + (roots ? 4.0 * n : 0.0));
  if (roots)
    jump_count_kernel<<<grid_for(n, 2048), BLOCK, 0, st>>>(n, parent, seg_end, roots, err);
  else
    jump_kernel<<<ceil_div_u32(n, BLOCK), BLOCK, 0, st>>>(n, parent, err);
  KMST_HIP(hipGetLastError());
}

// Pointer jumping with the round's bookkeeping folded in (the state scan only
// needs the hook, so it runs first): once v has reached its root r,
// label[v] = r's new id (NONE32 for finished vertices), an emitter writes its
// edge key into the MST list at mst_base + its slot, and with gpar the
// parent array over the original ids is updated as gpar_kernel does.
// TAGGED: the scan left every surviving root's new id in its parent entry
// with ROOT_TAG set, so the load that reaches the root also yields the id;
// roots keep that entry, finished vertices point at themselves untagged.
// Otherwise root_id[r] holds the ids.
template <bool TAGGED>
__global__ __launch_bounds__(BLOCK) void jump_label_kernel(
    u64 n, u32* parent, const u32* __restrict__ root_id, const u64* __restrict__ state,
    const u64* __restrict__ pos, const u64* __restrict__ best_val, u32* __restrict__ label,
    u64* __restrict__ mst, u64 mst_base, const u32* __restrict__ rep, u32* __restrict__ gpar,
    u32* __restrict__ rep_next, u32* err) {
  const u64 v = u64(blockIdx.x) * BLOCK + threadIdx.x;
  if (v >= n) return;
  // the vertex's own words are loaded before the walk, so that their
  // latency overlaps the walk's dependent loads
  const u64 sv = state[v];
  const u64 pv = pos[v];
  const u64 bv = best_val[v];
  const u32 name = gpar ? (rep ? rep[v] : static_cast<u32>(v)) : 0u;
  u32 x = parent[v];
  u32 id;
  if (TAGGED && (x & ROOT_TAG)) {
    id = x & ~ROOT_TAG;
    x = static_cast<u32>(v);
  } else {
    u32 y = parent[x];
    u32 guard = 0;
    while (x != y && !(TAGGED && (y & ROOT_TAG))) {
      parent[v] = y;
      x = y;
      y = parent[x];
      if (++guard > (1u << 26)) {
        atomicOr(err, ERR_SPIN);
        return;
      }
    }
    parent[v] = x;
    if (TAGGED) id = (y & ROOT_TAG) ? (y & ~ROOT_TAG) : NONE32;
    else id = root_id[x];
  }
  label[v] = id;
  if (sv & ST_EMIT) mst[mst_base + (pv >> 32)] = bv;
  if (gpar && sv != 0) {
    gpar[name] = rep ? rep[x] : x;
    if (sv & ST_ROOT) rep_next[static_cast<u32>(pv)] = name;
  }
}

void jump_and_label(hipStream_t st, u64 n, u32* parent, const u32* root_id, const u64* state,
                    const u64* pos, const u64* best_val, u32* label, u64* mst, u64 mst_base,
                    const u32* rep, u32* gpar, u32* rep_next, u32* err) {
  if (n == 0) return;
  KTScope t(st, K_JUMP, (root_id ? 40.0 : 36.0) * n + (gpar ? 16.0 * n : 0.0));
  if (root_id)
    jump_label_kernel<false><<<ceil_div_u32(n, BLOCK), BLOCK, 0, st>>>(
        n, parent, root_id, state, pos, best_val, label, mst, mst_base, rep, gpar, rep_next, err);
  else
    jump_label_kernel<true><<<ceil_div_u32(n, BLOCK), BLOCK, 0, st>>>(
        n, parent, nullptr, state, pos, best_val, label, mst, mst_base, rep, gpar, rep_next, err);
  KMST_HIP(hipGetLastError());
}

// label[v] = new consecutive id of v's root (NONE32 for finished vertices);
// emitters write their edge key into the MST list at mst_base + their slot.
__global__ __launch_bounds__(BLOCK) void label_kernel(u64 n, const u32* __restrict__ parent,
                                                      const u32* __restrict__ root_id,
                                                      const u64* __restrict__ state,
                                                      const u64* __restrict__ pos,
                                                      const u64* __restrict__ best_val,
                                                      u32* __restrict__ label,
                                                      u64* __restrict__ mst, u64 mst_base) {
  const u64 v = u64(blockIdx.x) * BLOCK + threadIdx.x;
  if (v >= n) return;
  label[v] = root_id[parent[v]];
  if (state[v] & ST_EMIT) mst[mst_base + (pos[v] >> 32)] = best_val[v];
}

void make_labels(hipStream_t st, u64 n, const u32* parent, const u32* root_id, const u64* state,
                 const u64* pos, const u64* best_val, u32* label, u64* mst, u64 mst_base) {
  if (n == 0) return;
  KTScope t(st, K_LABEL, 28.0 * n);
  label_kernel<<<ceil_div_u32(n, BLOCK), BLOCK, 0, st>>>(n, parent, root_id, state, pos,
                                                          best_val, label, mst, mst_base);
  KMST_HIP(hipGetLastError());
}

// ============================================================== K4+K5
// Relabel both endpoints, drop self-loops and dead records, keep order; one
// pass: each wave owns a contiguous sub-tile (coalesced rows of 64 records),
// ballots give the in-wave order, a decoupled look-back gives the tile offset.
// Dynamic tile id in dispatch order from a counter that only grows.
__device__ __forceinline__ u32 lb_tile_id(unsigned long long* counter, unsigned long long base) {
  __shared__ u32 s_tile;
  if (threadIdx.x == 0) s_tile = static_cast<u32>(atomicAdd(counter, 1ull) - base);
  __syncthreads();
  return s_tile;
}

// Loads of one wave's sub-tile: record e = wbase + k * WAVE + lane for k <
// CMP_IPT.  The index is clamped to m - 1, so every load is unconditional and
// all of them can be in flight together; callers mask the e >= m slots.
__device__ __forceinline__ void load_records_sd(const u32* __restrict__ s_in,
                                                const u32* __restrict__ d_in, u64 wbase, u64 m,
                                                u32 (&ss)[CMP_IPT], u32 (&dd)[CMP_IPT]) {
  const u64 lane = lane_id();
#pragma unroll
  for (int k = 0; k < CMP_IPT; ++k) {
    const u64 e = min(wbase + u64(k) * WAVE + lane, m - 1);
    ss[k] = s_in[e];
    dd[k] = d_in[e];
  }
}

__device__ __forceinline__ void load_records(const u32* __restrict__ s_in,
                                             const u32* __restrict__ d_in, const KeySrc& keys,
                                             u64 wbase, u64 m, u32 (&ss)[CMP_IPT],
                                             u32 (&dd)[CMP_IPT], u64 (&vv)[CMP_IPT]) {
  const u64 lane = lane_id();
  load_records_sd(s_in, d_in, wbase, m, ss, dd);
  if (keys.v) {
#pragma unroll
    for (int k = 0; k < CMP_IPT; ++k) vv[k] = keys.v[min(wbase + u64(k) * WAVE + lane, m - 1)];
  } else {
    u32 ww[CMP_IPT];
#pragma unroll
    for (int k = 0; k < CMP_IPT; ++k) ww[k] = keys.w[min(wbase + u64(k) * WAVE + lane, m - 1)];
#pragma unroll
    for (int k = 0; k < CMP_IPT; ++k) vv[k] = pack_key(ww[k], ss[k], dd[k], keys.vbits);
  }
}

// With TWINS the records are symmetric (every (a, b, key) has its (b, a, key)
// twin, or a lighter parallel record that makes both redundant): only the
// forward records (src < dst) are read past their ids and gathered, and each
// kept one is written in both directions, the wave's forward copies first,
// then its reverse copies in the same order.  Half the random label gathers.
template <int THREADS, bool TWINS>
__global__ __launch_bounds__(THREADS) void relabel_compact_kernel(
    const u32* __restrict__ s_in, const u32* __restrict__ d_in, const KeySrc keys, u64 key_min,
    u64 key_max, u64 m, const u32* __restrict__ label, const u64* __restrict__ giant,
    u32* __restrict__ s_out,
    u32* __restrict__ d_out, u64* __restrict__ v_out, u64* lb_status, u32 lb_epoch,
    unsigned long long* lb_counter, unsigned long long lb_base, u32 tiles, u32* count_dev,
    u32* err) {
  constexpr int NW = THREADS / WAVE;
  __shared__ u32 s_wave_cnt[NW];
  __shared__ u64 s_excl;
  __shared__ u32 s_ok;
  const u32 tile = lb_tile_id(lb_counter, lb_base);
  const int lane = lane_id(), w = wave_id();
  const u64 wbase = u64(tile) * (THREADS * CMP_IPT) + u64(w) * (WAVE * CMP_IPT);
  u32 ls[CMP_IPT], ld[CMP_IPT];
  u64 vv[CMP_IPT];
  u64 bal[CMP_IPT];
  u32 cnt = 0;
  // phased so that each thread has all its loads of one kind in flight at
  // once: (src, dst), then the bitmap, then keys, then the label gathers
  u32 se[CMP_IPT], de[CMP_IPT];
  load_records_sd(s_in, d_in, wbase, m, se, de);
  bool live[CMP_IPT];
#pragma unroll
  for (int k = 0; k < CMP_IPT; ++k) {
    const u64 e = wbase + u64(k) * WAVE + lane;
    const bool fwd = !TWINS || se[k] < de[k];
    // both ends in the marked component: a self-loop, decided without the
    // key or the label gathers
    const bool inside = fwd && giant && ((giant[se[k] >> 6] >> (se[k] & 63)) &
                                         (giant[de[k] >> 6] >> (de[k] & 63)) & 1ull);
    live[k] = e < m && fwd && !inside;
  }
#pragma unroll
  for (int k = 0; k < CMP_IPT; ++k) {
    const u64 e = wbase + u64(k) * WAVE + lane;
    vv[k] = live[k] ? load_key(keys, e, se[k], de[k]) : VAL_NONE;
    live[k] = live[k] && vv[k] != VAL_NONE && vv[k] >= key_min && vv[k] <= key_max;
  }
#pragma unroll
  for (int k = 0; k < CMP_IPT; ++k) {
    ls[k] = live[k] ? label[se[k]] : 0u;
    ld[k] = live[k] ? label[de[k]] : 0u;
  }
#pragma unroll
  for (int k = 0; k < CMP_IPT; ++k) {
    bal[k] = __ballot(live[k] && ls[k] != ld[k]);
    cnt += static_cast<u32>(__popcll(bal[k]));
  }
  if (lane == 0) s_wave_cnt[w] = TWINS ? 2 * cnt : cnt;
  __syncthreads();
  if (w == 0) {
    u64 agg = 0;
#pragma unroll
    for (int ww = 0; ww < NW; ++ww) agg += s_wave_cnt[ww];
    bool ok;
    const u64 excl = lookback_exclusive(lb_status, lb_epoch, tile, agg, err, ok);
    if (lane == 0) {
      s_excl = excl;
      s_ok = ok ? 1u : 0u;
      if (tile == tiles - 1) *count_dev = static_cast<u32>(excl + agg);
    }
  }
  __syncthreads();
  if (!s_ok) return;
  u64 pos = s_excl;
  for (int ww = 0; ww < w; ++ww) pos += s_wave_cnt[ww];
#pragma unroll
  for (int k = 0; k < CMP_IPT; ++k) {
    if ((bal[k] >> lane) & 1ull) {
      const u64 p = pos + mbcnt(bal[k]);
      s_out[p] = ls[k];
      d_out[p] = ld[k];
      v_out[p] = vv[k];
      if (TWINS) {
        s_out[p + cnt] = ld[k];
        d_out[p + cnt] = ls[k];
        v_out[p + cnt] = vv[k];
      }
    }
    pos += static_cast<u32>(__popcll(bal[k]));
  }
}

void LookbackWork::bind(hipStream_t st, u64* p, u64 m) {
  const u64 tiles = ceil_div_u32(m ? m : 1, CMP_TILE);
  if (p != words || tiles > tiles_cap) {
    // fresh memory: zero the status words and the tile counter once
    words = p;
    tiles_cap = tiles;
    fill_u64(st, p, tiles + 1, 0);
    epoch = 0;
    tiles_issued = 0;
  }
}

u32 LookbackWork::next_call(hipStream_t st, u32 tiles) {
  if (tiles > tiles_cap) throw Error(KMST_EINVAL, "look-back workspace too small");
  if (++epoch == (1u << LB_EPOCH_BITS)) {
    // epochs wrapped: clear the status words (the counter keeps growing)
    fill_u64(st, words, tiles_cap, 0);
    epoch = 1;
  }
  return epoch;
}

void relabel_compact(hipStream_t st, const u32* s_in, const u32* d_in, const KeySrc& keys,
                     u64 key_min, u64 m, const u32* label, u32* s_out, u32* d_out, u64* v_out,
                     LookbackWork& lb, u32* count_dev, u32* err, const u64* giant, u64 n,
                     u64 key_max, bool twins) {
  if (keys.twin_lo) throw Error(KMST_EINVAL, "relabel needs stored or packed keys");
  if (m == 0) {
    fill_u32(st, count_dev, 1, 0);
    return;
  }
  // the look-back passes a tile's offset on in a chain whose cost per tile
  // bounds large compactions: from RELABEL_WIDE_MIN records on a tile has 512
  // threads (8192 records), half as many tiles
  const bool wide = m >= RELABEL_WIDE_MIN;
  const u32 tiles = ceil_div_u32(m, u64(wide ? 512 : BLOCK) * CMP_IPT);
  const u32 epoch = lb.next_call(st, tiles);
  // reads: (src, dst) 8 B, the key (4 B weight or 8 B stored) and two 4-byte
  // label gathers -- with twins the key and the gathers of the forward half
  // only; with the component bitmap only (src, dst) and the bitmap are
  // counted (a lower bound).  The 16-byte writes of kept records are added by
  // the caller.
  const double per_read = (keys.v ? 8.0 : 4.0) + 8.0;
  KTScope t(st, K_RELABEL_SCATTER,
            giant ? 8.0 * m + double(n) / 8.0 : 8.0 * m + per_read * (twins ? 0.5 : 1.0) * m);
#define KMST_RELABEL(T, TW)                                                                    \
  relabel_compact_kernel<T, TW><<<tiles, T, 0, st>>>(                                          \
      s_in, d_in, keys, key_min, key_max, m, label, giant, s_out, d_out, v_out, lb.words, epoch, \
      reinterpret_cast<unsigned long long*>(lb.words + lb.tiles_cap), lb.tiles_issued, tiles,  \
      count_dev, err)
  if (wide && twins) KMST_RELABEL(512, true);
  else if (wide) KMST_RELABEL(512, false);
  else if (twins) KMST_RELABEL(BLOCK, true);
  else KMST_RELABEL(BLOCK, false);
#undef KMST_RELABEL
  KMST_HIP(hipGetLastError());
  lb.tiles_issued += tiles;
}

// ============================================================== K6 sort
// Stable LSD radix sort with 8-bit digits and one kernel per pass
// ("onesweep"): one read of the keys builds the digit histograms of every
// pass, their exclusive scans give each digit's global base, and inside a
// scatter pass each tile learns how many records of each digit the tiles
// before it hold through a decoupled look-back per digit.  Each wave
// owns a contiguous sub-tile and ranks its items in index order with 8-ballot
// digit matching, so a pass is stable; the tile is staged in LDS in (digit,
// input order) and leaves in runs of consecutive addresses: (src, dst) first,
// then the keys through the same LDS buffer.  No host synchronisation.
static_assert(RADIX == BLOCK, "one thread per digit");

// Sort key: (s << dbits) | d, or s alone when dbits is 0 (d, if present, is
// then only carried along).
template <bool HAS_D>
__device__ __forceinline__ u64 sort_key_of(u32 s, u32 d, u32 dbits) {
  return (HAS_D && dbits) ? ((u64(s) << dbits) | d) : u64(s);
}

// Digit of pass p: width[p] key bits from bit shift[p] (widths of at most
// RADIX_BITS).  With balanced digits a key of b bits takes ceil(b / 8) passes
// of nearly equal widths (20 bits: 7 + 7 + 6), so narrower keys get fewer
// buckets per pass and longer runs per bucket in every scatter tile.
DigitPlan plan_digits(u32 total_bits, bool balanced) {
  DigitPlan pl{};
  pl.passes = (total_bits + RADIX_BITS - 1) / RADIX_BITS;
  u32 shift = 0;
  for (u32 p = 0; p < pl.passes; ++p) {
    u32 w = RADIX_BITS;
    if (balanced) w = total_bits / pl.passes + (p < total_bits % pl.passes ? 1u : 0u);
    else if (shift + w > total_bits) w = total_bits - shift;
    pl.shift[p] = static_cast<u8>(shift);
    pl.width[p] = static_cast<u8>(w);
    shift += w;
  }
  return pl;
}

__device__ __forceinline__ u32 plan_digit(u64 key, const DigitPlan& pl, u32 p) {
  return static_cast<u32>((key >> pl.shift[p]) & ((1u << pl.width[p]) - 1));
}

// ghist[p * RADIX + x] = number of records whose digit p is x.  Per-wave
// LDS histograms.  Each thread counts UP_RPT consecutive keys and adds runs
// of equal digits at once: records grouped by src arrive in runs (a vertex's
// records, relabelled alike), so most LDS atomics fold away.
constexpr int UP_RPT = 8;

// With vec (s and d 16-byte aligned) a thread's UP_RPT keys come in two
// 16-byte loads per array.
template <bool HAS_D>
__global__ __launch_bounds__(BLOCK) void radix_upsweep_kernel(const u32* __restrict__ s,
                                                              const u32* __restrict__ d, u64 m,
                                                              u32 dbits, const DigitPlan pl,
                                                              u32* __restrict__ ghist, bool vec) {
  static_assert(UP_RPT == 8, "two 16-byte loads per array");
  const u32 passes = pl.passes;
  __shared__ u32 h[WAVES][MAX_PASSES][RADIX];
  for (int i = threadIdx.x; i < WAVES * MAX_PASSES * RADIX; i += BLOCK) (&h[0][0][0])[i] = 0;
  __syncthreads();
  const int w = wave_id();
  const u64 stride = u64(gridDim.x) * (BLOCK * UP_RPT);
  for (u64 base = u64(blockIdx.x) * (BLOCK * UP_RPT) + u64(threadIdx.x) * UP_RPT; base < m;
       base += stride) {
    u64 key[UP_RPT];
    if (vec && base + UP_RPT <= m) {
      const uint4 s0 = *reinterpret_cast<const uint4*>(s + base);
      const uint4 s1 = *reinterpret_cast<const uint4*>(s + base + 4);
      const u32 sv[UP_RPT] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
      u32 dv[UP_RPT] = {};
      if (HAS_D) {
        const uint4 d0 = *reinterpret_cast<const uint4*>(d + base);
        const uint4 d1 = *reinterpret_cast<const uint4*>(d + base + 4);
        dv[0] = d0.x; dv[1] = d0.y; dv[2] = d0.z; dv[3] = d0.w;
        dv[4] = d1.x; dv[5] = d1.y; dv[6] = d1.z; dv[7] = d1.w;
      }
#pragma unroll
      for (int k = 0; k < UP_RPT; ++k) key[k] = sort_key_of<HAS_D>(sv[k], dv[k], dbits);
    } else {
#pragma unroll
      for (int k = 0; k < UP_RPT; ++k) {
        const u64 e = min(base + k, m - 1);
        key[k] = sort_key_of<HAS_D>(s[e], HAS_D ? d[e] : 0u, dbits);
      }
    }
    const u32 valid = static_cast<u32>(min<u64>(UP_RPT, m - base));
    for (u32 p = 0; p < passes; ++p) {
      u32 cur = plan_digit(key[0], pl, p), run = 1;
#pragma unroll
      for (int k = 1; k < UP_RPT; ++k) {
        if (u32(k) >= valid) break;
        const u32 dg = plan_digit(key[k], pl, p);
        if (dg == cur) {
          ++run;
        } else {
          atomicAdd(&h[w][p][cur], run);
          cur = dg;
          run = 1;
        }
      }
      atomicAdd(&h[w][p][cur], run);
    }
  }
  __syncthreads();
  for (u32 i = threadIdx.x; i < passes * RADIX; i += BLOCK) {
    u32 c = 0;
#pragma unroll
    for (int ww = 0; ww < WAVES; ++ww) c += (&h[ww][0][0])[i];
    if (c) atomicAdd(ghist + i, c);
  }
}

// base[p * RADIX + x] = number of records whose digit p is below x; the
// histograms are cleared for the next sort (they start zeroed at bind).
__global__ __launch_bounds__(BLOCK) void radix_digit_base_kernel(u32* __restrict__ ghist,
                                                                 u32 passes,
                                                                 u32* __restrict__ base) {
  __shared__ u32 tmp[WAVES];
  for (u32 p = 0; p < passes; ++p) {
    u32 total;
    const u32 x = ghist[p * RADIX + threadIdx.x];
    base[p * RADIX + threadIdx.x] = block_exclusive_sum(x, total, tmp);
    ghist[p * RADIX + threadIdx.x] = 0;
  }
}

// Sort status word: bits 63..62 flag (1 aggregate, 2 inclusive prefix),
// bits 61..48 the pass epoch, bits 47..0 the count.  A word from another
// epoch reads as "not published", so the array is zeroed only when it is
// allocated.
constexpr u32 SORT_EPOCH_BITS = 14;
constexpr u64 SORT_VAL = (u64{1} << 48) - 1;

__device__ __forceinline__ u64 sort_status(u64 flag, u32 epoch, u64 value) {
  return (flag << 62) | (u64(epoch) << 48) | (value & SORT_VAL);
}

// Look-back of one digit over the tiles before `tile`: reads SORT_LB_WIN
// predecessors at a time (their loads in flight together), sums aggregates
// back to the nearest inclusive prefix, returns the number of records with
// this digit in those tiles and publishes the inclusive count.  Bounded like
// lookback_exclusive.
constexpr int SORT_LB_WIN = 4;

__device__ __forceinline__ u64 digit_lookback(u64* status, u32 epoch, u32 tile, u32 digit,
                                              u64 count, u32* err, bool& ok) {
  ok = true;
  u64* me = status + u64(tile) * RADIX + digit;
  if (tile == 0) {
    lb_store(me, sort_status(2, epoch, count));
    return 0;
  }
  lb_store(me, sort_status(1, epoch, count));
  u64 excl = 0;
  i64 t = static_cast<i64>(tile) - 1;  // next tile to read
  u32 spins = 0;
  while (true) {
    u64 win[SORT_LB_WIN];
#pragma unroll
    for (int j = 0; j < SORT_LB_WIN; ++j)
      win[j] = (t - j >= 0) ? lb_load(status + u64(t - j) * RADIX + digit)
                            : sort_status(2, epoch, 0);  // before tile 0: an empty prefix
    int used = 0;
    bool done = false;
#pragma unroll
    for (int j = 0; j < SORT_LB_WIN; ++j) {
      if (done || used < j) continue;  // stop at the first inclusive or unpublished word
      const u64 w = win[j];
      const bool current = ((w >> 48) & ((1u << SORT_EPOCH_BITS) - 1)) == epoch;
      const u64 flag = current ? (w >> 62) : 0;
      if (flag == 0) continue;
      excl += w & SORT_VAL;
      ++used;
      done = flag == 2;
    }
    if (done) break;
    t -= used;
    if (used == 0) {
      if (++spins > LB_SPIN_LIMIT) {
        atomicOr(err, ERR_SPIN);
        ok = false;
        return 0;
      }
      __builtin_amdgcn_s_sleep(1);
    }
  }
  lb_store(me, sort_status(2, epoch, excl + count));
  return excl;
}

// THREADS = 256 or 512 threads per tile of THREADS * SORT_IPT records; the
// first RADIX threads own one digit each for the scan and the look-back.
template <int THREADS, bool HAS_D, bool HAS_V, bool DEDUP>
__global__ __launch_bounds__(THREADS) void radix_onesweep_kernel(
    const u32* __restrict__ s_in, const u32* __restrict__ d_in, const u64* __restrict__ v_in,
    u32* __restrict__ s_out, u32* __restrict__ d_out, u64* __restrict__ v_out, u64 m, u32 dbits,
    u32 shift, u32 mask, const u32* __restrict__ digit_base, u64* status, u32 epoch,
    unsigned long long* tile_counter, unsigned long long tile_base_id, u32* err) {
  constexpr int NW = THREADS / WAVE;
  constexpr int TILE = THREADS * SORT_IPT;
  static_assert(THREADS >= RADIX, "one thread per digit");
  __shared__ u32 wcnt[NW][RADIX];
  __shared__ u32 delta[RADIX];  // global position - tile position, per digit
  __shared__ u64 stage[TILE];
  __shared__ u32 scan_tmp[NW];
  __shared__ u32 s_ok;
  for (int i = threadIdx.x; i < NW * RADIX; i += THREADS) (&wcnt[0][0])[i] = 0;
  if (threadIdx.x == 0) s_ok = 1;
  // dynamic tile ids in dispatch order from a counter that only grows
  __shared__ u32 s_tile;
  if (threadIdx.x == 0) s_tile = static_cast<u32>(atomicAdd(tile_counter, 1ull) - tile_base_id);
  __syncthreads();
  const u32 tile = s_tile;
  const int lane = lane_id(), w = wave_id();
  const u64 tile_base = u64(tile) * TILE;
  const u64 base = tile_base + u64(w) * (WAVE * SORT_IPT);
  u32 ks[SORT_IPT], kd[SORT_IPT], dig[SORT_IPT], rank[SORT_IPT];
  u64 kv[SORT_IPT];
#pragma unroll
  for (int k = 0; k < SORT_IPT; ++k) {
    const u64 e = min(base + u64(k) * WAVE + lane, m - 1);
    ks[k] = s_in[e];
    kd[k] = HAS_D ? d_in[e] : 0u;
    kv[k] = HAS_V ? v_in[e] : 0ull;
  }
#pragma unroll
  for (int k = 0; k < SORT_IPT; ++k) {
    const bool valid = base + u64(k) * WAVE + lane < m;
    const u64 key = sort_key_of<HAS_D>(ks[k], kd[k], dbits);
    const u32 digit = valid ? static_cast<u32>((key >> shift) & mask) : u32(RADIX);
    u64 peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < RADIX_BITS; ++b) {
      const bool bit = (digit >> b) & 1u;
      const u64 bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    dig[k] = digit;
    rank[k] = 0;
    if (valid) {
      const u32 old = wcnt[w][digit];
      rank[k] = old + mbcnt(peers);
      if (lane == __ffsll(static_cast<long long>(peers)) - 1)
        wcnt[w][digit] = old + static_cast<u32>(__popcll(peers));
    }
  }
  __syncthreads();
  {
    // thread t < RADIX owns digit t: tile-local start, per-wave offsets,
    // global base (the exclusive sum spans the first RADIX threads; the other
    // waves add zeros after them)
    const u32 dg = threadIdx.x;
    u32 tot = 0;
    if (dg < RADIX) {
#pragma unroll
      for (int ww = 0; ww < NW; ++ww) tot += wcnt[ww][dg];
    }
    u32 all;
    const u32 start = block_exclusive_sum(tot, all, scan_tmp);
    if (dg < RADIX) {
      u32 run = start;
#pragma unroll
      for (int ww = 0; ww < NW; ++ww) {
        const u32 c = wcnt[ww][dg];
        wcnt[ww][dg] = run;
        run += c;
      }
      if (dg <= mask) {  // digits beyond the pass's width hold no records
        bool ok;
        const u64 before = digit_lookback(status, epoch, tile, dg, tot, err, ok);
        delta[dg] = digit_base[dg] + static_cast<u32>(before) - start;
        if (!ok) s_ok = 0;
      }
    }
  }
  __syncthreads();
  if (!s_ok) return;
  const u32 count = static_cast<u32>((m - tile_base) < TILE ? (m - tile_base) : TILE);
  u32* st_s = reinterpret_cast<u32*>(stage);
  u32* st_d = st_s + TILE;
#pragma unroll
  for (int k = 0; k < SORT_IPT; ++k) {
    if (dig[k] < RADIX) {
      const u32 lp = wcnt[w][dig[k]] + rank[k];
      st_s[lp] = ks[k];
      if (HAS_D) st_d[lp] = kd[k];
    }
  }
  __syncthreads();
  // Position j of the sorted tile is handled by lane j % 64 of wave
  // (j / 64) % NW, so each wave owns aligned groups of 64 positions.
  u32 pos[SORT_IPT];
  u64 heads[SORT_IPT];  // with DEDUP: lanes that start a run of equal (src, dst)
#pragma unroll
  for (int r = 0; r < SORT_IPT; ++r) {
    const u32 j = r * THREADS + threadIdx.x;
    pos[r] = 0;
    bool head = true;
    if (j < count) {
      const u32 sj = st_s[j];
      const u32 dj = HAS_D ? st_d[j] : 0u;
      const u32 digit = static_cast<u32>((sort_key_of<HAS_D>(sj, dj, dbits) >> shift) & mask);
      pos[r] = delta[digit] + j;
      s_out[pos[r]] = sj;
      if (HAS_D) d_out[pos[r]] = dj;
      if (DEDUP) head = (lane == 0) || st_s[j - 1] != sj || st_d[j - 1] != dj;
    }
    heads[r] = DEDUP ? __ballot(head) : 0ull;
  }
  if (!HAS_V) return;
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SORT_IPT; ++k)
    if (dig[k] < RADIX) stage[wcnt[w][dig[k]] + rank[k]] = kv[k];
  __syncthreads();
#pragma unroll
  for (int r = 0; r < SORT_IPT; ++r) {
    const u32 j = r * THREADS + threadIdx.x;
    u64 vj = (j < count) ? stage[j] : VAL_NONE;
    if (DEDUP) {
      // the head of each run keeps the run's lightest key, the others are
      // marked dead.  Runs are cut at the 64-position groups; a longer run
      // keeps one survivor per group, which later rounds remove.  Segmented
      // suffix minimum over the group: after the step with offset o, lane l
      // holds the minimum over [l, min(l + 2o, end of l's run)).
      const u64 above = (lane == WAVE - 1) ? 0ull : (heads[r] & (~0ull << (lane + 1)));
      const u32 end = above ? static_cast<u32>(__ffsll(static_cast<long long>(above)) - 1) : WAVE;
      u64 x = vj;
#pragma unroll
      for (int off = 1; off < WAVE; off <<= 1) {
        const u64 y = __shfl_down(static_cast<unsigned long long>(x), off, WAVE);
        if (u32(lane) + off < end) x = y < x ? y : x;
      }
      vj = ((heads[r] >> lane) & 1ull) ? x : VAL_NONE;
    }
    if (j < count) v_out[pos[r]] = vj;
  }
}

u64 sort_work_words(u64 m) {
  const u64 tiles = ceil_div_u32(m ? m : 1, SORT_TILE);
  return 2 * (tiles * RADIX + 1) + 2 * u64(MAX_PASSES) * RADIX;
}

void SortWork::bind(hipStream_t st, u32* p, u64 m) {
  const u64 tiles = ceil_div_u32(m ? m : 1, SORT_TILE);
  if (p != words || tiles > tiles_cap) {
    // fresh memory: zero the status words, the tile counter and the digit
    // histograms once
    words = p;
    tiles_cap = tiles;
    fill_u64(st, reinterpret_cast<u64*>(p), tiles * RADIX + 1, 0);
    fill_u32(st, p + 2 * (tiles * RADIX + 1), u64(MAX_PASSES) * RADIX, 0);
    epoch = 0;
    tiles_issued = 0;
  }
}

u32 sort_threads(const SortWork& w, u64 m) {
  if (w.threads == 256 || w.threads == 512) return w.threads;
  return m >= w.wide_min ? 512u : 256u;
}

template <bool HAS_D, bool HAS_V, bool DEDUP>
static void launch_onesweep(hipStream_t st, u32 threads, u32 tiles, const u32* si, const u32* di,
                            const u64* vi, u32* so, u32* dout, u64* vo, u64 m, u32 dbits, u32 shift,
                            u32 mask, const u32* base, SortWork& w, u32* err) {
  u64* status = reinterpret_cast<u64*>(w.words);
  auto* counter = reinterpret_cast<unsigned long long*>(status + w.tiles_cap * RADIX);
  if (threads == 512)
    radix_onesweep_kernel<512, HAS_D, HAS_V, DEDUP><<<tiles, 512, 0, st>>>(
        si, di, vi, so, dout, vo, m, dbits, shift, mask, base, status, w.epoch, counter,
        w.tiles_issued, err);
  else
    radix_onesweep_kernel<256, HAS_D, HAS_V, DEDUP><<<tiles, 256, 0, st>>>(
        si, di, vi, so, dout, vo, m, dbits, shift, mask, base, status, w.epoch, counter,
        w.tiles_issued, err);
}

bool radix_sort(hipStream_t st, u32* s, u32* d, u64* v, u32* s_alt, u32* d_alt, u64* v_alt, u64 m,
                u32 dbits, u32 total_bits, SortWork& work, u32* err, bool dedup_last) {
  if (m <= 1 || total_bits == 0) return false;
  const u32 threads = sort_threads(work, m);
  const u32 tiles = ceil_div_u32(m, u64(threads) * SORT_IPT);
  if (tiles > work.tiles_cap) throw Error(KMST_EINVAL, "sort workspace too small");
  if (total_bits > 64) throw Error(KMST_EINVAL, "sort key wider than 64 bits");
  const DigitPlan pl = plan_digits(total_bits, work.balanced);
  const u32 passes = pl.passes;
  u32* ghist = work.words + 2 * (work.tiles_cap * RADIX + 1);  // zero between sorts
  u32* base = ghist + MAX_PASSES * RADIX;
  {
    const bool use_d = d && dbits;  // d is only carried along otherwise
    KTScope t(st, K_SORT_HIST, (use_d ? 8.0 : 4.0) * m);
    // few, fat blocks: every block adds its histograms to the same global bins.
    // From 2^26 records on, 1024 blocks hide more load latency than their
    // extra global adds cost (plain Boruvka's 2^29-record sorts: 1.65 -> 1.15
    // ms of histograms per solve); below, 512 (filter-Boruvka at GNM 2^28:
    // 0.29 ms with 512, 0.30 with 1024, 0.38 with 2048)
    const u32 grid = grid_for((m + UP_RPT - 1) / UP_RPT, m >= (u64{1} << 26) ? 1024 : 512);
    const bool vec = ((reinterpret_cast<uintptr_t>(s) |
                       (use_d ? reinterpret_cast<uintptr_t>(d) : 0)) & 15) == 0;
    if (use_d) radix_upsweep_kernel<true><<<grid, BLOCK, 0, st>>>(s, d, m, dbits, pl, ghist, vec);
    else radix_upsweep_kernel<false><<<grid, BLOCK, 0, st>>>(s, d, m, dbits, pl, ghist, vec);
    KMST_HIP(hipGetLastError());
    radix_digit_base_kernel<<<1, BLOCK, 0, st>>>(ghist, passes, base);
    KMST_HIP(hipGetLastError());
  }
  const bool can_dedup = dedup_last && d && v;
  bool alt = false;
  for (u32 p = 0; p < passes; ++p) {
    u32* si = alt ? s_alt : s;
    u32* di = alt ? d_alt : d;
    u64* vi = alt ? v_alt : v;
    u32* so = alt ? s : s_alt;
    u32* dout = alt ? d : d_alt;
    u64* vo = alt ? v : v_alt;
    if (++work.epoch == (1u << SORT_EPOCH_BITS)) {
      // epochs wrapped: clear the status words (the counter keeps growing)
      fill_u64(st, reinterpret_cast<u64*>(work.words), work.tiles_cap * RADIX, 0);
      work.epoch = 1;
    }
    KTScope t(st, K_SORT_SCATTER, 2.0 * m * (4.0 + (d ? 4.0 : 0.0) + (v ? 8.0 : 0.0)));
    const u32 shift = pl.shift[p], mask = (1u << pl.width[p]) - 1;
    const u32* b = base + p * RADIX;
    const bool dd = can_dedup && p + 1 == passes;
    if (dd)
      launch_onesweep<true, true, true>(st, threads, tiles, si, di, vi, so, dout, vo, m, dbits,
                                        shift, mask, b, work, err);
    else if (d && v)
      launch_onesweep<true, true, false>(st, threads, tiles, si, di, vi, so, dout, vo, m, dbits,
                                         shift, mask, b, work, err);
    else if (d)
      launch_onesweep<true, false, false>(st, threads, tiles, si, di, vi, so, dout, vo, m, dbits,
                                          shift, mask, b, work, err);
    else if (v)
      launch_onesweep<false, true, false>(st, threads, tiles, si, di, vi, so, dout, vo, m, dbits,
                                          shift, mask, b, work, err);
    else
      launch_onesweep<false, false, false>(st, threads, tiles, si, di, vi, so, dout, vo, m, dbits,
                                           shift, mask, b, work, err);
    KMST_HIP(hipGetLastError());
    work.tiles_issued += tiles;
    alt = !alt;
  }
  return alt;
}

// The same sort for a single u64 array ordered by its low bits (the MST keys
// by (min, max) at the end of a solve): 8-byte records instead of 16.
__global__ __launch_bounds__(BLOCK) void radix_upsweep_u64_kernel(const u64* __restrict__ v,
                                                                  u64 m, u64 mask,
                                                                  const DigitPlan pl,
                                                                  u32* __restrict__ ghist) {
  const u32 passes = pl.passes;
  __shared__ u32 h[WAVES][MAX_PASSES][RADIX];
  for (int i = threadIdx.x; i < WAVES * MAX_PASSES * RADIX; i += BLOCK) (&h[0][0][0])[i] = 0;
  __syncthreads();
  const int w = wave_id();
  const u64 stride = u64(gridDim.x) * (BLOCK * UP_RPT);
  for (u64 base = u64(blockIdx.x) * (BLOCK * UP_RPT); base < m; base += stride) {
    u64 key[UP_RPT];
#pragma unroll
    for (int k = 0; k < UP_RPT; ++k)
      key[k] = v[min(base + u64(k) * BLOCK + threadIdx.x, m - 1)] & mask;
#pragma unroll
    for (int k = 0; k < UP_RPT; ++k) {
      if (base + u64(k) * BLOCK + threadIdx.x >= m) continue;
      for (u32 p = 0; p < passes; ++p)
        atomicAdd(&h[w][p][plan_digit(key[k], pl, p)], 1u);
    }
  }
  __syncthreads();
  for (u32 i = threadIdx.x; i < passes * RADIX; i += BLOCK) {
    u32 c = 0;
#pragma unroll
    for (int ww = 0; ww < WAVES; ++ww) c += (&h[ww][0][0])[i];
    if (c) atomicAdd(ghist + i, c);
  }
}

template <int THREADS>
__global__ __launch_bounds__(THREADS) void radix_onesweep_u64_kernel(
    const u64* __restrict__ v_in, u64* __restrict__ v_out, u64 m, u64 mask, u32 shift,
    u32 dmask, const u32* __restrict__ digit_base, u64* status, u32 epoch,
    unsigned long long* tile_counter, unsigned long long tile_base_id, u32* err) {
  constexpr int NW = THREADS / WAVE;
  constexpr int TILE = THREADS * SORT_IPT;
  static_assert(THREADS >= RADIX, "one thread per digit");
  __shared__ u32 wcnt[NW][RADIX];
  __shared__ u32 delta[RADIX];
  __shared__ u64 stage[TILE];
  __shared__ u32 scan_tmp[NW];
  __shared__ u32 s_ok;
  __shared__ u32 s_tile;
  for (int i = threadIdx.x; i < NW * RADIX; i += THREADS) (&wcnt[0][0])[i] = 0;
  if (threadIdx.x == 0) {
    s_ok = 1;
    s_tile = static_cast<u32>(atomicAdd(tile_counter, 1ull) - tile_base_id);
  }
  __syncthreads();
  const u32 tile = s_tile;
  const int lane = lane_id(), w = wave_id();
  const u64 tile_base = u64(tile) * TILE;
  const u64 base = tile_base + u64(w) * (WAVE * SORT_IPT);
  u64 kv[SORT_IPT];
  u32 dig[SORT_IPT], rank[SORT_IPT];
#pragma unroll
  for (int k = 0; k < SORT_IPT; ++k) kv[k] = v_in[min(base + u64(k) * WAVE + lane, m - 1)];
#pragma unroll
  for (int k = 0; k < SORT_IPT; ++k) {
    const bool valid = base + u64(k) * WAVE + lane < m;
    const u32 digit = valid ? static_cast<u32>(((kv[k] & mask) >> shift) & dmask) : u32(RADIX);
    u64 peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < RADIX_BITS; ++b) {
      const bool bit = (digit >> b) & 1u;
      const u64 bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    dig[k] = digit;
    rank[k] = 0;
    if (valid) {
      const u32 old = wcnt[w][digit];
      rank[k] = old + mbcnt(peers);
      if (lane == __ffsll(static_cast<long long>(peers)) - 1)
        wcnt[w][digit] = old + static_cast<u32>(__popcll(peers));
    }
  }
  __syncthreads();
  {
    const u32 dg = threadIdx.x;
    u32 tot = 0;
    if (dg < RADIX) {
#pragma unroll
      for (int ww = 0; ww < NW; ++ww) tot += wcnt[ww][dg];
    }
    u32 all;
    const u32 start = block_exclusive_sum(tot, all, scan_tmp);
    if (dg < RADIX) {
      u32 run = start;
#pragma unroll
      for (int ww = 0; ww < NW; ++ww) {
        const u32 c = wcnt[ww][dg];
        wcnt[ww][dg] = run;
        run += c;
      }
      if (dg <= dmask) {  // digits beyond the pass's width hold no keys
        bool ok;
        const u64 before = digit_lookback(status, epoch, tile, dg, tot, err, ok);
        delta[dg] = digit_base[dg] + static_cast<u32>(before) - start;
        if (!ok) s_ok = 0;
      }
    }
  }
  __syncthreads();
  if (!s_ok) return;
#pragma unroll
  for (int k = 0; k < SORT_IPT; ++k)
    if (dig[k] < RADIX) stage[wcnt[w][dig[k]] + rank[k]] = kv[k];
  __syncthreads();
  const u32 count = static_cast<u32>((m - tile_base) < TILE ? (m - tile_base) : TILE);
  for (u32 j = threadIdx.x; j < count; j += THREADS) {
    const u64 x = stage[j];
    v_out[delta[static_cast<u32>(((x & mask) >> shift) & dmask)] + j] = x;
  }
}

bool radix_sort_u64(hipStream_t st, u64* v, u64* v_alt, u64 m, u32 bits, SortWork& work,
                    u32* err, u32 lo) {
  if (m <= 1 || bits == 0) return false;
  const u32 threads = sort_threads(work, m);
  const u32 tiles = ceil_div_u32(m, u64(threads) * SORT_IPT);
  if (tiles > work.tiles_cap) throw Error(KMST_EINVAL, "sort workspace too small");
  if (lo + bits > 64) throw Error(KMST_EINVAL, "sort key wider than 64 bits");
  DigitPlan pl = plan_digits(bits, work.balanced);
  for (u32 p = 0; p < pl.passes; ++p) pl.shift[p] = static_cast<u8>(pl.shift[p] + lo);
  const u32 passes = pl.passes;
  // bits above the key are masked off, bits below it are never a digit
  const u64 mask = lo + bits >= 64 ? ~0ull : ((u64{1} << (lo + bits)) - 1);
  u32* ghist = work.words + 2 * (work.tiles_cap * RADIX + 1);  // zero between sorts
  u32* base = ghist + MAX_PASSES * RADIX;
  {
    KTScope t(st, K_SORT_HIST, 8.0 * m);
    radix_upsweep_u64_kernel<<<grid_for((m + UP_RPT - 1) / UP_RPT, 512), BLOCK, 0, st>>>(
        v, m, mask, pl, ghist);
    KMST_HIP(hipGetLastError());
    radix_digit_base_kernel<<<1, BLOCK, 0, st>>>(ghist, passes, base);
    KMST_HIP(hipGetLastError());
  }
  bool alt = false;
  u64* status = reinterpret_cast<u64*>(work.words);
  auto* counter = reinterpret_cast<unsigned long long*>(status + work.tiles_cap * RADIX);
  for (u32 p = 0; p < passes; ++p) {
    if (++work.epoch == (1u << SORT_EPOCH_BITS)) {
      fill_u64(st, status, work.tiles_cap * RADIX, 0);
      work.epoch = 1;
    }
    KTScope t(st, K_SORT_SCATTER, 16.0 * m);
    const u32 shift = pl.shift[p], dmask = (1u << pl.width[p]) - 1;
    if (threads == 512)
      radix_onesweep_u64_kernel<512><<<tiles, 512, 0, st>>>(
          alt ? v_alt : v, alt ? v : v_alt, m, mask, shift, dmask, base + p * RADIX, status,
          work.epoch, counter, work.tiles_issued, err);
    else
      radix_onesweep_u64_kernel<256><<<tiles, 256, 0, st>>>(
          alt ? v_alt : v, alt ? v : v_alt, m, mask, shift, dmask, base + p * RADIX, status,
          work.epoch, counter, work.tiles_issued, err);
    KMST_HIP(hipGetLastError());
    work.tiles_issued += tiles;
    alt = !alt;
  }
  return alt;
}

// ============================================================== K7 dedup
// After the (src, dst) sort, keep the lightest record of each run of equal
// (src, dst) (EdgeProcessor::remove_duplicates, duplicate_detection.hpp:26-64)
// and mark the rest dead.  Runs are cut every 64 records so a head thread
// scans a bounded range; a long run keeps one survivor per 64 records, which
// later rounds remove.
__global__ __launch_bounds__(BLOCK) void dedup_kernel(const u32* __restrict__ s,
                                                      const u32* __restrict__ d,
                                                      u64* __restrict__ v, u64 m) {
  const u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x;
  if (i >= m) return;
  const u32 si = s[i], di = d[i];
  const bool head = (i % WAVE == 0) || s[i - 1] != si || d[i - 1] != di;
  if (!head) return;
  u64 j = i + 1;
  while (j < m && (j % WAVE) != 0 && s[j] == si && d[j] == di) ++j;
  if (j == i + 1) return;
  u64 mn = v[i];
  for (u64 k = i + 1; k < j; ++k) {
    const u64 x = v[k];
    mn = x < mn ? x : mn;
  }
  v[i] = mn;
  for (u64 k = i + 1; k < j; ++k) v[k] = VAL_NONE;
}

void dedup_mark(hipStream_t st, const u32* s, const u32* d, u64* v, u64 m) {
  if (m == 0) return;
  KTScope t(st, K_DEDUP, 16.0 * m);
  dedup_kernel<<<ceil_div_u32(m, BLOCK), BLOCK, 0, st>>>(s, d, v, m);
  KMST_HIP(hipGetLastError());
}

// ============================================================== K9 output
// MST keys -> canonical (min, max, w): packed keys decode directly,
// positional keys gather the twin record (GetMstEdge, misc.hpp:145-206).
struct MstDecode {
  u32 vbits;
  bool packed;
  const u32* in_src;
  const u32* in_dst;
  u32* out_s;
  u32* out_d;
  u32* out_w;
};

// Writes key's edge at output position i; returns its weight.
__device__ __forceinline__ u32 mst_write(const MstDecode& o, u64 key, u64 i) {
  u32 a, b, x;
  if (o.packed) {
    const u64 mask = (u64{1} << o.vbits) - 1;
    a = static_cast<u32>((key >> o.vbits) & mask);
    b = static_cast<u32>(key & mask);
    x = static_cast<u32>(key >> (2 * o.vbits));
  } else {
    // positional key (w << 32) | uid: the weight is the key's, the ends are
    // gathered from the record at uid
    const u32 uid = static_cast<u32>(key);
    const u32 s = o.in_src[uid], d = o.in_dst[uid];
    a = s < d ? s : d;
    b = s < d ? d : s;
    x = static_cast<u32>(key >> 32);
  }
  o.out_s[i] = a;
  o.out_d[i] = b;
  o.out_w[i] = x;
  return x;
}

// Adds the workgroup's weight sums to *total (one atomic per workgroup).
template <int THREADS>
__device__ __forceinline__ void add_block_sum(u64 sum, u64* s_sum, u64* total) {
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) sum += __shfl_xor(sum, o, WAVE);
  if (lane_id() == 0) s_sum[wave_id()] = sum;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 t = 0;
    for (int w = 0; w < THREADS / WAVE; ++w) t += s_sum[w];
    if (t) atomicAdd(reinterpret_cast<unsigned long long*>(total), t);
  }
}

__global__ __launch_bounds__(BLOCK) void mst_output_kernel(const u64* __restrict__ keys,
                                                           u64 count, MstDecode o, u64* total) {
  __shared__ u64 s_sum[WAVES];
  u64 sum = 0;
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < count; i += u64(gridDim.x) * BLOCK)
    sum += mst_write(o, keys[i], i);
  add_block_sum<BLOCK>(sum, s_sum, total);
}

void mst_output(hipStream_t st, const u64* keys, u64 count, u32 vbits, bool packed,
                const u32* in_src, const u32* in_dst, u32* out_s, u32* out_d, u32* out_w,
                u64* total_dev) {
  if (count == 0) return;
  // the key, the output triple, and for positional keys two 4-byte gathers
  KTScope t(st, K_GATHER, (packed ? 20.0 : 28.0) * count);
  const MstDecode o{vbits, packed, in_src, in_dst, out_s, out_d, out_w};
  mst_output_kernel<<<grid_for(count, 1024), BLOCK, 0, st>>>(keys, count, o, total_dev);
  KMST_HIP(hipGetLastError());
}

// Output order from keys sorted by group only.  A key's place inside its
// group is the number of group members with a smaller order field (an MSF
// holds each (min, max) once, so fields and places are distinct); the group's
// first key sits at the key's sorted position minus the members before it.
// A workgroup stages its BLOCK keys' fields with MST_GROUP_SMALL more on each
// side in LDS, and every thread reads its neighbours there, up to
// MST_GROUP_SMALL either way: groups hold one or two keys at GNM and RGG2D.
// The first key of a longer group (a hub's edges) lists the group.  Positions
// outside [0, count) hold all ones, which no group shares (fields are
// narrower than 64 bits).
__global__ __launch_bounds__(BLOCK) void mst_rank_kernel(const u64* __restrict__ keys, u64 count,
                                                         u64 fmask, u32 lbits, MstDecode o,
                                                         u32* __restrict__ list,
                                                         u32* __restrict__ list_count,
                                                         u64* total) {
  constexpr u32 HALO = MST_GROUP_SMALL;
  constexpr u32 SPAN = BLOCK + 2 * HALO;
  __shared__ u64 f_lds[SPAN];
  __shared__ u64 s_sum[WAVES];
  u64 sum = 0;
  const u32 t = threadIdx.x;
  for (u64 base = u64(blockIdx.x) * BLOCK; base < count; base += u64(gridDim.x) * BLOCK) {
    __syncthreads();  // the previous tile's reads are done
    for (u32 j = t; j < SPAN; j += BLOCK) {
      const u64 q = base + j;  // position q - HALO
      f_lds[j] = (q >= HALO && q - HALO < count) ? (keys[q - HALO] & fmask) : ~0ull;
    }
    __syncthreads();
    const u64 p = base + t;
    if (p >= count) continue;
    const u32 c = HALO + t;
    const u64 f = f_lds[c], g = f >> lbits;
    u32 before = 0, after = 0, rank = 0;
    while (before < HALO) {
      const u64 x = f_lds[c - 1 - before];
      if ((x >> lbits) != g) break;
      ++before;
      rank += x < f ? 1u : 0u;
    }
    while (after < HALO) {
      const u64 x = f_lds[c + 1 + after];
      if ((x >> lbits) != g) break;
      ++after;
      rank += x < f ? 1u : 0u;
    }
    // a side is closed when its scan stopped at another group within HALO
    const bool closed = before < HALO && after < HALO;
    if (closed && before + 1 + after <= MST_GROUP_SMALL) {
      sum += mst_write(o, keys[p], p - before + rank);
    } else if (before == 0) {
      list[atomicAdd(list_count, 1u)] = static_cast<u32>(p);
    }
  }
  add_block_sum<BLOCK>(sum, s_sum, total);
}

// Listed groups: one workgroup per group finds the group's length (up to
// MST_GROUP_MAX + MST_CHUNK keys), sorts the group's keys by order field with
// a bitonic network in LDS (padded to a power of two with all-ones keys,
// whose field exceeds every real one) and writes them out.  A longer group is
// left to the host (flag).
__global__ __launch_bounds__(MST_CHUNK) void mst_rank_long_kernel(
    const u64* __restrict__ keys, u64 count, u64 fmask, u32 lbits, MstDecode o,
    const u32* __restrict__ list, const u32* __restrict__ list_count, u32* too_long, u64* total) {
  __shared__ u64 sk[MST_GROUP_MAX];
  __shared__ u64 s_sum[MST_CHUNK / WAVE];
  __shared__ u32 s_end;
  const u32 t = threadIdx.x;
  const u32 groups = *list_count;
  u64 sum = 0;
  for (u32 e = blockIdx.x; e < groups; e += gridDim.x) {
    const u64 start = list[e];
    const u64 g = (keys[start] & fmask) >> lbits;
    if (t == 0) s_end = NONE32;
    __syncthreads();
    u32 len = NONE32;  // the group's length, the same in every thread
    for (u32 b = 0; b < MST_GROUP_MAX + MST_CHUNK; b += MST_CHUNK) {
      const u64 j = start + b + t;
      if (j >= count || ((keys[j] & fmask) >> lbits) != g) atomicMin(&s_end, b + t);
      __syncthreads();
      len = s_end;
      __syncthreads();
      if (len != NONE32) break;
    }
    if (len > MST_GROUP_MAX) {
      if (t == 0) atomicOr(too_long, 1u);
      continue;
    }
    u32 size = MST_CHUNK;
    while (size < len) size <<= 1;
    for (u32 i = t; i < size; i += MST_CHUNK) sk[i] = i < len ? keys[start + i] : ~0ull;
    __syncthreads();
    for (u32 k = 2; k <= size; k <<= 1) {
      for (u32 j = k >> 1; j > 0; j >>= 1) {
        for (u32 i = t; i < size; i += MST_CHUNK) {
          const u32 x = i ^ j;
          if (x > i) {
            const u64 a = sk[i], b = sk[x];
            const bool up = (i & k) == 0;
            if (((a & fmask) > (b & fmask)) == up) {
              sk[i] = b;
              sk[x] = a;
            }
          }
        }
        __syncthreads();
      }
    }
    for (u32 i = t; i < len; i += MST_CHUNK) sum += mst_write(o, sk[i], start + i);
    __syncthreads();  // sk is reloaded for the next group
  }
  add_block_sum<MST_CHUNK>(sum, s_sum, total);
}

void mst_order(hipStream_t st, const u64* keys, u64 count, u32 vbits, bool packed, u32 lbits,
               const u32* in_src, const u32* in_dst, u32* out_s, u32* out_d, u32* out_w,
               u64* total_dev, u32* list, u32* list_count, u32* too_long) {
  if (count == 0) return;
  if (count > 0xFFFFFFFFull) throw Error(KMST_EINVAL, "more than 2^32 - 1 MST edges");
  const u32 fbits = packed ? 2 * vbits : 32;
  const u64 fmask = fbits >= 64 ? ~0ull : ((u64{1} << fbits) - 1);
  const MstDecode o{vbits, packed, in_src, in_dst, out_s, out_d, out_w};
  KTScope t(st, K_GATHER, (packed ? 20.0 : 28.0) * count);
  // few workgroups: each ends with one atomic on the single total word
  mst_rank_kernel<<<grid_for(count, 1024), BLOCK, 0, st>>>(keys, count, fmask, lbits, o, list,
                                                           list_count, total_dev);
  KMST_HIP(hipGetLastError());
  // listed groups only; without any the workgroups leave at once
  mst_rank_long_kernel<<<1024, MST_CHUNK, 0, st>>>(keys, count, fmask, lbits, o, list,
                                                   list_count, too_long, total_dev);
  KMST_HIP(hipGetLastError());
}

// One workgroup reads RUN_SAMPLES evenly spread sources of the next round's
// records (and the source before each); m comes from the relabel's count.
__global__ __launch_bounds__(1024) void run_starts_kernel(const u32* __restrict__ s,
                                                          const u32* m_dev, u32* count) {
  __shared__ u32 s_c[1024 / WAVE];
  const u64 m = *m_dev;
  u32 c = 0;
  for (u32 k = threadIdx.x; k < RUN_SAMPLES; k += 1024) {
    const u64 i = (u64(k) * m) / RUN_SAMPLES;  // < m when m > 0
    if (m > 0 && (i == 0 || s[i] != s[i - 1])) {
      // low half: run starts; high half: those whose source lies in another
      // 16-entry block (another 128-byte line of the per-vertex minima)
      c += (i == 0 || (s[i] >> 4) != (s[i - 1] >> 4)) ? 0x10001u : 1u;
    }
  }
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, WAVE);
  if (lane_id() == 0) s_c[wave_id()] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 t = 0;
    for (int w = 0; w < 1024 / WAVE; ++w) t += s_c[w];
    *count = t;
  }
}

void sample_run_starts(hipStream_t st, const u32* s, const u32* m_dev, u32* count) {
  KTScope t(st, K_OTHER, 8.0 * RUN_SAMPLES);
  run_starts_kernel<<<1, 1024, 0, st>>>(s, m_dev, count);
  KMST_HIP(hipGetLastError());
}

// ============================================================== filter
__global__ __launch_bounds__(BLOCK) void iota_kernel(u32* p, u64 n) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += u64(gridDim.x) * BLOCK)
    p[i] = static_cast<u32>(i);
}

void iota_u32(hipStream_t st, u32* p, u64 n) {
  if (n == 0) return;
  KTScope t(st, K_FILL, 4.0 * n);
  iota_kernel<<<grid_for(n), BLOCK, 0, st>>>(p, n);
  KMST_HIP(hipGetLastError());
}

// Parent-array update of one round (UpdateParentArray in filter_boruvka_step,
// include/algorithms/hybrid_boruvka_edgefilter.hpp:91-95): every live vertex c
// of this round, named rep[c] in the original numbering (identity in the first
// round), points at its root's name; every surviving root passes its name to
// its new id.
__global__ __launch_bounds__(BLOCK) void gpar_kernel(u64 n, const u32* __restrict__ rep,
                                                     const u32* __restrict__ parent,
                                                     const u64* __restrict__ state,
                                                     const u64* __restrict__ pos,
                                                     u32* __restrict__ gpar,
                                                     u32* __restrict__ rep_next) {
  const u64 c = u64(blockIdx.x) * BLOCK + threadIdx.x;
  if (c >= n) return;
  const u64 sc = state[c];
  if (sc == 0) return;  // no edge this round: keeps its parent-array entry
  const u32 name = rep ? rep[c] : static_cast<u32>(c);
  const u32 p = parent[c];
  gpar[name] = rep ? rep[p] : p;
  if (sc & ST_ROOT) rep_next[static_cast<u32>(pos[c])] = name;
}

void update_parent_array(hipStream_t st, u64 n, const u32* rep, const u32* parent,
                         const u64* state, const u64* pos, u32* gpar, u32* rep_next) {
  if (n == 0) return;
  KTScope t(st, K_GPAR, 32.0 * n);
  gpar_kernel<<<ceil_div_u32(n, BLOCK), BLOCK, 0, st>>>(n, rep, parent, state, pos, gpar,
                                                         rep_next);
  KMST_HIP(hipGetLastError());
}

// k evenly spaced weight samples (NONE32 for dead records); the pivot only
// changes speed, never the result.
__global__ __launch_bounds__(BLOCK) void sample_kernel(const u64* __restrict__ v,
                                                       const u32* __restrict__ wt, u64 m, u32 k,
                                                       u32 wshift, u32* __restrict__ out) {
  const u32 j = blockIdx.x * BLOCK + threadIdx.x;
  if (j >= k) return;
  const u64 i = (u64(j) * m) / k + m / (2ull * k);
  const u64 ii = i < m ? i : m - 1;
  if (!v) {
    out[j] = wt[ii];
    return;
  }
  const u64 x = v[ii];
  out[j] = (x == VAL_NONE) ? NONE32 : static_cast<u32>(x >> wshift);
}

void sample_weights(hipStream_t st, const u64* v, const u32* wt, u64 m, u32 k, u32 wshift,
                    u32* out) {
  if (k == 0 || m == 0) return;
  KTScope t(st, K_SAMPLE, 12.0 * k);
  sample_kernel<<<ceil_div_u32(k, BLOCK), BLOCK, 0, st>>>(v, wt, m, k, wshift, out);
  KMST_HIP(hipGetLastError());
}

// *out = how many of k evenly spread records (the positions of
// sample_weights) join ids at least `span` apart.  One workgroup.
__global__ __launch_bounds__(BLOCK) void sample_far_kernel(const u32* __restrict__ s,
                                                           const u32* __restrict__ d, u64 m,
                                                           u32 k, u32 span, u32* __restrict__ out) {
  __shared__ u32 s_sum[WAVES];
  u32 far = 0;
  for (u32 j = threadIdx.x; j < k; j += BLOCK) {
    const u64 i = (u64(j) * m) / k + m / (2ull * k);
    const u64 ii = i < m ? i : m - 1;
    const u32 a = s[ii], b = d[ii];
    far += (a > b ? a - b : b - a) >= span ? 1u : 0u;
  }
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) far += __shfl_xor(far, o, WAVE);
  if (lane_id() == 0) s_sum[wave_id()] = far;
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 t = 0;
    for (int w = 0; w < WAVES; ++w) t += s_sum[w];
    *out = t;
  }
}

void sample_far(hipStream_t st, const u32* s, const u32* d, u64 m, u32 k, u32 span, u32* out) {
  if (k == 0 || m == 0) return;
  KTScope t(st, K_SAMPLE, 8.0 * k);
  sample_far_kernel<<<1, BLOCK, 0, st>>>(s, d, m, k, span, out);
  KMST_HIP(hipGetLastError());
}

// Two-way split by weight (partition_edges, filter_helpers.hpp:101-139):
// light (w <= pivot) records to out[0, L) in input order (they stay sorted by
// (src, dst)); heavy records to out[L, m) in reverse order (they are
// relabelled and re-sorted before use).  One pass with a decoupled look-back
// over the light counts; heavy offsets follow from them.
__global__ __launch_bounds__(BLOCK) void split_kernel(
    const u32* __restrict__ s_in, const u32* __restrict__ d_in, const KeySrc keys, u64 m,
    u32 pivot, u32 wshift, bool write_heavy, u32* __restrict__ s_out, u32* __restrict__ d_out,
    u64* __restrict__ v_out, u64* lb_status, u32 lb_epoch, unsigned long long* lb_counter,
    unsigned long long lb_base, u32 tiles, u32* light_dev, u32* err) {
  __shared__ u32 s_wave_light[WAVES];
  __shared__ u64 s_excl;
  __shared__ u32 s_ok;
  const u32 tile = lb_tile_id(lb_counter, lb_base);
  const int lane = lane_id(), w = wave_id();
  const u64 tile_base = u64(tile) * CMP_TILE;
  const u64 wbase = tile_base + u64(w) * (WAVE * CMP_IPT);
  u32 ss[CMP_IPT], dd[CMP_IPT];
  u64 vv[CMP_IPT];
  u64 bal[CMP_IPT];
  u32 cnt = 0;
  load_records(s_in, d_in, keys, wbase, m, ss, dd, vv);
#pragma unroll
  for (int k = 0; k < CMP_IPT; ++k) {
    const u64 e = wbase + u64(k) * WAVE + lane;
    // dead records (VAL_NONE) always count as heavy
    const bool light =
        e < m && vv[k] != VAL_NONE && static_cast<u32>(vv[k] >> wshift) <= pivot;
    bal[k] = __ballot(light);
    cnt += static_cast<u32>(__popcll(bal[k]));
  }
  if (lane == 0) s_wave_light[w] = cnt;
  __syncthreads();
  if (w == 0) {
    u64 agg = 0;
#pragma unroll
    for (int ww = 0; ww < WAVES; ++ww) agg += s_wave_light[ww];
    bool ok;
    const u64 excl = lookback_exclusive(lb_status, lb_epoch, tile, agg, err, ok);
    if (lane == 0) {
      s_excl = excl;
      s_ok = ok ? 1u : 0u;
      if (tile == tiles - 1) *light_dev = static_cast<u32>(excl + agg);
    }
  }
  __syncthreads();
  if (!s_ok) return;
  u64 lpos = s_excl;
  u64 before = u64(w) * (WAVE * CMP_IPT);  // records of this tile before this wave
  for (int ww = 0; ww < w; ++ww) lpos += s_wave_light[ww];
  // heavy rank = records before - light records before
  u64 hrank = (tile_base - s_excl) + (before - (lpos - s_excl));
#pragma unroll
  for (int k = 0; k < CMP_IPT; ++k) {
    const u64 e = wbase + u64(k) * WAVE + lane;
    const u64 valid = __ballot(e < m);
    const bool light = (bal[k] >> lane) & 1ull;
    if (e < m && (light || write_heavy)) {
      const u64 p = light ? lpos + mbcnt(bal[k]) : (m - 1) - (hrank + mbcnt(valid & ~bal[k]));
      s_out[p] = ss[k];
      d_out[p] = dd[k];
      v_out[p] = vv[k];
    }
    const u32 nl = static_cast<u32>(__popcll(bal[k]));
    lpos += nl;
    hrank += static_cast<u32>(__popcll(valid)) - nl;
  }
}

void split_light_heavy(hipStream_t st, const u32* s_in, const u32* d_in, const KeySrc& keys,
                       u64 m, u32 pivot, u32 wshift, bool write_heavy, u32* s_out, u32* d_out,
                       u64* v_out, LookbackWork& lb, u32* light_count_dev, u32* err) {
  if (m == 0) {
    fill_u32(st, light_count_dev, 1, 0);
    return;
  }
  const u32 tiles = ceil_div_u32(m, CMP_TILE);
  const u32 epoch = lb.next_call(st, tiles);
  // reads the records; writes (light, plus heavy when kept) are added by
  // the caller once the light count is known
  KTScope t(st, K_SPLIT_SCATTER, (keys.v ? 16.0 : 12.0) * m);
  split_kernel<<<tiles, BLOCK, 0, st>>>(
      s_in, d_in, keys, m, pivot, wshift, write_heavy, s_out, d_out, v_out, lb.words, epoch,
      reinterpret_cast<unsigned long long*>(lb.words + lb.tiles_cap), lb.tiles_issued, tiles,
      light_count_dev, err);
  KMST_HIP(hipGetLastError());
  lb.tiles_issued += tiles;
}

// Vertex compaction before a sparse base case, in passes over the records
// only: the records are sorted by src and symmetric, so the base case's
// vertices are its sources, and the runs of equal src numbered in order give
// them [0, nc) in increasing original id.  run_start_flags marks every run's
// first record; after an exclusive scan of the marks, number_runs writes
// new_id[src] and rep[new id] = src at every run start.
__global__ __launch_bounds__(BLOCK) void run_start_flags_kernel(const u32* __restrict__ s, u64 m,
                                                                u32* __restrict__ flag) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < m; i += u64(gridDim.x) * BLOCK)
    flag[i] = (i == 0 || s[i - 1] != s[i]) ? 1u : 0u;
}

__global__ __launch_bounds__(BLOCK) void number_runs_kernel(const u32* __restrict__ s, u64 m,
                                                            const u32* __restrict__ id,
                                                            u32* __restrict__ new_id,
                                                            u32* __restrict__ rep) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < m; i += u64(gridDim.x) * BLOCK) {
    const u32 v = s[i];
    if (i == 0 || s[i - 1] != v) {
      new_id[v] = id[i];
      rep[id[i]] = v;
    }
  }
}

void run_start_flags(hipStream_t st, const u32* s, u64 m, u32* flag) {
  if (m == 0) return;
  KTScope t(st, K_OTHER, 8.0 * m);
  run_start_flags_kernel<<<grid_for(m), BLOCK, 0, st>>>(s, m, flag);
  KMST_HIP(hipGetLastError());
}

void number_runs(hipStream_t st, const u32* s, u64 m, const u32* id, u32* new_id, u32* rep) {
  if (m == 0) return;
  KTScope t(st, K_OTHER, 8.0 * m);
  number_runs_kernel<<<grid_for(m), BLOCK, 0, st>>>(s, m, id, new_id, rep);
  KMST_HIP(hipGetLastError());
}

// Number of parent-array roots that had edges in the input: the reference
// skips the heavy half when the light edges left at most one such component
// (global_num_vertices, distributed_parent_array.hpp:62-72, used at
// hybrid_boruvka_edgefilter.hpp:289-296).
__global__ __launch_bounds__(BLOCK) void count_roots_kernel(const u32* __restrict__ gpar,
                                                            const u32* __restrict__ seg_end, u64 n,
                                                            u32* count) {
  __shared__ u32 s_c[WAVES];
  u32 c = 0;
  // 4 vertices per thread and step with 16-byte loads (the arrays come from
  // the cached allocator, so they are 16-byte aligned)
  const u64 n4 = n / 4;
  const u64 stride = u64(gridDim.x) * BLOCK;
  for (u64 q = u64(blockIdx.x) * BLOCK + threadIdx.x; q < n4; q += stride) {
    const uint4 g = reinterpret_cast<const uint4*>(gpar)[q];
    const uint4 h = reinterpret_cast<const uint4*>(seg_end)[q];
    const u32 v = static_cast<u32>(q * 4);
    c += (g.x == v && h.x) + (g.y == v + 1 && h.y) + (g.z == v + 2 && h.z) + (g.w == v + 3 && h.w);
  }
  if (blockIdx.x == 0 && threadIdx.x < n - n4 * 4) {
    const u64 v = n4 * 4 + threadIdx.x;
    c += (gpar[v] == v && seg_end[v]) ? 1u : 0u;
  }
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, WAVE);
  if (lane_id() == 0) s_c[wave_id()] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 t = 0;
    for (int w = 0; w < WAVES; ++w) t += s_c[w];
    if (t) atomicAdd(count, t);
  }
}

void count_roots(hipStream_t st, const u32* gpar, const u32* seg_end, u64 n, u32* count_dev) {
  fill_u32(st, count_dev, 1, 0);
  if (n == 0) return;
  KTScope t(st, K_OTHER, 8.0 * n);
  count_roots_kernel<<<grid_for((n + 3) / 4, 2048), BLOCK, 0, st>>>(gpar, seg_end, n, count_dev);
  KMST_HIP(hipGetLastError());
}

// Vertices with records outside the largest component -- whose root is not
// `root`, or, with the bitmap `inside` (each partition range's largest
// component), whose bit is clear -- with their
// record ranges [seg_start, seg_end) in the (src, dst)-sorted input (seg_end
// is 0 for vertices without records).  Each block owns OUT_TILES tiles of
// OUT_VPT * BLOCK consecutive vertices, thread t those at t + k * BLOCK of a
// tile (coalesced loads).  A first pass keeps each thread's outside vertices
// as one bit mask per tile; the block then takes one slot range with one
// global atomic (a quarter of the same-address atomics of one tile per
// block: RGG2D 2^24 lists 0.40 -> 0.37 ms, GNM unchanged) and a second pass
// writes the marked vertices, re-reading only their ranges.
// The list holds a block's vertices thread by thread and tile by tile, the
// blocks in the order of their atomics.  counters[0] = list length,
// counters[2..3] = u64 total range length.
constexpr int OUT_VPT = 16;
constexpr int OUT_TILES = 4;

__global__ __launch_bounds__(BLOCK) void outside_list_kernel(
    const u32* __restrict__ gpar, const u32* __restrict__ seg_start,
    const u32* __restrict__ seg_end, u64 n, u32 root, const u64* __restrict__ inside,
    u32* __restrict__ list, u32* __restrict__ seg_lo, u32* __restrict__ seg_hi, u32* counters) {
  static_assert(OUT_VPT <= 32, "one bit per vertex of a tile in a u32 mask");
  __shared__ u32 s_scan[WAVES];
  __shared__ u32 s_base;
  __shared__ u64 s_len[WAVES];
  const u64 b0 = u64(blockIdx.x) * (OUT_TILES * BLOCK * OUT_VPT) + threadIdx.x;
  u32 mask[OUT_TILES];
  u32 mine = 0;
#pragma unroll
  for (int tt = 0; tt < OUT_TILES; ++tt) {
    const u64 v0 = b0 + u64(tt) * (BLOCK * OUT_VPT);
    u32 h[OUT_VPT];
    bool in[OUT_VPT];
#pragma unroll
    for (int k = 0; k < OUT_VPT; ++k) {
      const u64 v = min(v0 + u64(k) * BLOCK, n - 1);
      h[k] = seg_end[v];
      in[k] = inside ? ((inside[v >> 6] >> (v & 63)) & 1ull) != 0 : gpar[v] == root;
    }
    u32 mk = 0;
#pragma unroll
    for (int k = 0; k < OUT_VPT; ++k)
      if (v0 + u64(k) * BLOCK < n && h[k] && !in[k]) mk |= 1u << k;
    mask[tt] = mk;
    mine += static_cast<u32>(__popc(mk));
  }
  u32 total;
  const u32 off = block_exclusive_sum(mine, total, s_scan);
  if (threadIdx.x == 0) s_base = total ? atomicAdd(counters, total) : 0u;
  __syncthreads();
  u32 slot = s_base + off;
  u64 len = 0;
#pragma unroll
  for (int tt = 0; tt < OUT_TILES; ++tt) {
    const u64 v0 = b0 + u64(tt) * (BLOCK * OUT_VPT);
    for (u32 mk = mask[tt]; mk; mk &= mk - 1) {
      const u32 v = static_cast<u32>(v0 + u64(__ffs(mk) - 1) * BLOCK);
      const u32 lo = seg_start[v], hi = seg_end[v];
      list[slot] = v;
      seg_lo[slot] = lo;
      seg_hi[slot] = hi;
      len += hi - lo;
      ++slot;
    }
  }
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) len += __shfl_xor(len, o, WAVE);
  if (lane_id() == 0) s_len[wave_id()] = len;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 t = 0;
    for (int w = 0; w < WAVES; ++w) t += s_len[w];
    if (t)
      atomicAdd(reinterpret_cast<unsigned long long*>(counters + 2),
                static_cast<unsigned long long>(t));
  }
}

void outside_list(hipStream_t st, const u32* gpar, const u32* seg_start, const u32* seg_end,
                  u64 n, u32 root, const u64* inside, u32* list, u32* seg_lo, u32* seg_hi,
                  u32* counters) {
  fill_u32(st, counters, 4, 0);
  if (n == 0) return;
  KTScope t(st, K_OTHER, inside ? 4.0 * n + n / 8.0 : 8.0 * n);
  const u32 blocks = ceil_div_u32(n, u64(OUT_TILES) * BLOCK * OUT_VPT);
  outside_list_kernel<<<blocks, BLOCK, 0, st>>>(gpar, seg_start, seg_end, n, root, inside, list,
                                                seg_lo, seg_hi, counters);
  KMST_HIP(hipGetLastError());
}

// Each block owns BLOCK consecutive listed vertices u and walks their records
// as one flattened range (a thread finds its vertex by binary search in LDS).
// Records (u, d, key) with key in [kmin, kmax] and gpar[d] != gpar[u] leave as
// (gpar[u], gpar[d], key); when d is inside the largest component (its root
// is `root`, or its bit in `inside` is set: d is not listed, so its own range
// is not walked) the reverse record is emitted as well.  A counting
// pass sizes the block's output, one global atomic places it, and an emitting
// pass (whose reads hit the cache) writes it.  Output order is arbitrary;
// positions at or past cap raise ERR_BOUNDS.  With lazy positional keys
// (keys.twin_lo) [kmin, kmax] are weight bounds shifted by 32, and only the
// emitted records' keys are made (their pair_position search in the emitting pass).
__global__ __launch_bounds__(BLOCK) void outside_gather_kernel(
    const u32* __restrict__ list, const u32* __restrict__ seg_lo, const u32* __restrict__ seg_hi,
    u32 n_list, const u32* __restrict__ in_s, const u32* __restrict__ in_d, const KeySrc keys,
    u64 kmin, u64 kmax,
    const u32* __restrict__ gpar, u32 root, const u64* __restrict__ inside,
    u32* __restrict__ s_out, u32* __restrict__ d_out, u64* __restrict__ v_out, u64 cap,
    u32* count, u32* err) {
  __shared__ u64 s_start[BLOCK + 1];
  __shared__ u32 s_u[BLOCK], s_gu[BLOCK], s_lo[BLOCK];
  __shared__ u64 s_scan[WAVES];
  __shared__ u32 s_cnt[WAVES];
  __shared__ u32 s_base, s_cursor;
  const u32 i0 = blockIdx.x * BLOCK;
  const u32 nv = min<u32>(BLOCK, n_list - i0);
  u64 len = 0;
  if (threadIdx.x < nv) {
    const u32 u = list[i0 + threadIdx.x];
    const u32 lo = seg_lo[i0 + threadIdx.x];
    s_u[threadIdx.x] = u;
    s_gu[threadIdx.x] = gpar[u];
    s_lo[threadIdx.x] = lo;
    len = seg_hi[i0 + threadIdx.x] - lo;
  }
  u64 total;
  const u64 start = block_exclusive_sum(len, total, s_scan);
  s_start[threadIdx.x] = start;
  if (threadIdx.x == 0) {
    s_start[BLOCK] = total;
    s_cursor = 0;
  }
  __syncthreads();
  const int lane = lane_id();
  // record t of the block's range: vertex j = last with s_start[j] <= t
  auto visit = [&](u64 t, bool need_key, bool& keep, bool& both, u32& gu, u32& gd, u64& key) {
    keep = both = false;
    if (t >= total) return;
    u32 lo = 0, hi = nv;  // s_start[lo] <= t < s_start[hi]
    while (hi - lo > 1) {
      const u32 mid = (lo + hi) >> 1;
      if (s_start[mid] <= t) lo = mid;
      else hi = mid;
    }
    const u64 e = u64(s_lo[lo]) + (t - s_start[lo]);
    const u32 u = s_u[lo];
    gu = s_gu[lo];
    const u32 d = in_d[e];
    u32 w = 0;
    if (keys.twin_lo) {
      w = keys.w[e];
      if (w < (kmin >> 32) || w > (kmax >> 32)) return;
    } else {
      key = load_key(keys, e, u, d);
      if (key < kmin || key > kmax) return;
    }
    gd = gpar[d];
    keep = gd != gu;
    both = keep && (inside ? ((inside[d >> 6] >> (d & 63)) & 1ull) != 0 : gd == root);
    if (keys.twin_lo && keep && need_key) {
      const u32 x = min(u, d), y = max(u, d);
      const u64 j = pair_position(in_d, y, keys.twin_lo[x], keys.twin_hi[x]);
      if (j == NO_TWIN) {  // the input check saw every twin: not reached
        atomicOr(err, ERR_ASYMMETRIC);
        keep = both = false;
      }
      key = (u64(w) << 32) | j;
    }
  };
  u32 mine = 0;
  for (u64 t0 = 0; t0 < total; t0 += BLOCK) {
    bool keep, both;
    u32 gu, gd;
    u64 key;
    visit(t0 + threadIdx.x, false, keep, both, gu, gd, key);
    mine += (keep ? 1u : 0u) + (both ? 1u : 0u);
  }
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) mine += __shfl_xor(mine, o, WAVE);
  if (lane == 0) s_cnt[wave_id()] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 t = 0;
    for (int w = 0; w < WAVES; ++w) t += s_cnt[w];
    s_base = t ? atomicAdd(count, t) : 0u;
  }
  __syncthreads();
  for (u64 t0 = 0; t0 < total; t0 += BLOCK) {
    bool keep, both;
    u32 gu = 0, gd = 0;
    u64 key = 0;
    visit(t0 + threadIdx.x, true, keep, both, gu, gd, key);
    const u64 bk = __ballot(keep), bb = __ballot(both);
    const u32 nw = static_cast<u32>(__popcll(bk) + __popcll(bb));
    if (nw == 0) continue;  // uniform per wave
    u32 wbase = 0;
    if (lane == 0) wbase = atomicAdd(&s_cursor, nw);
    wbase = __shfl(wbase, 0, WAVE) + s_base;
    if (keep) {
      const u64 p = u64(wbase) + mbcnt(bk);
      if (p < cap) {
        s_out[p] = gu;
        d_out[p] = gd;
        v_out[p] = key;
      } else {
        atomicOr(err, ERR_BOUNDS);
      }
    }
    if (both) {
      const u64 p = u64(wbase) + static_cast<u32>(__popcll(bk)) + mbcnt(bb);
      if (p < cap) {
        s_out[p] = gd;
        d_out[p] = gu;
        v_out[p] = key;
      } else {
        atomicOr(err, ERR_BOUNDS);
      }
    }
  }
}

void outside_gather(hipStream_t st, const u32* list, const u32* seg_lo, const u32* seg_hi,
                    u32 n_list, u64 total, const u32* in_s, const u32* in_d, const KeySrc& keys,
                    u64 kmin, u64 kmax,
                    const u32* gpar, u32 root, const u64* inside, u32* s_out, u32* d_out,
                    u64* v_out, u64 cap, u32* count, u32* err) {
  fill_u32(st, count, 1, 0);
  if (n_list == 0) return;
  // reads of the ranges (dst + key) and the root gathers; writes are added by
  // the caller
  KTScope t(st, K_RELABEL_SCATTER, (keys.v ? 16.0 : 8.0) * double(total) + 12.0 * n_list);
  outside_gather_kernel<<<ceil_div_u32(n_list, BLOCK), BLOCK, 0, st>>>(
      list, seg_lo, seg_hi, n_list, in_s, in_d, keys, kmin, kmax, gpar, root, inside, s_out, d_out,
      v_out, cap, count, err);
  KMST_HIP(hipGetLastError());
}

// bits[v / 64] bit v % 64 = (gpar[v] == root), for v < n.
__global__ __launch_bounds__(BLOCK) void component_bits_kernel(const u32* __restrict__ gpar, u64 n,
                                                               u32 root, u64* __restrict__ bits) {
  const u64 v = u64(blockIdx.x) * BLOCK + threadIdx.x;
  const u64 b = __ballot(v < n && gpar[v] == root);
  if (lane_id() == 0 && v < n) bits[v >> 6] = b;
}

// One block per partition range r = [part[r], part[r + 1]): giant[r] = the
// most frequent parent-array root among up to BLOCK evenly spaced vertices of
// the range that have records (NONE32 when none has).
__global__ __launch_bounds__(BLOCK) void part_giant_kernel(const u32* __restrict__ gpar,
                                                           const u32* __restrict__ seg_end,
                                                           const u32* __restrict__ part,
                                                           u32* __restrict__ giant) {
  __shared__ u32 s_x[BLOCK];
  __shared__ u64 s_best[WAVES];
  const u32 r = blockIdx.x;
  const u32 lo = part[r], len = part[r + 1] - lo;
  const u32 samples = min<u32>(len, BLOCK);
  u32 x = NONE32;
  if (threadIdx.x < samples) {
    const u32 v = lo + static_cast<u32>((u64(threadIdx.x) * len) / samples);
    if (seg_end[v]) x = gpar[v];
  }
  s_x[threadIdx.x] = x;
  __syncthreads();
  u32 c = 0;
  if (x != NONE32)
    for (u32 j = 0; j < samples; ++j) c += s_x[j] == x ? 1u : 0u;
  u64 key = c ? (u64(c) << 32) | x : 0;  // most samples wins (any root would be correct)
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    const u64 y = __shfl_xor(key, o, WAVE);
    key = y > key ? y : key;
  }
  if (lane_id() == 0) s_best[wave_id()] = key;
  __syncthreads();
  if (threadIdx.x == 0) {
    u64 b = 0;
    for (int w = 0; w < WAVES; ++w) b = s_best[w] > b ? s_best[w] : b;
    giant[r] = b ? static_cast<u32>(b) : NONE32;
  }
}

// bits[v / 64] bit v % 64 = (gpar[v] == giant[range of v]) for v < n.
__global__ __launch_bounds__(BLOCK) void part_bits_kernel(const u32* __restrict__ gpar, u64 n,
                                                          const u32* __restrict__ part, u32 count,
                                                          const u32* __restrict__ giant,
                                                          u64* __restrict__ bits) {
  const u64 v = u64(blockIdx.x) * BLOCK + threadIdx.x;
  bool in = false;
  if (v < n && v < part[count]) {
    u32 lo = 0, hi = count;  // last range r with part[r] <= v
    while (hi - lo > 1) {
      const u32 mid = (lo + hi) >> 1;
      if (part[mid] <= v) lo = mid;
      else hi = mid;
    }
    const u32 g = giant[lo];
    in = g != NONE32 && gpar[v] == g;
  }
  const u64 b = __ballot(in);
  if (lane_id() == 0 && v < n) bits[v >> 6] = b;
}

void partition_giant_bits(hipStream_t st, const u32* gpar, const u32* seg_end, u64 n,
                          const u32* part, u32 count, u32* giant, u64* bits) {
  if (n == 0 || count == 0) return;
  KTScope t(st, K_OTHER, 4.0 * n + n / 8.0 + 8.0 * BLOCK * count);
  part_giant_kernel<<<count, BLOCK, 0, st>>>(gpar, seg_end, part, giant);
  KMST_HIP(hipGetLastError());
  part_bits_kernel<<<ceil_div_u32(n, BLOCK), BLOCK, 0, st>>>(gpar, n, part, count, giant, bits);
  KMST_HIP(hipGetLastError());
}

void component_bits(hipStream_t st, const u32* gpar, u64 n, u32 root, u64* bits) {
  if (n == 0) return;
  KTScope t(st, K_OTHER, 4.0 * n + n / 8.0);
  component_bits_kernel<<<ceil_div_u32(n, BLOCK), BLOCK, 0, st>>>(gpar, n, root, bits);
  KMST_HIP(hipGetLastError());
}

// Scatter of the two lightest weight classes: class 0 (w <= q0) to
// out[off0[u] + ...] and class 1 (q0 < w <= q1) to out[base1 + off1[u] + ...],
// both in input order (off0 / off1: exclusive scans of the unit counts of the
// weights pass).  One wave per 256-record unit, no waits between units.  With
// CHECK the same pass validates the input as validate_kernel does (ids, order,
// symmetry hashes, forward count, every vertex's record range), so the input
// is read once for both.  With LAZY (positional keys) the keys of the written
// records are made here: (w << 32) | pair_position of (min, max) in the
// ranges keys.twin_lo / twin_hi (the record's own position for an s < d
// record that no copy of itself precedes), so that no key array of the whole
// input is built.
// A block takes a chunk of consecutive units per step (two per wave, one
// with the input check, which needs the registers).  The chunk's class 0 and
// class 1 records are contiguous in the output, so they are staged in LDS and
// leave in coalesced runs.  With packed keys the raw weight is staged and the
// key made on the way out, for the written records only: the per-record work
// of this pass (the checks, the symmetry hash, the class tests) costs about as
// much issue time as its HBM traffic takes.
template <bool CHECK>
struct ScatterChunk {
  static constexpr int WAVE_UNITS = CHECK ? 1 : 2;
  static constexpr int UNITS = WAVES * WAVE_UNITS;
  static constexpr int RECORDS = UNITS * int(UNIT_RECORDS);
};

// popcount(mask & lanes-below-me) + acc
__device__ __forceinline__ u32 mbcnt_acc(u64 mask, u32 acc) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<u32>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<u32>(mask), acc));
}

template <bool VEC, bool CHECK, bool LAZY>
__global__ __launch_bounds__(BLOCK) void class_scatter_kernel(
    const u32* __restrict__ s_in, const u32* __restrict__ d_in, const KeySrc keys, u64 m,
    u32 wshift, u32 q0, u32 q1, const u32* __restrict__ off0, const u32* __restrict__ off1,
    u64 base1, u64 out_cap, u32* __restrict__ s_out, u32* __restrict__ d_out,
    u64* __restrict__ v_out, u32* err, const InputCheck chk) {
  constexpr int CS_WAVE_UNITS = ScatterChunk<CHECK>::WAVE_UNITS;
  constexpr int CS_CHUNK = ScatterChunk<CHECK>::UNITS;
  constexpr int CS_CHUNK_RECORDS = ScatterChunk<CHECK>::RECORDS;
  __shared__ u32 st_s[CS_CHUNK_RECORDS];
  __shared__ u32 st_d[CS_CHUNK_RECORDS];
  __shared__ u64 st_v[CS_CHUNK_RECORDS];
  __shared__ u64 s_red[WAVES][3];
  // lazy keys: every wave's searches of the chunk, by stage position
  constexpr int TASKS = LAZY ? CS_WAVE_UNITS * int(UNIT_RECORDS) : 1;
  __shared__ u32 s_task[WAVES][TASKS];
  const int lane = lane_id(), wv = wave_id();
  const u64 units = record_units_dev(m);
  const u64 chunks = (units + CS_CHUNK - 1) / CS_CHUNK;
  const u64 total0 = base1, total1 = out_cap - base1;  // class sizes
  u32 bad = 0, fwd = 0;
  u64 hf = 0, hr = 0;
  // stored keys (never with the input check, which needs the weights)
  const bool stored = !CHECK && keys.v != nullptr;
  // packed keys: the raw weight is staged, the key made on the way out
  const bool late_key = !LAZY && !stored;
  // a chunk's records of this wave.  Without the input check the next
  // chunk's are loaded before the current one is processed, so their latency
  // overlaps the staging and the workgroup barriers (1.33 -> 1.07 ms at RGG2D
  // 2^24).  With the check it gains nothing (GNM 2^24 / 2^28: 1.91 against
  // 1.93 ms) and takes 70 registers instead of 49.
  constexpr bool PREFETCH = !CHECK;
  auto load_chunk = [&](u64 chunk, u32 (&ls)[CS_WAVE_UNITS][4], u32 (&ld)[CS_WAVE_UNITS][4],
                        u32 (&lw)[CS_WAVE_UNITS][4], u64 (&lv)[CS_WAVE_UNITS][4]) {
    const u64 first = chunk * CS_CHUNK;
#pragma unroll
    for (int u = 0; u < CS_WAVE_UNITS; ++u) {
      const u64 r0 = (first + wv * CS_WAVE_UNITS + u) * UNIT_RECORDS + 4 * u64(lane);
#pragma unroll
      for (int e = 0; e < 4; ++e) lw[u][e] = 0;
      if (VEC && r0 + 3 < m) {
        const uint4 a = *reinterpret_cast<const uint4*>(s_in + r0);
        const uint4 b = *reinterpret_cast<const uint4*>(d_in + r0);
        ls[u][0] = a.x; ls[u][1] = a.y; ls[u][2] = a.z; ls[u][3] = a.w;
        ld[u][0] = b.x; ld[u][1] = b.y; ld[u][2] = b.z; ld[u][3] = b.w;
        if (stored) {
          const uint4 c = *reinterpret_cast<const uint4*>(keys.v + r0);
          const uint4 e = *reinterpret_cast<const uint4*>(keys.v + r0 + 2);
          lv[u][0] = (u64(c.y) << 32) | c.x;
          lv[u][1] = (u64(c.w) << 32) | c.z;
          lv[u][2] = (u64(e.y) << 32) | e.x;
          lv[u][3] = (u64(e.w) << 32) | e.z;
        } else {
          const uint4 c = *reinterpret_cast<const uint4*>(keys.w + r0);
          lw[u][0] = c.x; lw[u][1] = c.y; lw[u][2] = c.z; lw[u][3] = c.w;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const u64 i = min(r0 + e, m - 1);
          ls[u][e] = s_in[i];
          ld[u][e] = d_in[i];
          if (stored) lv[u][e] = keys.v[i];
          else lw[u][e] = keys.w[i];
        }
      }
    }
  };
  u32 ss[CS_WAVE_UNITS][4] = {}, dd[CS_WAVE_UNITS][4] = {}, ww[CS_WAVE_UNITS][4] = {};
  u64 vv[CS_WAVE_UNITS][4] = {};
  if (PREFETCH && blockIdx.x < chunks) load_chunk(blockIdx.x, ss, dd, ww, vv);
  for (u64 chunk = blockIdx.x; chunk < chunks; chunk += gridDim.x) {  // block-uniform
    const u64 first = chunk * CS_CHUNK;
    const u64 end = min(first + CS_CHUNK, units);
    // the chunk's output runs: class 0 [a0, a0 + n0), class 1 [a1, a1 + n1)
    const u64 a0 = off0[first], a1 = off1[first];
    const u64 n0 = (end < units ? u64(off0[end]) : total0) - a0;
    const u64 n1 = (end < units ? u64(off1[end]) : total1) - a1;
    u32 ns[CS_WAVE_UNITS][4] = {}, nd[CS_WAVE_UNITS][4] = {}, nw[CS_WAVE_UNITS][4] = {};
    u64 nv[CS_WAVE_UNITS][4] = {};
    if (!PREFETCH) load_chunk(chunk, ss, dd, ww, vv);
    else if (chunk + gridDim.x < chunks) load_chunk(chunk + gridDim.x, ns, nd, nw, nv);
    // the last record of the wave's first unit (lane 63's), for the check
    const u32 last_s = __shfl(ss[0][3], WAVE - 1, WAVE), last_d = __shfl(dd[0][3], WAVE - 1, WAVE);
    u32 ntask = 0;  // wave-uniform
#pragma unroll
    for (int u = 0; u < CS_WAVE_UNITS; ++u) {
      const u64 unit = first + wv * CS_WAVE_UNITS + u;  // wave-uniform
      if (unit >= units) break;
      const u64 r0 = unit * UNIT_RECORDS + 4 * u64(lane);
      bool dup[4] = {false, false, false, false};  // a copy of the record precedes it
      if (CHECK) {
        // the record before lane l's first one is lane l - 1's last
        u32 ps = __shfl_up(ss[u][3], 1, WAVE), pd = __shfl_up(dd[u][3], 1, WAVE);
        if (lane == 0) {
          if (u > 0) {
            ps = last_s;
            pd = last_d;
          } else if (r0 > 0 && r0 < m) {
            ps = s_in[r0 - 1];
            pd = d_in[r0 - 1];
          }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const u64 i = r0 + e;
          if (i >= m) break;
          const u32 s = ss[u][e], d = dd[u][e];
          const bool has_prev = i > 0;
          if (e > 0) {
            ps = ss[u][e - 1];
            pd = dd[u][e - 1];
          }
          if (s >= chk.n || d >= chk.n || s == d) bad |= ERR_INVALID;
          if (has_prev && (ps > s || (ps == s && pd > d))) bad |= ERR_UNSORTED;
          dup[e] = has_prev && ps == s && pd == d;
          if (chk.seg_end && (!has_prev || ps != s)) {
            // record i opens s's range and closes the previous vertex's
            if (s < chk.n) chk.seg_start[s] = static_cast<u32>(i);
            if (has_prev && ps < chk.n) chk.seg_end[ps] = static_cast<u32>(i);
          }
          if (chk.seg_end && i + 1 == m && s < chk.n) chk.seg_end[s] = static_cast<u32>(m);
          const bool f = s < d;
          // CHECK runs with packed or lazy keys only: ww holds the weights
          const u64 h = sym_hash(f ? s : d, f ? d : s, ww[u][e]);
          fwd += f ? 1u : 0u;
          hf += f ? h : 0ull;
          hr += f ? 0ull : h;
        }
      }
      u32 cls[4];
      u64 b0[4], b1[4];
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const u32 w = stored ? static_cast<u32>(vv[u][e] >> wshift) : ww[u][e];
        cls[e] = (r0 + e >= m) ? 2u : (w <= q0 ? 0u : (w <= q1 ? 1u : 2u));
        b0[e] = __ballot(cls[e] == 0);
        b1[e] = __ballot(cls[e] == 1);
      }
      // lazy keys: an s < d record that no copy of itself precedes keys its
      // own position; the others get a pair_position task (ids checked), run
      // after the chunk is staged
      bool need[4] = {false, false, false, false};
      u64 bt[4] = {0, 0, 0, 0};
      if (LAZY) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const u32 s = ss[u][e], d = dd[u][e];
          const bool own = s < d && !dup[e];
          need[e] = cls[e] < 2 && !own && s != d && s < chk.n && d < chk.n;
          vv[u][e] = (u64(ww[u][e]) << 32) | (own ? r0 + e : u64{0xFFFFFFFFu});
          bt[e] = __ballot(need[e]);
        }
      }
      u32 tslot = ntask;
      // stage positions: lane l's records follow those of lanes < l
      u32 p0 = static_cast<u32>(off0[unit] - a0), p1 = static_cast<u32>(n0 + off1[unit] - a1);
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        p0 = mbcnt_acc(b0[e], p0);
        p1 = mbcnt_acc(b1[e], p1);
      }
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        if (cls[e] < 2) {
          const u32 p = cls[e] == 0 ? p0++ : p1++;
          if (p < u32(CS_CHUNK_RECORDS)) {
            st_s[p] = ss[u][e];
            st_d[p] = dd[u][e];
            st_v[p] = late_key ? u64(ww[u][e]) : vv[u][e];
            if (LAZY && need[e]) s_task[wv][tslot + mbcnt(bt[e])] = static_cast<u32>(p);
          } else {
            atomicOr(err, ERR_BOUNDS);
            if (LAZY && need[e]) s_task[wv][tslot + mbcnt(bt[e])] = NONE32;  // no search
          }
        }
        if (LAZY) tslot += static_cast<u32>(__popcll(bt[e]));
      }
      ntask = tslot;
    }
    __syncthreads();
    if (LAZY) {
      // the wave's pair_position searches, one per lane
      for (u32 t = lane; t < ntask; t += WAVE) {
        const u32 p = s_task[wv][t];
        if (p == NONE32) continue;
        const u32 s = st_s[p], d = st_d[p];
        const u32 wi = static_cast<u32>(st_v[p] >> 32);
        const u32 x = min(s, d), y = max(s, d);
        // unchecked ranges: keep the search inside the arrays
        const u64 lo = min<u64>(keys.twin_lo[x], m);
        const u64 hi = min<u64>(max<u64>(keys.twin_hi[x], lo), m);
        const u64 j = pair_position(d_in, y, lo, hi);
        if (j == NO_TWIN) {
          bad |= ERR_ASYMMETRIC;
          st_v[p] = VAL_NONE;
        } else {
          st_v[p] = (u64(wi) << 32) | j;
        }
      }
      __syncthreads();
    }
    // coalesced runs out of the stage
    const u64 nst = min<u64>(n0 + n1, CS_CHUNK_RECORDS);
    for (u32 j = threadIdx.x; j < nst; j += BLOCK) {
      const u64 p = j < n0 ? a0 + j : base1 + a1 + (j - n0);
      if (p < out_cap) {
        s_out[p] = st_s[j];
        d_out[p] = st_d[j];
        v_out[p] = late_key ? pack_key(static_cast<u32>(st_v[j]), st_s[j], st_d[j], keys.vbits)
                            : st_v[j];
      } else {
        atomicOr(err, ERR_BOUNDS);
      }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < CS_WAVE_UNITS && PREFETCH; ++u) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        ss[u][e] = ns[u][e];
        dd[u][e] = nd[u][e];
        ww[u][e] = nw[u][e];
        vv[u][e] = nv[u][e];
      }
    }
  }
  if (!CHECK) return;
#pragma unroll
  for (int o = WAVE / 2; o > 0; o >>= 1) {
    fwd += __shfl_xor(fwd, o, WAVE);
    hf += __shfl_xor(hf, o, WAVE);
    hr += __shfl_xor(hr, o, WAVE);
  }
  if (lane == 0) {
    s_red[wv][0] = fwd;
    s_red[wv][1] = hf;
    s_red[wv][2] = hr;
  }
  if (bad) atomicOr(err, bad);
  __syncthreads();
  if (threadIdx.x < 3) {
    u64 t = 0;
    for (int ww_ = 0; ww_ < WAVES; ++ww_) t += s_red[ww_][threadIdx.x];
    if (threadIdx.x == 0) {
      if (t) atomicAdd(chk.fwd_count, static_cast<u32>(t));
    } else if (t) {
      atomicAdd(reinterpret_cast<unsigned long long*>(chk.hsum + threadIdx.x - 1),
                static_cast<unsigned long long>(t));
    }
  }
}

void class_scatter(hipStream_t st, const u32* s_in, const u32* d_in, const KeySrc& keys, u64 m,
                   u32 wshift, u32 q0, u32 q1, const u32* off0, const u32* off1, u64 base1,
                   u64 written, u32* s_out, u32* d_out, u64* v_out, u32* err,
                   const InputCheck* check) {
  if (m == 0) return;
  if (check && keys.v) throw Error(KMST_EINVAL, "input check needs packed or lazy keys");
  const bool lazy = keys.twin_lo != nullptr;
  if (lazy && !check) throw Error(KMST_EINVAL, "lazy positional keys come with the input check");
  KTScope t(st, K_SPLIT_SCATTER, (keys.v ? 16.0 : 12.0) * m + 16.0 * double(written));
  const u64 chunk = check ? ScatterChunk<true>::UNITS : ScatterChunk<false>::UNITS;
  // more workgroups than fit at once: the dispatcher then balances the
  // chunks, whose time varies (GNM 2^24 / 2^28: 2.08 ms with 4096, 1.95 with
  // 8192, 1.93 with 16384, 1.89 with 32768 and 65536)
  const u32 grid = static_cast<u32>(std::min<u64>((record_units(m) + chunk - 1) / chunk, 65536));
  const bool vec = ((reinterpret_cast<uintptr_t>(s_in) | reinterpret_cast<uintptr_t>(d_in) |
                     reinterpret_cast<uintptr_t>(keys.v ? static_cast<const void*>(keys.v)
                                                        : static_cast<const void*>(keys.w))) &
                    15) == 0;
  const InputCheck chk = check ? *check : InputCheck{};
#define KMST_CLASS_SCATTER(V, C, L)                                                              \
  class_scatter_kernel<V, C, L><<<grid, BLOCK, 0, st>>>(s_in, d_in, keys, m, wshift, q0, q1,     \
                                                        off0, off1, base1, written, s_out, d_out, \
                                                        v_out, err, chk)
  if (vec && lazy) KMST_CLASS_SCATTER(true, true, true);
  else if (lazy) KMST_CLASS_SCATTER(false, true, true);
  else if (vec && check) KMST_CLASS_SCATTER(true, true, false);
  else if (vec) KMST_CLASS_SCATTER(true, false, false);
  else if (check) KMST_CLASS_SCATTER(false, true, false);
  else KMST_CLASS_SCATTER(false, false, false);
#undef KMST_CLASS_SCATTER
  KMST_HIP(hipGetLastError());
}

}  // namespace kmst
