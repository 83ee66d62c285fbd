// kmst_common.hpp -- shared types, error handling and wave/block primitives
// for the libkmst HIP kernels (gfx950, 64-lane wavefronts).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/kmst.h"

namespace kmst {

using u8 = std::uint8_t;
using u32 = std::uint32_t;
using u64 = std::uint64_t;
using i64 = std::int64_t;

constexpr int WAVE = 64;
constexpr int BLOCK = 256;            // 4 waves per workgroup
constexpr int WAVES = BLOCK / WAVE;
constexpr u32 NONE32 = 0xFFFFFFFFu;
constexpr u64 VAL_NONE = ~u64{0};     // "no edge" / dead record key

// Error bits a kernel can raise in the context's device error word.
enum ErrBit : u32 {
  ERR_UNSORTED = 1u,
  ERR_INVALID = 2u,
  ERR_ASYMMETRIC = 4u,
  ERR_SPIN = 8u,      // a bounded wait or bounded loop hit its limit
  ERR_BOUNDS = 16u,   // a computed output position fell outside its buffer
};

struct Error : std::runtime_error {
  kmst_status status;
  Error(kmst_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

#define KMST_HIP(expr)                                                              \
  do {                                                                              \
    hipError_t _e = (expr);                                                         \
    if (_e != hipSuccess)                                                           \
      throw ::kmst::Error(KMST_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

inline u32 ceil_div_u32(u64 a, u64 b) { return static_cast<u32>((a + b - 1) / b); }
inline u32 bits_for(u64 n) {  // bits needed to represent values in [0, n)
  u32 b = 0;
  while (b < 64 && (u64{1} << b) < n) ++b;
  return b;
}

// ------------------------------------------------------- per-kernel timing
// Every launch site brackets its kernel with a KTScope carrying the kernel's
// algorithmic bytes; when a KernelTimer is installed (a context with timing
// on) the scope records a HIP event pair on the launch stream.
enum KernelKind : int {
  K_FILL, K_SCAN_REDUCE, K_SCAN_DOWN, K_VALIDATE, K_OFFSETS, K_KEYS, K_ARGMIN, K_ARGMIN_FIXUP,
  K_HOOK, K_JUMP, K_LABEL, K_GPAR, K_RELABEL_COUNT, K_RELABEL_SCATTER, K_SORT_HIST,
  K_SORT_SCATTER, K_DEDUP, K_SAMPLE, K_SPLIT_COUNT, K_SPLIT_SCATTER, K_GATHER, K_GEN,
  K_PERSIST, K_OTHER, K_COUNT
};
const char* kernel_kind_name(int k);

struct KernelTotals {
  double launches = 0, ms = 0, bytes = 0;
};

struct KernelTimer {
  struct Mark {
    int kind;
    double bytes;
  };
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
  std::vector<Mark> marks;
  size_t used = 0;
  int depth = 0;  // open scopes: a scope opened inside another is not timed
  std::map<int, KernelTotals> totals;
  void begin(hipStream_t st, int kind, double bytes);
  void end(hipStream_t st) noexcept;
  void collect();  // after the stream is synchronised
  // adds to the algorithmic bytes of the most recent launch of `kind`
  // (outputs whose size is known only after the launch)
  void adjust_last(int kind, double delta) {
    for (size_t i = marks.size(); i-- > 0;)
      if (marks[i].kind == kind) {
        marks[i].bytes += delta;
        return;
      }
  }
  void reset() { totals.clear(); }
  ~KernelTimer();
};

extern thread_local KernelTimer* g_ktimer;

struct KTScope {
  hipStream_t st;
  bool on;
  KTScope(hipStream_t s, int kind, double bytes) : st(s), on(g_ktimer != nullptr) {
    if (on) g_ktimer->begin(s, kind, bytes);
  }
  ~KTScope() {
    if (on) g_ktimer->end(st);
  }
  KTScope(const KTScope&) = delete;
  KTScope& operator=(const KTScope&) = delete;
};

// SplitMix64 finaliser: the hash behind every deterministic generator.
__host__ __device__ __forceinline__ u64 mix64(u64 x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// ---------------------------------------------------------------- device side
__device__ __forceinline__ int lane_id() { return threadIdx.x & (WAVE - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / WAVE; }

// popcount(mask & lanes-below-me)
__device__ __forceinline__ u32 mbcnt(u64 mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<u32>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<u32>(mask), 0u));
}

template <typename T>
__device__ __forceinline__ T wave_inclusive_sum(T x) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < WAVE; d <<= 1) {
    T y = __shfl_up(x, d, WAVE);
    if (lane >= d) x += y;
  }
  return x;
}

// Exclusive prefix sum over the workgroup; `lds` holds >= WAVES entries.
template <typename T>
__device__ __forceinline__ T block_exclusive_sum(T x, T& total, T* lds) {
  const int lane = lane_id(), wid = wave_id();
  const T incl = wave_inclusive_sum(x);
  if (lane == WAVE - 1) lds[wid] = incl;
  __syncthreads();
  T off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < WAVES; ++w) {
    const T v = lds[w];
    off += (w < wid) ? v : T(0);
    tot += v;
  }
  __syncthreads();
  total = tot;
  return off + incl - x;
}

// ------------------------------------------------ decoupled look-back
// Tile status word: bits 63..62 flag (0 none, 1 aggregate, 2 inclusive
// prefix), bits 61..48 the call's epoch, bits 47..0 value.  A word of another
// epoch reads as "none", so the status array is cleared only when it is
// (re)allocated or the epochs wrap.  Tiles take dynamic ids in dispatch
// order, so a tile only ever waits for tiles that are already running, and
// every wait is bounded: on expiry the kernel raises ERR_SPIN and the tile
// writes nothing, which the host turns into an error.
constexpr u64 LB_AGG = u64{1} << 62;
constexpr u64 LB_INC = u64{2} << 62;
constexpr u32 LB_EPOCH_BITS = 14;
constexpr u64 LB_VAL = (u64{1} << 48) - 1;
constexpr u32 LB_SPIN_LIMIT = 1u << 22;

__device__ __forceinline__ u64 lb_load(const u64* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void lb_store(u64* p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 lb_word(u64 flag, u32 epoch, u64 value) {
  return flag | (u64(epoch) << 48) | (value & LB_VAL);
}

// Called by all 64 lanes of one wave; returns the exclusive prefix of `tile`
// (wave-uniform) and publishes the inclusive one.  `ok` turns false when the
// bounded wait expired; the caller must then write nothing.
__device__ __forceinline__ u64 lookback_exclusive(u64* status, u32 epoch, u32 tile, u64 aggregate,
                                                  u32* err, bool& ok) {
  ok = true;
  const int lane = threadIdx.x & (WAVE - 1);
  if (tile == 0) {
    if (lane == 0) lb_store(status, lb_word(LB_INC, epoch, aggregate));
    return 0;
  }
  if (lane == 0) lb_store(status + tile, lb_word(LB_AGG, epoch, aggregate));
  u64 excl = 0;
  i64 end = static_cast<i64>(tile);  // window covers tiles [end-64, end)
  u32 spins = 0;
  while (true) {
    const i64 idx = end - 1 - lane;
    const u64 w = (idx >= 0) ? lb_load(status + idx) : lb_word(LB_INC, epoch, 0);
    const bool current = ((w >> 48) & ((1u << LB_EPOCH_BITS) - 1)) == epoch;
    const u64 flag = current ? (w >> 62) : 0;
    const u64 inc_mask = __ballot(flag == 2);
    const u64 none_mask = __ballot(flag == 0);
    const int first_inc = inc_mask ? (__ffsll(static_cast<long long>(inc_mask)) - 1) : WAVE;
    const u64 upto = (first_inc >= WAVE - 1) ? ~u64{0} : ((u64{2} << first_inc) - 1);
    if (none_mask & upto) {
      if (++spins > LB_SPIN_LIMIT) {
        if (lane == 0) atomicOr(err, ERR_SPIN);
        ok = false;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    u64 v = (lane <= first_inc) ? (w & LB_VAL) : 0;
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, WAVE);
    excl += v;
    if (first_inc < WAVE) break;
    end -= WAVE;
  }
  if (lane == 0) lb_store(status + tile, lb_word(LB_INC, epoch, excl + aggregate));
  return excl;
}

// Where a record's strict key comes from: stored keys, or (packed-key mode on
// the read-only input) computed from the weight array as
// (w << 2b) | (min(s, d) << b) | max(s, d) with b = vbits.  With twin_lo set
// (positional mode on the read-only input, taken only by the kernels that
// read the input directly: the class scatter and the outside gather) the key
// is (w << 32) | the position of the first record of the pair (min, max),
// found by a search in min's record range [twin_lo[min], twin_hi[min]).
struct KeySrc {
  const u64* v = nullptr;
  const u32* w = nullptr;
  u32 vbits = 0;
  const u32* twin_lo = nullptr;
  const u32* twin_hi = nullptr;
};

__device__ __forceinline__ u64 pack_key(u32 w, u32 s, u32 d, u32 vbits) {
  const u64 lo = s < d ? s : d, hi = s < d ? d : s;
  return (u64(w) << (2 * vbits)) | (lo << vbits) | hi;
}

__device__ __forceinline__ u64 load_key(const KeySrc& k, u64 i, u32 s, u32 d) {
  return k.v ? k.v[i] : pack_key(k.w[i], s, d, k.vbits);
}

// Called by all lanes of a wave holding consecutive records: folds the keys
// of lanes with equal sources (k == VAL_NONE: nothing to offer) and lets the
// last lane of each run offer the minimum.  A lane only ever folds keys of
// its own source, so the value offered is always one of that source's keys,
// and the last lane of a run sees at least the whole run.
__device__ __forceinline__ void offer_min(u64* best, u32 s, u64 k) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < WAVE; o <<= 1) {
    const u32 so = __shfl_up(s, o, WAVE);
    const bool same = lane >= o && so == s;
    // runs are contiguous: without a same-source lane o back, no run is
    // longer than o, and the steps so far (1 + 2 + ... + o/2 lanes) cover it
    if (!__any(same)) break;
    const u64 ko = __shfl_up(k, o, WAVE);
    if (same && ko < k) k = ko;
  }
  const u32 sn = __shfl_down(s, 1, WAVE);
  if ((lane == WAVE - 1 || sn != s) && k != VAL_NONE) {
    unsigned long long* p = reinterpret_cast<unsigned long long*>(best + s);
    if (k < *p) atomicMin(p, static_cast<unsigned long long>(k));
  }
}

}  // namespace kmst
