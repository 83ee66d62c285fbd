// kmst_persist.hip -- every remaining Boruvka round of a small graph in one
// persistent, cooperatively launched kernel.
//
// The round loop of the solver sizes each launch from counts it reads back,
// so a small round costs a dozen dependent launches and a host round trip
// (about 120 us at a few thousand records).  Here one workgroup per CU stays
// resident for the whole loop and the phases of a round are separated by a
// grid barrier instead:
//
//   B  bind   every record whose key equals its source's minimum names the
//             other endpoint (the minima themselves come from phase R);
//   H  hook   a vertex with a minimum points at the other endpoint, or stays
//             a root when the choice is mutual and it is the smaller id (the
//             lower-endpoint rule of determine_mst_edges,
//             include/algorithms/gbbs_reimplementation.hpp:100-130); the
//             non-root of each pair appends the edge's key to the MST;
//   J  jump   pointer jumping to the root (update_parents,
//             include/algorithms/gbbs_reimplementation.hpp:162-169), clears
//             the minimum and updates the parent array of the filter
//             recursion;
//   R  relabel both endpoints, drop self-loops (relabel_edges /
//             filter_out_self_loops, gbbs_reimplementation.hpp:184-219) and,
//             in the same pass, offer every kept record's key to its new
//             source's minimum (compute_min_edges,
//             include/algorithms/gbbs_reimplementation.hpp:72-95).
//
// The minima are 64-bit atomic minima of the strict key, so the result does
// not depend on record order: ids keep their values across rounds (a root is
// named by its own id), no sort is needed, and the MSF is the unique one
// every other path computes.  A wave first folds the keys of lanes with the
// same source (runs of one source stay adjacent: the input is grouped by
// source and the compaction keeps the order inside each wave's chunk), so a
// vertex with many records receives one atomic per run, not one per record.
#include "kmst_kernels.hpp"

namespace kmst {

namespace {

constexpr int PB = 1024;  // threads per workgroup: 16 waves on one CU
constexpr int PW = PB / WAVE;
constexpr int PU = 4;     // rows of 64 records a wave loads before it uses any
constexpr u64 PCHUNK = u64(WAVE) * PU;
constexpr u64 BAR_TICKS = 1000000000ull;  // 10 s of the 100 MHz wall clock

// control words (u32) at PersistArgs::ctl
enum : int { PC_BAR = 0, PC_M = 1 /* three slots */, PC_MST = 4, PC_ROUNDS = 5, PC_WHY = 6,
             PC_WHY_ROUND = 7 };
// why the kernel stopped early (bits of ctl[PC_WHY])
enum : u32 { WHY_CYCLE = 1, WHY_ROUNDS = 2, WHY_TIMEOUT = 4, WHY_PEER = 8, WHY_BIND = 16,
             WHY_CAP = 32, WHY_ENTRY = 64 };

__device__ __forceinline__ void why(u32* ctl, u32 bit, u32 r) {
  atomicOr(ctl + PC_WHY, bit);
  atomicMax(ctl + PC_WHY_ROUND, r);
}

// All workgroups arrive, then all leave.  Arrivals go to one of
// ceil(G / BAR_GROUP) group counters (128 bytes apart) and the last arrival of
// a group adds one to the root counter, so no address takes more than
// BAR_GROUP arrivals (same-address atomics serialise).  Counters only grow:
// after the k-th barrier a group counter holds k * (its workgroups) and the
// root k * (groups).  Every wave first waits
// until its own global stores have completed (the workgroup barrier waits for
// LDS traffic only), then thread 0 releases the CU's writes with an
// agent-scope fence, arrives, waits, and acquires for the whole CU before the
// workgroup leaves.  The wait is bounded by the wall clock; on expiry, or
// when another workgroup has already raised ERR_SPIN, the barrier fails and
// the caller leaves the kernel, so every workgroup ends.
constexpr u32 BAR_GROUP = 16;
constexpr u32 BAR_GROUP_STRIDE = 32;  // u32 words between group counters
constexpr u32 BAR_GROUPS_MAX = 16;    // up to 256 workgroups
static_assert(BAR_GROUPS_MAX * BAR_GROUP_STRIDE == PERSIST_GROUP_WORDS, "group counter area");
__device__ __forceinline__ bool grid_sync(u32* bar, u32* groups, u32& k, u32* err, u32 r) {
  __shared__ u32 s_ok;
  if (gridDim.x == 1) {
    // one workgroup: its barrier orders its own global accesses
    __syncthreads();
    return true;
  }
  ++k;
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores are done
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 ok = 1;
    __threadfence();
    const u32 G = gridDim.x, g = blockIdx.x / BAR_GROUP;
    const u32 in_group = min(BAR_GROUP, G - g * BAR_GROUP);
    if (atomicAdd(groups + g * BAR_GROUP_STRIDE, 1u) + 1 == k * in_group) atomicAdd(bar, 1u);
    const u32 target = k * ((G + BAR_GROUP - 1) / BAR_GROUP);
    const u64 t0 = wall_clock64();
    while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      const bool peer =
          __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & ERR_SPIN;
      if (peer || wall_clock64() - t0 > BAR_TICKS) {
        atomicOr(err, ERR_SPIN);
        why(bar, peer ? WHY_PEER : WHY_TIMEOUT, r);
        ok = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
    __threadfence();
    s_ok = ok;
  }
  __syncthreads();
  return s_ok != 0;
}

// Workgroup-aggregated reservation of output slots: every wave passes its
// count (wave-uniform) and gets its first slot; one atomic per workgroup and
// call (per-wave atomics on one counter serialise).  Called by all threads;
// the two LDS sets alternate, so a wave that runs ahead into the next call
// never overwrites values another wave has yet to read.
struct Reserve {
  u32 cnt[2][PW];
  u32 base[2];
};
__device__ __forceinline__ u32 block_reserve(u32 wave_count, u32* counter, Reserve& s,
                                             u32& parity) {
  const int wid = threadIdx.x / WAVE;
  if (lane_id() == 0) s.cnt[parity][wid] = wave_count;
  __syncthreads();
  if (threadIdx.x == 0) {
    u32 t = 0;
    for (int w = 0; w < PW; ++w) {
      const u32 c = s.cnt[parity][w];
      s.cnt[parity][w] = t;
      t += c;
    }
    s.base[parity] = t ? atomicAdd(counter, t) : 0u;
  }
  __syncthreads();
  const u32 first = s.base[parity] + s.cnt[parity][wid];
  parity ^= 1u;
  return first;
}

__global__ __launch_bounds__(PB) void persistent_rounds_kernel(PersistArgs a) {
  __shared__ u32 s_live[PW], s_emit[PW];
  __shared__ Reserve s_res;
  u32 parity = 0;
  const int lane = lane_id(), wid = threadIdx.x / WAVE;
  const u64 gtid = u64(blockIdx.x) * PB + threadIdx.x, nthr = u64(gridDim.x) * PB;
  const u64 gwave = u64(blockIdx.x) * PW + wid, nwaves = u64(gridDim.x) * PW;
  const u64 n = a.n;
  u32* bar = a.ctl + PC_BAR;
  u32 target = 0;
  u32 r = 0;
  if (gtid == 0 && __hip_atomic_load(a.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    why(a.ctl, WHY_ENTRY, 0);

  for (u64 v = gtid; v < n; v += nthr) {
    a.best[v] = VAL_NONE;
    a.bdst[v] = NONE32;
  }
  if (!grid_sync(bar, a.groups, target, a.err, r)) return;
  // minima of the first round
  for (u64 base = gwave * PCHUNK; base < a.m0; base += nwaves * PCHUNK) {
    u32 ss[PU], dd[PU];
    u64 kk[PU];
#pragma unroll
    for (int u = 0; u < PU; ++u) {
      const u64 i = base + u * WAVE + lane;
      ss[u] = i < a.m0 ? a.s0[i] : NONE32;
      dd[u] = i < a.m0 ? a.d0[i] : NONE32;
    }
#pragma unroll
    for (int u = 0; u < PU; ++u) {
      const u64 i = base + u * WAVE + lane;
      kk[u] = (i < a.m0 && ss[u] != dd[u]) ? load_key(a.k0, i, ss[u], dd[u]) : VAL_NONE;
    }
#pragma unroll
    for (int u = 0; u < PU; ++u) offer_min(a.best, ss[u], kk[u]);
  }
  if (!grid_sync(bar, a.groups, target, a.err, r)) return;

  const u32* in_s = a.s0;
  const u32* in_d = a.d0;
  KeySrc keys = a.k0;
  u64 m = a.m0;
  while (true) {
    u64* lg = a.log + PERSIST_LOG_WORDS * r;
    const bool stamp = blockIdx.x == 0 && threadIdx.x == 0;
    if (stamp) {
      lg[0] = m;
      lg[2] = wall_clock64();
    }
    // B: the other endpoint of every vertex's minimum
    for (u64 i0 = gtid; i0 < m; i0 += PU * nthr) {
      u32 ss[PU], dd[PU];
      u64 kk[PU], bb[PU];
#pragma unroll
      for (int u = 0; u < PU; ++u) {
        const u64 i = i0 + u * nthr;
        ss[u] = i < m ? in_s[i] : 0u;
        dd[u] = i < m ? in_d[i] : 0u;
      }
#pragma unroll
      for (int u = 0; u < PU; ++u) {
        const u64 i = i0 + u * nthr;
        kk[u] = i < m ? load_key(keys, i, ss[u], dd[u]) : VAL_NONE;
        bb[u] = i < m ? a.best[ss[u]] : 0;
      }
#pragma unroll
      for (int u = 0; u < PU; ++u)
        if (kk[u] != VAL_NONE && ss[u] != dd[u] && kk[u] == bb[u]) a.bdst[ss[u]] = dd[u];
    }
    if (!grid_sync(bar, a.groups, target, a.err, r)) return;

    if (stamp) lg[3] = wall_clock64();
    // H: hook, MST emission (workgroup-uniform loop: one slot reservation
    // per workgroup and step)
    u32 live = 0, emit = 0;
    for (u64 bb = u64(blockIdx.x) * PB; bb < n; bb += nthr) {
      const u64 v = bb + threadIdx.x;
      bool em = false;
      u64 bv = VAL_NONE;
      if (v < n) {
        bv = a.best[v];
        if (bv == VAL_NONE) {
          a.parent[v] = static_cast<u32>(v);
        } else {
          const u32 o = a.bdst[v];
          if (o >= n) {
            // no record bound the minimum: never so for a consistent state;
            // the vertex stays a root and emits nothing, the host reports it
            atomicOr(a.err, ERR_BOUNDS);
            why(a.ctl, WHY_BIND, r);
            a.parent[v] = static_cast<u32>(v);
          } else {
            const bool root = v < o && a.best[o] == bv;  // no gather above the target
            a.parent[v] = root ? static_cast<u32>(v) : o;
            em = !root;
          }
          ++live;
        }
      }
      const u64 emask = __ballot(em);
      const u32 first =
          block_reserve(static_cast<u32>(__popcll(emask)), a.ctl + PC_MST, s_res, parity);
      if (em) {
        const u64 at = a.mst_base + first + mbcnt(emask);
        if (at < a.mst_cap)
          a.mst[at] = bv;
        else
          atomicOr(a.err, ERR_BOUNDS);
      }
      emit += em ? 1u : 0u;
    }
    {
      // per-round log: live vertices (low half), MST edges added (high half)
      u32 lw = live, ew = emit;
#pragma unroll
      for (int o = WAVE / 2; o > 0; o >>= 1) {
        lw += __shfl_xor(lw, o, WAVE);
        ew += __shfl_xor(ew, o, WAVE);
      }
      if (lane == 0) {
        s_live[wid] = lw;
        s_emit[wid] = ew;
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        u64 lt = 0, et = 0;
        for (int w = 0; w < PW; ++w) {
          lt += s_live[w];
          et += s_emit[w];
        }
        if (lt | et)
          atomicAdd(reinterpret_cast<unsigned long long*>(lg + 1),
                    static_cast<unsigned long long>((et << 32) | lt));
      }
    }
    if (!grid_sync(bar, a.groups, target, a.err, r)) return;
    if (stamp) lg[4] = wall_clock64();

    // J: pointer jumping; the minima are cleared for the next round
    for (u64 v = gtid; v < n; v += nthr) {
      if (a.best[v] == VAL_NONE) continue;
      // a path to the root has fewer than n steps; a longer walk is a cycle
      u32 x = a.parent[v];
      u32 y = a.parent[x];
      u64 guard = 0;
      while (x != y) {
        a.parent[v] = y;
        x = y;
        y = a.parent[x];
        if (++guard > n) {
          atomicOr(a.err, ERR_SPIN);
          why(a.ctl, WHY_CYCLE, r);
          break;
        }
      }
      a.parent[v] = x;
      a.best[v] = VAL_NONE;
      a.bdst[v] = NONE32;
      if (a.gpar) {
        const u32 name = a.rep ? a.rep[v] : static_cast<u32>(v);
        a.gpar[name] = a.rep ? a.rep[x] : x;
      }
    }
    // the count slot of the next round is free: every workgroup read it two
    // barriers ago at the latest
    if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl[PC_M + (r + 1) % 3] = 0;
    if (!grid_sync(bar, a.groups, target, a.err, r)) return;
    if (stamp) lg[5] = wall_clock64();

    // R: relabel, drop self-loops, compact (each wave's chunk of PU rows
    // keeps its order), next minima
    const int ob = r & 1;
    u32* os = a.s[ob];
    u32* od = a.d[ob];
    u64* ov = a.v[ob];
    u32* cnt = a.ctl + PC_M + r % 3;
    for (u64 bb = u64(blockIdx.x) * (PB * PU); bb < m; bb += nthr * PU) {
      // workgroup-uniform steps of PW chunks; this wave takes chunk wid
      const u64 base = bb + u64(wid) * PCHUNK;
      u32 ss[PU], dd[PU], ns[PU], nd[PU];
      u64 kk[PU], bal[PU];
#pragma unroll
      for (int u = 0; u < PU; ++u) {
        const u64 i = base + u * WAVE + lane;
        ss[u] = i < m ? in_s[i] : 0u;
        dd[u] = i < m ? in_d[i] : 0u;
      }
#pragma unroll
      for (int u = 0; u < PU; ++u) {
        const u64 i = base + u * WAVE + lane;
        ns[u] = i < m ? a.parent[ss[u]] : NONE32;
        nd[u] = i < m ? a.parent[dd[u]] : NONE32;
      }
      u32 total = 0;
#pragma unroll
      for (int u = 0; u < PU; ++u) {
        // records a dedup pass marked dead (key VAL_NONE) are dropped too
        const u64 i = base + u * WAVE + lane;
        kk[u] = (i < m && ns[u] != nd[u]) ? load_key(keys, i, ss[u], dd[u]) : VAL_NONE;
        const bool keep = kk[u] != VAL_NONE;
        bal[u] = __ballot(keep);
        total += static_cast<u32>(__popcll(bal[u]));
        if (!keep) ns[u] = NONE32;
      }
      u32 at0 = block_reserve(total, cnt, s_res, parity);
#pragma unroll
      for (int u = 0; u < PU; ++u) {
        if (kk[u] != VAL_NONE) {
          const u64 at = u64(at0) + mbcnt(bal[u]);
          if (at < a.cap) {
            os[at] = ns[u];
            od[at] = nd[u];
            ov[at] = kk[u];
          } else {
            atomicOr(a.err, ERR_BOUNDS);
            why(a.ctl, WHY_CAP, r);
          }
        }
        at0 += static_cast<u32>(__popcll(bal[u]));
      }
#pragma unroll
      for (int u = 0; u < PU; ++u) offer_min(a.best, ns[u], kk[u]);
    }
    if (!grid_sync(bar, a.groups, target, a.err, r)) return;
    if (stamp) lg[6] = wall_clock64();

    m = __hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    in_s = os;
    in_d = od;
    keys = KeySrc{ov, nullptr, 0};
    ++r;
    if (m == 0) break;
    // more kept records than the output holds (never so for a consistent
    // state: a round keeps at most the records it read): ERR_BOUNDS is
    // raised above, and every workgroup reads the same count after the
    // barrier, so all of them leave here before anything reads past cap
    if (m > a.cap) break;
    if (r >= PERSIST_MAX_ROUNDS) {
      // Boruvka at least halves the vertices with records every round
      if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicOr(a.err, ERR_SPIN);
        why(a.ctl, WHY_ROUNDS, r);
      }
      break;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) a.ctl[PC_ROUNDS] = r;
}

}  // namespace

u32 persistent_blocks_max(int device) {
  static int cached_dev = -1;
  static u32 cached = 0;
  if (cached_dev == device) return cached;
  int coop = 0, cus = 0, per_cu = 0;
  KMST_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, device));
  KMST_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  KMST_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(persistent_rounds_kernel), PB, 0));
  // one workgroup per CU: enough threads for the sizes run here, and the
  // fewest arrivals per barrier
  cached = (coop && per_cu > 0) ? static_cast<u32>(cus) : 0u;
  cached_dev = device;
  return cached;
}

u32 persistent_threads() { return PB; }

double persistent_ticks_per_us(int device) {
  static int cached_dev = -1;
  static double cached = 0;
  if (cached_dev == device) return cached;
  int khz = 0;
  KMST_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device));
  cached = khz > 0 ? khz / 1000.0 : 100.0;
  cached_dev = device;
  return cached;
}

void persistent_rounds(hipStream_t st, const PersistArgs& a, u32 blocks) {
  if (blocks == 0) throw Error(KMST_EDEVICE, "persistent rounds: no cooperative launch");
  if (reinterpret_cast<u64*>(a.ctl) + PERSIST_CTL_WORDS / 2 != a.log)
    throw Error(KMST_EINVAL, "persistent rounds: ctl must directly precede log");
  if (a.log + PERSIST_LOG_TOTAL != reinterpret_cast<u64*>(a.groups))
    throw Error(KMST_EINVAL, "persistent rounds: the barrier groups must directly follow log");
  if (blocks > BAR_GROUP * BAR_GROUPS_MAX)
    throw Error(KMST_EINVAL, "persistent rounds: too many workgroups");
  fill_u64(st, reinterpret_cast<u64*>(a.ctl),
           PERSIST_CTL_WORDS / 2 + PERSIST_LOG_TOTAL + PERSIST_GROUP_WORDS / 2, 0);
  KTScope t(st, K_PERSIST, 0.0);
  if (blocks == 1) {
    // one workgroup has no grid barrier, so it needs no co-residency
    // guarantee: an ordinary launch, which the host does not wait on
    persistent_rounds_kernel<<<1, PB, 0, st>>>(a);
    KMST_HIP(hipGetLastError());
    return;
  }
  PersistArgs arg = a;
  void* args[] = {&arg};
  KMST_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(persistent_rounds_kernel),
                                      dim3(blocks), dim3(PB), args, 0, st));
}

}  // namespace kmst
