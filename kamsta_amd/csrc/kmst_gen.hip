// kmst_gen.hip -- deterministic input generators (the role KaGen plays for
// the reference driver, benchmarks/mst_benchmarks.hpp:116-199).
//
// GNM(n, m, seed, max_weight): candidate c (c = 0, 1, 2, ...) is the pair
// (a, b) = (hi32(r) * n >> 32, lo32(r) * n >> 32) of r = mix(mix(seed) + c);
// self-loops are rejected and the graph is the first m distinct {a, b} in
// counter order, i.e. a uniform m-subset of the pairs.  The weight of {u, v}
// (u < v) is 1 + mix(mix(seed ^ W) ^ (u << 32 | v)) mod max_weight, equal in
// both directions.  Output: 2m records, symmetric, sorted by (src, dst).
// The device version (candidate sort + first-occurrence marks) and the host
// version produce identical arrays.
#include <algorithm>
#include <numeric>
#include <vector>

#include "kmst_kernels.hpp"
#include "kmst_solver.hpp"

namespace kmst {

constexpr u64 WEIGHT_SALT = 0x5745494748543A31ull;

__host__ __device__ __forceinline__ void gnm_candidate(u64 seed_mix, u64 n, u64 c, u32& a, u32& b) {
  const u64 r = mix64(seed_mix + c);
  const u32 x = static_cast<u32>(((r >> 32) * n) >> 32);
  const u32 y = static_cast<u32>(((r & 0xFFFFFFFFull) * n) >> 32);
  a = x < y ? x : y;
  b = x < y ? y : x;
}

__host__ __device__ __forceinline__ u32 gnm_weight(u64 wseed, u32 u, u32 v, u32 max_w) {
  return 1u + static_cast<u32>(mix64(wseed ^ ((u64(u) << 32) | v)) % max_w);
}

__global__ __launch_bounds__(BLOCK) void gnm_candidates_kernel(u64 seed_mix, u64 n, u64 k,
                                                               u32* __restrict__ a,
                                                               u32* __restrict__ b,
                                                               u64* __restrict__ ctr) {
  for (u64 c = u64(blockIdx.x) * BLOCK + threadIdx.x; c < k; c += u64(gridDim.x) * BLOCK) {
    u32 x, y;
    gnm_candidate(seed_mix, n, c, x, y);
    a[c] = x;
    b[c] = y;
    ctr[c] = c;
  }
}

// flag_by_ctr[ctr] = 1 for the first occurrence (lowest counter) of each
// distinct non-loop pair in the (pair, counter)-sorted candidates.
__global__ __launch_bounds__(BLOCK) void gnm_first_kernel(const u32* __restrict__ a,
                                                          const u32* __restrict__ b,
                                                          const u64* __restrict__ ctr, u64 k,
                                                          u32* __restrict__ flag_by_ctr) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < k; i += u64(gridDim.x) * BLOCK) {
    const bool first = a[i] != b[i] && (i == 0 || a[i - 1] != a[i] || b[i - 1] != b[i]);
    flag_by_ctr[ctr[i]] = first ? 1u : 0u;
  }
}

__global__ __launch_bounds__(BLOCK) void gnm_emit_kernel(u64 seed_mix, u64 wseed, u64 n, u64 k,
                                                         u64 m, u32 max_w,
                                                         const u32* __restrict__ flag,
                                                         const u32* __restrict__ rank,
                                                         u32* __restrict__ s, u32* __restrict__ d,
                                                         u64* __restrict__ w) {
  for (u64 c = u64(blockIdx.x) * BLOCK + threadIdx.x; c < k; c += u64(gridDim.x) * BLOCK) {
    if (!flag[c]) continue;
    const u64 j = rank[c];
    if (j >= m) continue;
    u32 x, y;
    gnm_candidate(seed_mix, n, c, x, y);
    const u32 wt = gnm_weight(wseed, x, y, max_w);
    s[2 * j] = x;
    d[2 * j] = y;
    w[2 * j] = wt;
    s[2 * j + 1] = y;
    d[2 * j + 1] = x;
    w[2 * j + 1] = wt;
  }
}

__global__ __launch_bounds__(BLOCK) void narrow_kernel(const u64* __restrict__ in, u64 n,
                                                       u32* __restrict__ out) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += u64(gridDim.x) * BLOCK)
    out[i] = static_cast<u32>(in[i]);
}

static void check_gnm_args(u64 n, u64 m, u32 max_w) {
  if (n < 2 && m > 0) throw Error(KMST_EINVAL, "GNM needs n >= 2");
  if (n >= (u64{1} << 32) - 1) throw Error(KMST_EINVAL, "GNM n >= 2^32-1");
  if (max_w == 0) throw Error(KMST_EINVAL, "max_weight == 0");
  if (m > n * (n - 1) / 2) throw Error(KMST_EINVAL, "m exceeds n(n-1)/2");
  if (2 * m >= (u64{1} << 32) - 1) throw Error(KMST_EINVAL, "2m >= 2^32-1");
}

static u64 initial_candidates(u64 n, u64 m) {
  // expected duplicates ~ m^2 / n^2 (pairs drawn from n^2/2), plus loops ~ m/n
  const double nn = double(n);
  const double dup = double(m) * double(m) / (nn * nn) + double(m) / nn;
  return m + static_cast<u64>(2.0 * dup) + m / 64 + 1024;
}

void gen_gnm_device(kmst_context& C, u64 n, u64 m, u32 max_w, u64 seed, u32* out_s, u32* out_d,
                    u32* out_w) {
  check_gnm_args(n, m, max_w);
  if (m == 0) return;
  hipStream_t st = C.stream;
  const u64 seed_mix = mix64(seed), wseed = mix64(seed ^ WEIGHT_SALT);
  u64 k = initial_candidates(n, m);
  const u32 pair_bits = bits_for(n);
  u32* ctr = static_cast<u32*>(C.counters.get(256));
  for (int attempt = 0; attempt < 40; ++attempt) {
    // Each of the six arrays serves as candidate storage (k entries), as the
    // emit target and as the alt buffer of the final record sort (2m
    // entries), depending on the parity of the pass counts; size all of
    // them for the larger role.
    const u64 cap = std::max(k, 2 * m);
    u32* a = static_cast<u32*>(C.gen_a.get(cap * 4));
    u32* b = static_cast<u32*>(C.gen_b.get(cap * 4));
    u64* cv = static_cast<u64*>(C.gen_c.get(cap * 8));
    u32* a2 = static_cast<u32*>(C.xs.get(cap * 4));
    u32* b2 = static_cast<u32*>(C.xd.get(cap * 4));
    u64* c2 = static_cast<u64*>(C.xv.get(cap * 8));
    u32* flag = static_cast<u32*>(C.gen_d.get(2 * k * 4));
    u32* rank = flag + k;
    const u64 sort_m = std::max(k, 2 * m);
    C.sort_work.bind(st, static_cast<u32*>(C.hist.get(sort_work_words(sort_m) * 4)), sort_m);
    u32* scan_scratch = static_cast<u32*>(C.scan_scratch.get(scan_scratch_elems(k) * 4 + 64));
    u32* err = ctr + 8;  // ctr[0] receives scan totals
    fill_u32(st, ctr, 64, 0);
    const u32 grid = static_cast<u32>(std::min<u64>((k + BLOCK - 1) / BLOCK, 65536));
    {
      KTScope t(st, K_GEN, 16.0 * k);
      gnm_candidates_kernel<<<grid, BLOCK, 0, st>>>(seed_mix, n, k, a, b, cv);
      KMST_HIP(hipGetLastError());
    }
    u32 *sa = a, *sb = b;
    u64* sc = cv;
    if (radix_sort(st, a, b, cv, a2, b2, c2, k, pair_bits, 2 * pair_bits, C.sort_work, err)) {
      sa = a2;
      sb = b2;
      sc = c2;
    }
    {
      KTScope t(st, K_GEN, 20.0 * k);
      gnm_first_kernel<<<grid, BLOCK, 0, st>>>(sa, sb, sc, k, flag);
      KMST_HIP(hipGetLastError());
    }
    scan_exclusive_u32(st, flag, rank, k, scan_scratch, ctr);
    u32 unique = 0;
    KMST_HIP(hipMemcpyAsync(&unique, ctr, 4, hipMemcpyDeviceToHost, st));
    KMST_HIP(hipStreamSynchronize(st));
    if (unique < m) {
      k = k * 2;
      continue;
    }
    // emit 2m unsorted records into (a2|a, b2|b, c2|cv) then sort by (src, dst)
    u32* rs = (sa == a) ? a2 : a;
    u32* rd = (sa == a) ? b2 : b;
    u64* rw = (sa == a) ? c2 : cv;
    u32* ts = (sa == a) ? a : a2;
    u32* td = (sa == a) ? b : b2;
    u64* tw = (sa == a) ? cv : c2;
    {
      KTScope t(st, K_GEN, 8.0 * k + 32.0 * m);
      gnm_emit_kernel<<<grid, BLOCK, 0, st>>>(seed_mix, wseed, n, k, m, max_w, flag, rank, rs, rd,
                                              rw);
      KMST_HIP(hipGetLastError());
    }
    const u64 N = 2 * m;
    if (radix_sort(st, rs, rd, rw, ts, td, tw, N, pair_bits, 2 * pair_bits, C.sort_work, err)) {
      rs = ts;
      rd = td;
      rw = tw;
    }
    KMST_HIP(hipMemcpyAsync(out_s, rs, N * 4, hipMemcpyDeviceToDevice, st));
    KMST_HIP(hipMemcpyAsync(out_d, rd, N * 4, hipMemcpyDeviceToDevice, st));
    {
      KTScope t(st, K_GEN, 12.0 * N);
      narrow_kernel<<<static_cast<u32>(std::min<u64>((N + BLOCK - 1) / BLOCK, 65536)), BLOCK, 0,
                      st>>>(rw, N, out_w);
      KMST_HIP(hipGetLastError());
    }
    u32 err_host = 0;
    KMST_HIP(hipMemcpyAsync(&err_host, err, 4, hipMemcpyDeviceToHost, st));
    KMST_HIP(hipStreamSynchronize(st));
    if (err_host) throw Error(KMST_EDEVICE, "GNM: sort wait expired");
    return;
  }
  throw Error(KMST_EINVAL, "GNM: could not draw m distinct pairs");
}

void gen_gnm_host(u64 n, u64 m, u32 max_w, u64 seed, u32* out_s, u32* out_d, u32* out_w) {
  check_gnm_args(n, m, max_w);
  if (m == 0) return;
  const u64 seed_mix = mix64(seed), wseed = mix64(seed ^ WEIGHT_SALT);
  u64 k = initial_candidates(n, m);
  for (int attempt = 0; attempt < 40; ++attempt) {
    std::vector<u64> key(k);
    for (u64 c = 0; c < k; ++c) {
      u32 x, y;
      gnm_candidate(seed_mix, n, c, x, y);
      key[c] = (u64(x) << 32) | y;
    }
    std::vector<u64> order(k);
    std::iota(order.begin(), order.end(), 0ull);
    std::stable_sort(order.begin(), order.end(), [&](u64 p, u64 q) { return key[p] < key[q]; });
    std::vector<unsigned char> first(k, 0);
    u64 unique = 0;
    for (u64 i = 0; i < k; ++i) {
      const u64 c = order[i];
      const bool loop = (key[c] >> 32) == (key[c] & 0xFFFFFFFFull);
      if (!loop && (i == 0 || key[order[i - 1]] != key[c])) {
        first[c] = 1;
        ++unique;
      }
    }
    if (unique < m) {
      k *= 2;
      continue;
    }
    std::vector<u64> rec;  // (src << 32 | dst), weight looked up afterwards
    rec.reserve(2 * m);
    u64 taken = 0;
    for (u64 c = 0; c < k && taken < m; ++c) {
      if (!first[c]) continue;
      rec.push_back(key[c]);
      rec.push_back(((key[c] & 0xFFFFFFFFull) << 32) | (key[c] >> 32));
      ++taken;
    }
    std::sort(rec.begin(), rec.end());
    for (u64 i = 0; i < rec.size(); ++i) {
      const u32 s = static_cast<u32>(rec[i] >> 32), d = static_cast<u32>(rec[i]);
      out_s[i] = s;
      out_d[i] = d;
      out_w[i] = gnm_weight(wseed, s < d ? s : d, s < d ? d : s, max_w);
    }
    return;
  }
  throw Error(KMST_EINVAL, "GNM: could not draw m distinct pairs");
}

}  // namespace kmst
