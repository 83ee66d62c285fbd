// kmst_gen_geo.hip -- deterministic geometric input generators, on the host
// and on the device (the role KaGen's RGG_2D and RHG play for the reference
// driver, benchmarks/mst_benchmarks.hpp:116-199).  Output: symmetric records
// sorted by (src, dst), weights 1 + H(seed, min, max) mod max_weight, so both
// directions agree.  The host and device versions produce identical arrays.
//
// RGG2D(n, avg_degree): points uniform in the unit square as 32-bit fixed
// point coordinates; {u, v} is an edge iff the integer squared distance is at
// most r^2 with r = sqrt(avg_degree / (pi n)) (in fixed point).  The test is
// exact integer arithmetic, so the graph does not depend on floating point.
// Vertex ids are assigned cell by cell (row-major over a grid of cells of
// side >= r, points of a cell in generation order), the way KaGen numbers the
// vertices of its geometric generators, so that neighbours have nearby ids.
//
// RHG(n, avg_degree, gamma): threshold random hyperbolic graph.  alpha =
// (gamma - 1) / 2, disk radius R = 2 ln((2/pi) xi^2 n / avg_degree) with
// xi = alpha / (alpha - 1/2); radii from the density alpha sinh(alpha r) /
// (cosh(alpha R) - 1) (inverse CDF: r = acosh(1 + u (cosh(alpha R) - 1)) /
// alpha), angles uniform.  {u, v} is an edge iff their hyperbolic distance is
// at most R, i.e. (law of cosines in haversine form, well conditioned for the
// tiny angles of boundary pairs)
//   2 sin^2(dtheta / 2) sinh r_u sinh r_v <= cosh R - cosh(r_u - r_v),
// evaluated with the lower id first, so the test is the same from both sides.
// Vertex ids follow the (angle, radius) order (neighbours have nearby ids).
// The per-vertex values and the test use the functions of namespace det
// below, built from IEEE add, multiply, divide and square root only, with
// floating-point contraction off: the host and the device compute every
// value bit for bit alike, so they decide every pair alike.  Candidates come
// from radial bands sorted by angle; a band's angular window is taken at its
// smallest member radius, where it is widest, with a slack on the right-hand
// side and a relative margin of 1e-9 on the angle, so that it holds every
// pair the test accepts.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <exception>
#include <mutex>
#include <thread>
#include <vector>

#include "kmst_kernels.hpp"
#include "kmst_solver.hpp"

#pragma clang fp contract(off)

namespace kmst {

// =================================================================== det
// Elementary functions from the mathematics, deterministic across host and
// device.  exp: Cody-Waite reduction by ln 2 (W. J. Cody and W. Waite,
// Software Manual for the Elementary Functions, 1980) and the Taylor series
// of e^r on |r| <= ln2 / 2; log: the exponent split off, then
// ln m = 2 atanh((m - 1) / (m + 1)) by its series on m in [sqrt(1/2),
// sqrt(2)); sin, sinh: Taylor series on the reduced argument.  Series
// coefficients are the nearest doubles to 1/j!.
namespace det {

constexpr double PI = 3.141592653589793;           // nearest double to pi
constexpr double HALF_PI = 1.5707963267948966;
constexpr double TWO_PI = 6.283185307179586;
constexpr double LN2_HI = 0.6931467056274414;      // ln 2, low 32 mantissa bits clear
constexpr double LN2_LO = 4.7493250390316726e-07;  // ln 2 - LN2_HI
constexpr double INV_LN2 = 1.4426950408889634;
constexpr double SQRT2 = 1.4142135623730951;

__host__ __device__ inline double pow2(int k) {  // 2^k, -1022 <= k <= 1023
  return __builtin_bit_cast(double, u64(k + 1023) << 52);
}

__host__ __device__ inline double exp(double x) {
  if (x > 709.0) return __builtin_bit_cast(double, u64{0x7FF0000000000000ull});
  if (x < -708.0) return 0.0;
  const double kf = floor(x * INV_LN2 + 0.5);
  const int k = static_cast<int>(kf);
  const double r = (x - kf * LN2_HI) - kf * LN2_LO;
  // e^r = sum r^j / j!, j <= 14 (|r| <= 0.35: the next term is below 2^-60)
  double p = 1.1470745597729725e-11;
  p = p * r + 1.6059043836821613e-10;
  p = p * r + 2.08767569878681e-09;
  p = p * r + 2.505210838544172e-08;
  p = p * r + 2.755731922398589e-07;
  p = p * r + 2.7557319223985893e-06;
  p = p * r + 2.48015873015873e-05;
  p = p * r + 0.0001984126984126984;
  p = p * r + 0.001388888888888889;
  p = p * r + 0.008333333333333333;
  p = p * r + 0.041666666666666664;
  p = p * r + 0.16666666666666666;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return p * pow2(k);
}

__host__ __device__ inline double log(double x) {  // x normal and positive
  const u64 b = __builtin_bit_cast(u64, x);
  int e = static_cast<int>((b >> 52) & 0x7FF) - 1023;
  double m = __builtin_bit_cast(double, (b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);
  if (m > SQRT2) {
    m = m * 0.5;
    e += 1;
  }
  const double s = (m - 1.0) / (m + 1.0);
  const double z = s * s;
  // atanh(s) / s = sum z^j / (2j + 1), j <= 11 (z <= 0.0295)
  double q = 1.0 / 23.0;
  q = q * z + 1.0 / 21.0;
  q = q * z + 1.0 / 19.0;
  q = q * z + 1.0 / 17.0;
  q = q * z + 1.0 / 15.0;
  q = q * z + 1.0 / 13.0;
  q = q * z + 1.0 / 11.0;
  q = q * z + 1.0 / 9.0;
  q = q * z + 1.0 / 7.0;
  q = q * z + 1.0 / 5.0;
  q = q * z + 1.0 / 3.0;
  q = q * z + 1.0;
  const double ef = static_cast<double>(e);
  return ef * LN2_HI + (ef * LN2_LO + 2.0 * s * q);
}

__host__ __device__ inline double sin(double x) {  // |x| <= pi
  const double sg = x < 0.0 ? -1.0 : 1.0;
  double a = x < 0.0 ? -x : x;
  if (a > HALF_PI) a = PI - a;
  const double z = a * a;
  // sin(a) / a = sum (-z)^j / (2j + 1)!, j <= 12 (a <= pi/2)
  double q = 6.446950284384474e-26;
  q = q * z - 3.868170170630684e-23;
  q = q * z + 1.9572941063391263e-20;
  q = q * z - 8.22063524662433e-18;
  q = q * z + 2.8114572543455206e-15;
  q = q * z - 7.647163731819816e-13;
  q = q * z + 1.6059043836821613e-10;
  q = q * z - 2.505210838544172e-08;
  q = q * z + 2.7557319223985893e-06;
  q = q * z - 0.0001984126984126984;
  q = q * z + 0.008333333333333333;
  q = q * z - 0.16666666666666666;
  q = q * z + 1.0;
  return sg * (a * q);
}

__host__ __device__ inline double sinh(double x) {  // x >= 0
  if (x < 1.0) {
    const double z = x * x;
    // sinh(x) / x = sum z^j / (2j + 1)!, j <= 11
    double q = 3.868170170630684e-23;
    q = q * z + 1.9572941063391263e-20;
    q = q * z + 8.22063524662433e-18;
    q = q * z + 2.8114572543455206e-15;
    q = q * z + 7.647163731819816e-13;
    q = q * z + 1.6059043836821613e-10;
    q = q * z + 2.505210838544172e-08;
    q = q * z + 2.7557319223985893e-06;
    q = q * z + 0.0001984126984126984;
    q = q * z + 0.008333333333333333;
    q = q * z + 0.16666666666666666;
    q = q * z + 1.0;
    return x * q;
  }
  return 0.5 * (exp(x) - exp(-x));
}

__host__ __device__ inline double cosh(double x) { return 0.5 * (exp(x) + exp(-x)); }

// acosh(1 + t), t >= 0: ln(1 + t + sqrt(t (2 + t)))
__host__ __device__ inline double acosh1p(double t) {
  return log(1.0 + t + sqrt(t * (2.0 + t)));
}

}  // namespace det

namespace {

constexpr u64 GEO_WEIGHT_SALT = 0x5745494748543A32ull;

__host__ __device__ inline double unit(u64 h) {  // [0, 1), exact
  return double(h >> 11) * (1.0 / 9007199254740992.0);
}

__host__ __device__ inline u32 geo_weight(u64 wseed, u32 s, u32 d, u32 max_w) {
  const u32 a = s < d ? s : d, b = s < d ? d : s;
  return 1u + static_cast<u32>(mix64(wseed ^ ((u64(a) << 32) | b)) % max_w);
}

// host threads for the generators: OMP_NUM_THREADS when set (the CPU share
// of the process), else the hardware's
unsigned host_threads() {
  const char* e = std::getenv("OMP_NUM_THREADS");
  long t = e ? std::strtol(e, nullptr, 10) : 0;
  if (t <= 0) t = static_cast<long>(std::thread::hardware_concurrency());
  return static_cast<unsigned>(std::min<long>(std::max<long>(t, 1), 64));
}

// f(lo, hi, t) over T contiguous chunks of [0, n), chunk t by thread t
// (one thread below min_n items)
template <typename F>
void parallel_chunks(u64 n, unsigned T, F&& f, u64 min_n = 4096) {
  if (T <= 1 || n < min_n) {
    f(u64{0}, n, 0u);
    return;
  }
  std::vector<std::thread> th;
  std::exception_ptr failure;
  std::mutex mu;
  for (unsigned t = 0; t < T; ++t) {
    th.emplace_back([&, t] {
      try {
        f(n * t / T, n * (t + 1) / T, t);
      } catch (...) {
        std::lock_guard<std::mutex> lk(mu);
        failure = std::current_exception();
      }
    });
  }
  for (auto& x : th) x.join();
  if (failure) std::rethrow_exception(failure);
}

// Per-thread pair lists (u << 32 | v), concatenated in thread order.
kmst_status emit_pairs(const std::vector<std::vector<u64>>& parts, u32 max_w, u64 seed, u32* src,
                       u32* dst, u32* w, u64 cap, u64* m_out) {
  u64 total = 0;
  std::vector<u64> start(parts.size() + 1, 0);
  for (size_t t = 0; t < parts.size(); ++t) start[t + 1] = start[t] + parts[t].size();
  total = start.back();
  *m_out = total;
  if (!src || cap < total) return KMST_EINVAL;
  const u64 wseed = mix64(seed ^ GEO_WEIGHT_SALT);
  parallel_chunks(
      parts.size(), static_cast<unsigned>(parts.size()),
      [&](u64 lo, u64 hi, unsigned) {
        for (u64 t = lo; t < hi; ++t)
          for (u64 i = 0; i < parts[t].size(); ++i) {
            const u64 o = start[t] + i;
            const u32 s = static_cast<u32>(parts[t][i] >> 32), d = static_cast<u32>(parts[t][i]);
            src[o] = s;
            dst[o] = d;
            w[o] = geo_weight(wseed, s, d, max_w);
          }
      },
      2);
  return KMST_OK;
}

// =================================================================== RGG2D
struct RggParams {
  u64 n, cells, r2_hi, r2_lo, sm;
};

RggParams rgg_params(u64 n, double avg_degree, u64 seed) {
  const double r = std::sqrt(avg_degree / (det::PI * double(n)));
  const u64 rr = static_cast<u64>(r * 4294967296.0);  // radius in fixed point
  RggParams p;
  p.n = n;
  const unsigned __int128 r2 = static_cast<unsigned __int128>(rr) * rr;  // (hi, lo)
  p.r2_hi = static_cast<u64>(r2 >> 64);
  p.r2_lo = static_cast<u64>(r2);
  u64 cells = std::max<u64>(1, static_cast<u64>(1.0 / r));
  if (cells > (u64{1} << 15)) cells = u64{1} << 15;
  p.cells = cells;
  p.sm = mix64(seed);
  return p;
}

__host__ __device__ inline u64 cell_of(u32 c, u64 cells) { return (u64(c) * cells) >> 32; }

// (dx^2 + dy^2 <= r^2) in 65-bit arithmetic
__host__ __device__ inline bool rgg_close(u32 xi, u32 yi, u32 xj, u32 yj, u64 r2_hi, u64 r2_lo) {
  const u64 dx = xi > xj ? u64(xi - xj) : u64(xj - xi);
  const u64 dy = yi > yj ? u64(yi - yj) : u64(yj - yi);
  const u64 sx = dx * dx, sy = dy * dy;
  const u64 lo = sx + sy;
  const u64 hi = lo < sx ? 1u : 0u;
  return hi < r2_hi || (hi == r2_hi && lo <= r2_lo);
}

kmst_status rgg2d_host(u64 n, double avg_degree, u32 max_w, u64 seed, u32* src, u32* dst, u32* w,
                       u64 cap, u64* m_out) {
  const RggParams P = rgg_params(n, avg_degree, seed);
  const u64 cells = P.cells;
  std::vector<u32> x(n), y(n);
  for (u64 i = 0; i < n; ++i) {
    const u64 h = mix64(P.sm + i);
    x[i] = static_cast<u32>(h >> 32);
    y[i] = static_cast<u32>(h);
  }
  std::vector<u64> start(cells * cells + 1, 0);
  for (u64 i = 0; i < n; ++i) start[cell_of(y[i], cells) * cells + cell_of(x[i], cells) + 1]++;
  for (u64 c = 0; c < cells * cells; ++c) start[c + 1] += start[c];
  // renumber: vertex id = position in cell order
  {
    std::vector<u64> fill(start.begin(), start.end() - 1);
    std::vector<u32> nx(n), ny(n);
    for (u64 i = 0; i < n; ++i) {
      const u64 p = fill[cell_of(y[i], cells) * cells + cell_of(x[i], cells)]++;
      nx[p] = x[i];
      ny[p] = y[i];
    }
    x.swap(nx);
    y.swap(ny);
  }
  const unsigned T = host_threads();
  std::vector<std::vector<u64>> parts(T);
  parallel_chunks(n, T, [&](u64 lo, u64 hi, unsigned t) {
    auto& o = parts[t];
    o.reserve(static_cast<size_t>(avg_degree * double(hi - lo) * 1.1));
    for (u64 i = lo; i < hi; ++i) {
      const long cx = static_cast<long>(cell_of(x[i], cells));
      const long cy = static_cast<long>(cell_of(y[i], cells));
      // neighbour cells in row-major order hold ascending ids
      for (long dy = -1; dy <= 1; ++dy)
        for (long dx = -1; dx <= 1; ++dx) {
          const long gx = cx + dx, gy = cy + dy;
          if (gx < 0 || gy < 0 || gx >= long(cells) || gy >= long(cells)) continue;
          const u64 c = u64(gy) * cells + u64(gx);
          for (u64 k = start[c]; k < start[c + 1]; ++k)
            if (k != i && rgg_close(x[i], y[i], x[k], y[k], P.r2_hi, P.r2_lo))
              o.push_back((i << 32) | k);
        }
    }
  });
  return emit_pairs(parts, max_w, seed, src, dst, w, cap, m_out);
}

__global__ __launch_bounds__(BLOCK) void rgg_points_kernel(u64 sm, u64 n, u64 cells,
                                                           u32* __restrict__ key,
                                                           u32* __restrict__ idx) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += u64(gridDim.x) * BLOCK) {
    const u64 h = mix64(sm + i);
    key[i] = static_cast<u32>(cell_of(static_cast<u32>(h), cells) * cells +
                              cell_of(static_cast<u32>(h >> 32), cells));
    idx[i] = static_cast<u32>(i);
  }
}

__global__ __launch_bounds__(BLOCK) void rgg_place_kernel(u64 sm, u64 n, const u32* __restrict__ idx,
                                                          u32* __restrict__ X, u32* __restrict__ Y) {
  for (u64 p = u64(blockIdx.x) * BLOCK + threadIdx.x; p < n; p += u64(gridDim.x) * BLOCK) {
    const u64 h = mix64(sm + idx[p]);
    X[p] = static_cast<u32>(h >> 32);
    Y[p] = static_cast<u32>(h);
  }
}

// start[c] = first position whose sorted key is >= c, for c in [0, nkeys]
__global__ __launch_bounds__(BLOCK) void key_starts_kernel(const u32* __restrict__ key, u64 n,
                                                           u64 nkeys, u32* __restrict__ start) {
  for (u64 p = u64(blockIdx.x) * BLOCK + threadIdx.x; p <= n; p += u64(gridDim.x) * BLOCK) {
    const u64 prev = p ? u64(key[p - 1]) + 1 : 0;
    const u64 cur = p < n ? u64(key[p]) : nkeys;
    for (u64 c = prev; c <= cur; ++c) start[c] = static_cast<u32>(p);
  }
}

template <bool EMIT>
__global__ __launch_bounds__(BLOCK) void rgg_adj_kernel(RggParams P, const u32* __restrict__ X,
                                                        const u32* __restrict__ Y,
                                                        const u32* __restrict__ start,
                                                        u64* __restrict__ deg,
                                                        u32* __restrict__ os, u32* __restrict__ od,
                                                        u32* __restrict__ ow, u64 wseed, u32 max_w) {
  const u64 cells = P.cells;
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < P.n; i += u64(gridDim.x) * BLOCK) {
    const u32 xi = X[i], yi = Y[i];
    const long cx = static_cast<long>(cell_of(xi, cells)), cy = static_cast<long>(cell_of(yi, cells));
    u64 at = EMIT ? deg[i] : 0;  // EMIT: deg holds the exclusive scan
    u64 found = 0;
    for (long dy = -1; dy <= 1; ++dy)
      for (long dx = -1; dx <= 1; ++dx) {
        const long gx = cx + dx, gy = cy + dy;
        if (gx < 0 || gy < 0 || gx >= long(cells) || gy >= long(cells)) continue;
        const u64 c = u64(gy) * cells + u64(gx);
        const u64 k1 = start[c + 1];
        for (u64 k = start[c]; k < k1; ++k) {
          if (k == i || !rgg_close(xi, yi, X[k], Y[k], P.r2_hi, P.r2_lo)) continue;
          if (EMIT) {
            os[at] = static_cast<u32>(i);
            od[at] = static_cast<u32>(k);
            ow[at] = geo_weight(wseed, static_cast<u32>(i), static_cast<u32>(k), max_w);
            ++at;
          }
          ++found;
        }
      }
    if (!EMIT) deg[i] = found;
  }
}

static u32 grid_of(u64 n) { return static_cast<u32>(std::min<u64>((n + BLOCK - 1) / BLOCK, 65536)); }

// Degrees (deg, n entries) -> exclusive scan in place; returns the total.
u64 scan_degrees(kmst_context& C, u64* deg, u64 n) {
  hipStream_t st = C.stream;
  u64* scratch = static_cast<u64*>(C.scan_scratch.get(scan_scratch_elems(n) * 8 + 64));
  u64* total = static_cast<u64*>(C.counters.get(256));
  scan_exclusive_u64(st, deg, deg, n, scratch, total);
  u64 h = 0;
  KMST_HIP(hipMemcpyAsync(&h, total, 8, hipMemcpyDeviceToHost, st));
  KMST_HIP(hipStreamSynchronize(st));
  return h;
}

// Stable sort of (key, val) pairs by the low `bits` bits of key; returns the
// arrays holding the result.
void sort_pairs_u32(kmst_context& C, u32*& key, u32*& val, u32* key_alt, u32* val_alt, u64 n,
                    u32 bits) {
  hipStream_t st = C.stream;
  C.sort_work.bind(st, static_cast<u32*>(C.hist.get(sort_work_words(n) * 4)), n);
  apply_tuning(C);
  u32* err = static_cast<u32*>(C.counters.get(256)) + 8;
  fill_u32(st, err, 1, 0);
  if (radix_sort(st, key, val, nullptr, key_alt, val_alt, nullptr, n, 0, bits, C.sort_work, err)) {
    key = key_alt;
    val = val_alt;
  }
}

void check_sort(kmst_context& C, const char* what) {
  u32 e = 0;
  KMST_HIP(hipMemcpyAsync(&e, static_cast<u32*>(C.counters.get(256)) + 8, 4,
                          hipMemcpyDeviceToHost, C.stream));
  KMST_HIP(hipStreamSynchronize(C.stream));
  if (e) throw Error(KMST_EDEVICE, std::string(what) + ": sort wait expired");
}

void rgg2d_device(kmst_context& C, u64 n, double avg_degree, u32 max_w, u64 seed, u32* src,
                  u32* dst, u32* w, u64 cap, u64* m_out) {
  hipStream_t st = C.stream;
  const RggParams P = rgg_params(n, avg_degree, seed);
  const u64 ncells = P.cells * P.cells;
  u32* key = static_cast<u32*>(C.gen_a.get(n * 4));
  u32* idx = static_cast<u32*>(C.gen_b.get(n * 4));
  u32* key2 = static_cast<u32*>(C.xs.get(n * 4));
  u32* idx2 = static_cast<u32*>(C.xd.get(n * 4));
  u32* X = static_cast<u32*>(C.ys.get(n * 4));
  u32* Y = static_cast<u32*>(C.yd.get(n * 4));
  u32* start = static_cast<u32*>(C.gen_d.get((ncells + 1) * 4));
  u64* deg = static_cast<u64*>(C.gen_c.get((n + 1) * 8));
  {
    KTScope t(st, K_GEN, 8.0 * n);
    rgg_points_kernel<<<grid_of(n), BLOCK, 0, st>>>(P.sm, n, P.cells, key, idx);
    KMST_HIP(hipGetLastError());
  }
  // points in cell order, generation order inside a cell (stable)
  sort_pairs_u32(C, key, idx, key2, idx2, n, bits_for(ncells));
  {
    KTScope t(st, K_GEN, 12.0 * n + 4.0 * ncells);
    key_starts_kernel<<<grid_of(n + 1), BLOCK, 0, st>>>(key, n, ncells, start);
    KMST_HIP(hipGetLastError());
    rgg_place_kernel<<<grid_of(n), BLOCK, 0, st>>>(P.sm, n, idx, X, Y);
    KMST_HIP(hipGetLastError());
  }
  check_sort(C, "RGG2D");
  const u64 wseed = mix64(seed ^ GEO_WEIGHT_SALT);
  {
    KTScope t(st, K_GEN, 16.0 * n);
    rgg_adj_kernel<false><<<grid_of(n), BLOCK, 0, st>>>(P, X, Y, start, deg, nullptr, nullptr,
                                                        nullptr, wseed, max_w);
    KMST_HIP(hipGetLastError());
  }
  const u64 total = scan_degrees(C, deg, n);
  *m_out = total;
  if (cap < total || !src || !dst || !w) throw Error(KMST_EINVAL, "RGG2D: output capacity too small");
  {
    KTScope t(st, K_GEN, 16.0 * n + 12.0 * total);
    rgg_adj_kernel<true><<<grid_of(n), BLOCK, 0, st>>>(P, X, Y, start, deg, src, dst, w, wseed,
                                                       max_w);
    KMST_HIP(hipGetLastError());
  }
  KMST_HIP(hipStreamSynchronize(st));
}

// =================================================================== RHG
struct RhgParams {
  u64 n;
  double alpha, R, coshR, cr, slack;
  u32 nb;
  u64 sm;
};

RhgParams rhg_params(u64 n, double avg_degree, double gamma, u64 seed) {
  RhgParams p;
  p.n = n;
  p.alpha = (gamma - 1.0) / 2.0;
  const double xi = p.alpha / (p.alpha - 0.5);
  p.R = 2.0 * det::log((2.0 / det::PI) * xi * xi * double(n) / avg_degree);
  p.coshR = det::cosh(p.R);
  p.cr = det::cosh(p.alpha * p.R) - 1.0;
  // slack on the test's right-hand side for the candidate windows (far above
  // the rounding of cosh R - cosh(.))
  p.slack = 1e-12 * p.coshR;
  p.nb = static_cast<u32>(std::max(1, static_cast<int>(std::ceil(p.R))));  // radial bands
  p.sm = mix64(seed ^ 0x524847ull);
  return p;
}

__host__ __device__ inline void rhg_point(const RhgParams& P, u64 i, double& theta, double& r) {
  const u64 h1 = mix64(P.sm + 2 * i), h2 = mix64(P.sm + 2 * i + 1);
  theta = det::TWO_PI * unit(h1);
  r = det::acosh1p(unit(h2) * P.cr) / P.alpha;
}

__host__ __device__ inline u32 rhg_band(const RhgParams& P, double r) {
  const int b = static_cast<int>(r / P.R * double(P.nb));
  return static_cast<u32>(b < 0 ? 0 : (b >= int(P.nb) ? int(P.nb) - 1 : b));
}

// the test, with a < b
__host__ __device__ inline bool rhg_adjacent(double ta, double ra, double sha, double tb, double rb,
                                             double shb, double coshR) {
  const double h = det::sin(0.5 * (ta - tb));
  return 2.0 * h * h * sha * shb <= coshR - det::cosh(ra - rb);
}

// Half-width of the angular window of a vertex (radius ri, sinh ri = shi)
// in a band of smallest radius rb; PI or more: the whole band.
__host__ __device__ inline double rhg_window(const RhgParams& P, double ri, double shi, double rb) {
  const double den = 2.0 * shi * det::sinh(rb);
  if (!(den > 0.0)) return det::PI;
  const double x = (P.coshR - det::cosh(ri - rb) + P.slack) / den;  // sin^2(dmax / 2)
  if (x >= 1.0) return det::PI;
  const double d = 2.0 * asin(sqrt(x > 0.0 ? x : 0.0)) * (1.0 + 1e-9) + 1e-15;
  return d < det::PI ? d : det::PI;
}

kmst_status rhg_host(u64 n, double avg_degree, double gamma, u32 max_w, u64 seed, u32* src,
                     u32* dst, u32* w, u64 cap, u64* m_out) {
  const RhgParams P = rhg_params(n, avg_degree, gamma, seed);
  const unsigned T = host_threads();
  struct Pt {
    double theta, r;
  };
  std::vector<Pt> pts(n);
  parallel_chunks(n, T, [&](u64 lo, u64 hi, unsigned) {
    for (u64 i = lo; i < hi; ++i) rhg_point(P, i, pts[i].theta, pts[i].r);
  });
  std::sort(pts.begin(), pts.end(), [](const Pt& a, const Pt& b) {
    return a.theta < b.theta || (a.theta == b.theta && a.r < b.r);
  });
  std::vector<double> shr(n);
  parallel_chunks(n, T, [&](u64 lo, u64 hi, unsigned) {
    for (u64 i = lo; i < hi; ++i) shr[i] = det::sinh(pts[i].r);
  });
  const u32 nb = P.nb;
  std::vector<std::vector<u32>> band(nb);  // ids in angle order
  std::vector<double> band_rmin(nb, P.R);
  for (u64 i = 0; i < n; ++i) {
    const u32 b = rhg_band(P, pts[i].r);
    band[b].push_back(static_cast<u32>(i));
    band_rmin[b] = std::min(band_rmin[b], pts[i].r);
  }
  std::vector<std::vector<u64>> parts(T);
  parallel_chunks(n, T, [&](u64 lo, u64 hi, unsigned t) {
    auto& o = parts[t];
    o.reserve(static_cast<size_t>(avg_degree * double(hi - lo) * 1.2));
    std::vector<u32> nbrs;
    for (u64 i = lo; i < hi; ++i) {
      nbrs.clear();
      const double ti = pts[i].theta;
      auto test = [&](u32 j) {
        if (j == i) return;
        const u64 a = std::min<u64>(i, j), b = std::max<u64>(i, j);
        if (rhg_adjacent(pts[a].theta, pts[a].r, shr[a], pts[b].theta, pts[b].r, shr[b], P.coshR))
          nbrs.push_back(j);
      };
      for (u32 b = 0; b < nb; ++b) {
        const auto& ids = band[b];
        if (ids.empty()) continue;
        const double dmax = rhg_window(P, pts[i].r, shr[i], band_rmin[b]);
        if (dmax >= det::PI) {
          for (u32 j : ids) test(j);
          continue;
        }
        // ids are sorted by angle: scan [lo, hi] with wrap-around
        auto scan = [&](double a0, double a1) {
          auto it = std::lower_bound(ids.begin(), ids.end(), a0,
                                     [&](u32 id, double v) { return pts[id].theta < v; });
          for (; it != ids.end() && pts[*it].theta <= a1; ++it) test(*it);
        };
        const double a0 = ti - dmax, a1 = ti + dmax;
        if (a0 < 0.0) {
          scan(0.0, a1);
          scan(a0 + det::TWO_PI, det::TWO_PI);
        } else if (a1 >= det::TWO_PI) {
          scan(a0, det::TWO_PI);
          scan(0.0, a1 - det::TWO_PI);
        } else {
          scan(a0, a1);
        }
      }
      std::sort(nbrs.begin(), nbrs.end());
      nbrs.erase(std::unique(nbrs.begin(), nbrs.end()), nbrs.end());
      for (u32 j : nbrs) o.push_back((i << 32) | j);
    }
  });
  return emit_pairs(parts, max_w, seed, src, dst, w, cap, m_out);
}

// ---- RHG on the device
__global__ __launch_bounds__(BLOCK) void rhg_points_kernel(RhgParams P, u32* __restrict__ kh,
                                                           u32* __restrict__ kl,
                                                           u64* __restrict__ idx) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < P.n; i += u64(gridDim.x) * BLOCK) {
    double t, r;
    rhg_point(P, i, t, r);
    const u64 b = __builtin_bit_cast(u64, r);  // r >= 0: bits order as values
    kh[i] = static_cast<u32>(b >> 32);
    kl[i] = static_cast<u32>(b);
    idx[i] = i;
  }
}

__global__ __launch_bounds__(BLOCK) void rhg_theta_keys_kernel(RhgParams P,
                                                               const u64* __restrict__ idx,
                                                               u32* __restrict__ kh,
                                                               u32* __restrict__ kl) {
  for (u64 p = u64(blockIdx.x) * BLOCK + threadIdx.x; p < P.n; p += u64(gridDim.x) * BLOCK) {
    double t, r;
    rhg_point(P, idx[p], t, r);
    const u64 b = __builtin_bit_cast(u64, t);  // theta >= 0
    kh[p] = static_cast<u32>(b >> 32);
    kl[p] = static_cast<u32>(b);
  }
}

// vertex p = idx[p]-th point: its angle, radius, sinh of the radius, band;
// the band's smallest radius by an atomic minimum of the radius bits
__global__ __launch_bounds__(BLOCK) void rhg_place_kernel(RhgParams P, const u64* __restrict__ idx,
                                                          double* __restrict__ th,
                                                          double* __restrict__ rr,
                                                          double* __restrict__ sh,
                                                          u32* __restrict__ band,
                                                          u32* __restrict__ pid,
                                                          unsigned long long* __restrict__ brmin) {
  for (u64 p = u64(blockIdx.x) * BLOCK + threadIdx.x; p < P.n; p += u64(gridDim.x) * BLOCK) {
    double t, r;
    rhg_point(P, idx[p], t, r);
    th[p] = t;
    rr[p] = r;
    sh[p] = det::sinh(r);
    const u32 b = rhg_band(P, r);
    band[p] = b;
    pid[p] = static_cast<u32>(p);
    atomicMin(brmin + b, static_cast<unsigned long long>(__builtin_bit_cast(u64, r)));
  }
}

__global__ __launch_bounds__(BLOCK) void fill_u64_bits_kernel(unsigned long long* p, u64 n, u64 v) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < n; i += u64(gridDim.x) * BLOCK) p[i] = v;
}

struct RhgAdj {
  RhgParams P;
  const double* th;
  const double* rr;
  const double* sh;
  const u32* blist;   // vertex ids grouped by band, ascending (angle order) inside a band
  const u32* bstart;  // nb + 1 entries
  const unsigned long long* brmin;
  u64* deg;           // count pass: degrees; emit pass: their exclusive scan
  u32* os;
  u32* od;
};

__device__ __forceinline__ u32 band_lower(const RhgAdj& a, u32 lo, u32 hi, double x) {
  // first k in [lo, hi) with th[blist[k]] >= x
  while (lo < hi) {
    const u32 mid = lo + (hi - lo) / 2;
    if (a.th[a.blist[mid]] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ u32 band_upper(const RhgAdj& a, u32 lo, u32 hi, double x) {
  // first k in [lo, hi) with th[blist[k]] > x
  while (lo < hi) {
    const u32 mid = lo + (hi - lo) / 2;
    if (a.th[a.blist[mid]] <= x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// One wave per vertex.  Lane l takes band g + (l & 31) of a group of 32
// bands, lanes 32..63 the wrapped part of the window; the candidates of all
// 64 ranges are then tested 64 at a time, each lane finding its range by a
// binary search over the lanes' exclusive range offsets.
template <bool EMIT>
__global__ __launch_bounds__(BLOCK) void rhg_adj_kernel(RhgAdj a) {
  const u64 p = u64(blockIdx.x) * WAVES + wave_id();
  if (p >= a.P.n) return;  // wave-uniform
  const int lane = lane_id();
  const double ti = a.th[p], ri = a.rr[p], shi = a.sh[p];
  const u64 base = EMIT ? a.deg[p] : 0;
  u64 found = 0;
  for (u32 g = 0; g < a.P.nb; g += 32) {
    const u32 b = g + static_cast<u32>(lane & 31);
    const bool wrapped = lane >= 32;
    u32 lo = 0, hi = 0;
    if (b < a.P.nb) {
      const u32 b0 = a.bstart[b], b1 = a.bstart[b + 1];
      if (b0 < b1) {
        const double rb = __builtin_bit_cast(double, static_cast<u64>(a.brmin[b]));
        const double dmax = rhg_window(a.P, ri, shi, rb);
        if (dmax >= det::PI) {
          if (!wrapped) {
            lo = b0;
            hi = b1;
          }
        } else {
          const double a0 = ti - dmax, a1 = ti + dmax;
          double x0 = 0.0, x1 = -1.0;  // empty unless set
          if (a0 < 0.0) {
            x0 = wrapped ? a0 + det::TWO_PI : 0.0;
            x1 = wrapped ? det::TWO_PI : a1;
          } else if (a1 >= det::TWO_PI) {
            x0 = wrapped ? 0.0 : a0;
            x1 = wrapped ? a1 - det::TWO_PI : det::TWO_PI;
          } else if (!wrapped) {
            x0 = a0;
            x1 = a1;
          }
          if (x1 >= x0) {
            lo = band_lower(a, b0, b1, x0);
            hi = band_upper(a, lo, b1, x1);
          }
        }
      }
    }
    const u32 len = hi - lo;
    const u32 incl = wave_inclusive_sum(len);
    const u32 excl = incl - len;
    const u32 total = __shfl(incl, WAVE - 1, WAVE);
    for (u32 t0 = 0; t0 < total; t0 += WAVE) {
      const u32 t = t0 + static_cast<u32>(lane);
      // the range holding candidate t: the last lane whose offset is <= t
      int o = 0;
#pragma unroll
      for (int step = 32; step > 0; step >>= 1) {
        const u32 e = __shfl(excl, o + step, WAVE);
        if (e <= t) o += step;
      }
      const u32 olo = __shfl(lo, o, WAVE), oex = __shfl(excl, o, WAVE);
      bool adj = false;
      u32 j = 0;
      if (t < total) {
        j = a.blist[olo + (t - oex)];
        if (j != p) {
          const u64 x = j < p ? j : p, y = j < p ? p : j;
          adj = rhg_adjacent(a.th[x], a.rr[x], a.sh[x], a.th[y], a.rr[y], a.sh[y], a.P.coshR);
        }
      }
      const u64 bal = __ballot(adj);
      if (EMIT && adj) {
        const u64 at = base + found + mbcnt(bal);
        a.os[at] = static_cast<u32>(p);
        a.od[at] = j;
      }
      found += static_cast<u64>(__popcll(bal));
    }
  }
  if (!EMIT && lane == 0) a.deg[p] = found;
}

__global__ __launch_bounds__(BLOCK) void geo_weights_kernel(const u32* __restrict__ s,
                                                            const u32* __restrict__ d, u64 m,
                                                            u64 wseed, u32 max_w,
                                                            u32* __restrict__ w) {
  for (u64 i = u64(blockIdx.x) * BLOCK + threadIdx.x; i < m; i += u64(gridDim.x) * BLOCK)
    w[i] = geo_weight(wseed, s[i], d[i], max_w);
}

void rhg_device(kmst_context& C, u64 n, double avg_degree, double gamma, u32 max_w, u64 seed,
                u32* src, u32* dst, u32* w, u64 cap, u64* m_out) {
  hipStream_t st = C.stream;
  const RhgParams P = rhg_params(n, avg_degree, gamma, seed);
  const u32 nb = P.nb;
  // order of the points: stable sorts by radius, then by angle
  u32* kh = static_cast<u32*>(C.gen_a.get(n * 4));
  u32* kl = static_cast<u32*>(C.gen_b.get(n * 4));
  u64* idx = static_cast<u64*>(C.gen_c.get(n * 8));
  u32* kh2 = static_cast<u32*>(C.xs.get(n * 4));
  u32* kl2 = static_cast<u32*>(C.xd.get(n * 4));
  u64* idx2 = static_cast<u64*>(C.xv.get(n * 8));
  C.sort_work.bind(st, static_cast<u32*>(C.hist.get(sort_work_words(n) * 4)), n);
  apply_tuning(C);
  u32* err = static_cast<u32*>(C.counters.get(256)) + 8;
  fill_u32(st, err, 1, 0);
  {
    KTScope t(st, K_GEN, 16.0 * n);
    rhg_points_kernel<<<grid_of(n), BLOCK, 0, st>>>(P, kh, kl, idx);
    KMST_HIP(hipGetLastError());
  }
  auto sort64 = [&] {  // 63-bit keys (non-negative doubles), index carried
    if (radix_sort(st, kh, kl, idx, kh2, kl2, idx2, n, 32, 63, C.sort_work, err)) {
      std::swap(kh, kh2);
      std::swap(kl, kl2);
      std::swap(idx, idx2);
    }
  };
  sort64();
  {
    KTScope t(st, K_GEN, 16.0 * n);
    rhg_theta_keys_kernel<<<grid_of(n), BLOCK, 0, st>>>(P, idx, kh, kl);
    KMST_HIP(hipGetLastError());
  }
  sort64();
  // per-vertex values in the final order; bands
  double* th = static_cast<double*>(C.val0.get(n * 24));
  double* rr = th + n;
  double* sh = rr + n;
  u32* band = kh;  // free after the sorts
  u32* pid = kl;
  auto* brmin = static_cast<unsigned long long*>(C.samples.get(4096 * 4));
  if (nb > 2048) throw Error(KMST_EINVAL, "RHG: too many radial bands");
  fill_u64_bits_kernel<<<1, BLOCK, 0, st>>>(brmin, nb, __builtin_bit_cast(u64, P.R));
  KMST_HIP(hipGetLastError());
  {
    KTScope t(st, K_GEN, 48.0 * n);
    rhg_place_kernel<<<grid_of(n), BLOCK, 0, st>>>(P, idx, th, rr, sh, band, pid, brmin);
    KMST_HIP(hipGetLastError());
  }
  u32* blist = pid;
  sort_pairs_u32(C, band, blist, kh2, kl2, n, bits_for(nb));
  u32* bstart = static_cast<u32*>(C.gen_d.get((u64(nb) + 1) * 4));
  key_starts_kernel<<<grid_of(n + 1), BLOCK, 0, st>>>(band, n, nb, bstart);
  KMST_HIP(hipGetLastError());
  check_sort(C, "RHG");
  u64* deg = static_cast<u64*>(C.rep.get((n + 1) * 8));
  RhgAdj a{P, th, rr, sh, blist, bstart, brmin, deg, nullptr, nullptr};
  const u32 grid = ceil_div_u32(n, WAVES);
  {
    KTScope t(st, K_GEN, 32.0 * n);
    rhg_adj_kernel<false><<<grid, BLOCK, 0, st>>>(a);
    KMST_HIP(hipGetLastError());
  }
  const u64 total = scan_degrees(C, deg, n);
  *m_out = total;
  if (cap < total || !src || !dst || !w) throw Error(KMST_EINVAL, "RHG: output capacity too small");
  if (total == 0) return;
  a.os = src;
  a.od = dst;
  {
    KTScope t(st, K_GEN, 32.0 * n + 8.0 * total);
    rhg_adj_kernel<true><<<grid, BLOCK, 0, st>>>(a);
    KMST_HIP(hipGetLastError());
  }
  // each vertex's records are contiguous, their dst in band order: sort the
  // records by (src, dst)
  const u32 vb = bits_for(n);
  u32* as = static_cast<u32*>(C.ys.get(total * 4));
  u32* ad = static_cast<u32*>(C.yd.get(total * 4));
  C.sort_work.bind(st, static_cast<u32*>(C.hist.get(sort_work_words(total) * 4)), total);
  fill_u32(st, err, 1, 0);
  if (radix_sort(st, src, dst, nullptr, as, ad, nullptr, total, vb, 2 * vb, C.sort_work, err)) {
    KMST_HIP(hipMemcpyAsync(src, as, total * 4, hipMemcpyDeviceToDevice, st));
    KMST_HIP(hipMemcpyAsync(dst, ad, total * 4, hipMemcpyDeviceToDevice, st));
  }
  {
    KTScope t(st, K_GEN, 12.0 * total);
    geo_weights_kernel<<<grid_of(total), BLOCK, 0, st>>>(src, dst, total,
                                                         mix64(seed ^ GEO_WEIGHT_SALT), max_w, w);
    KMST_HIP(hipGetLastError());
  }
  check_sort(C, "RHG");
}

bool geo_args_ok(u64 n, double avg_degree, u32 max_w, const u64* m_out) {
  return m_out && n >= 2 && n < (u64{1} << 31) && max_w != 0 && avg_degree > 0.0;
}

}  // namespace

}  // namespace kmst

using kmst::u64;

extern "C" {

kmst_status kmst_gen_rgg2d_host(uint64_t n, double avg_degree, uint32_t max_weight, uint64_t seed,
                                uint32_t* src, uint32_t* dst, uint32_t* weight, uint64_t capacity,
                                uint64_t* m_directed) {
  if (!kmst::geo_args_ok(n, avg_degree, max_weight, m_directed)) return KMST_EINVAL;
  try {
    return kmst::rgg2d_host(n, avg_degree, max_weight, seed, src, dst, weight, capacity,
                            m_directed);
  } catch (const std::bad_alloc&) {
    return KMST_ENOMEM;
  }
}

kmst_status kmst_gen_rhg_host(uint64_t n, double avg_degree, double gamma, uint32_t max_weight,
                              uint64_t seed, uint32_t* src, uint32_t* dst, uint32_t* weight,
                              uint64_t capacity, uint64_t* m_directed) {
  if (!kmst::geo_args_ok(n, avg_degree, max_weight, m_directed) || !(gamma > 2.0))
    return KMST_EINVAL;
  try {
    return kmst::rhg_host(n, avg_degree, gamma, max_weight, seed, src, dst, weight, capacity,
                          m_directed);
  } catch (const std::bad_alloc&) {
    return KMST_ENOMEM;
  }
}

kmst_status kmst_gen_rgg2d_device(kmst_context* ctx, uint64_t n, double avg_degree,
                                  uint32_t max_weight, uint64_t seed, uint32_t* src,
                                  uint32_t* dst, uint32_t* weight, uint64_t capacity,
                                  uint64_t* m_directed) {
  if (!ctx || !kmst::geo_args_ok(n, avg_degree, max_weight, m_directed)) return KMST_EINVAL;
  return kmst::guarded_call([&] {
    KMST_HIP(hipSetDevice(ctx->device));
    kmst::rgg2d_device(*ctx, n, avg_degree, max_weight, seed, src, dst, weight, capacity,
                       m_directed);
  });
}

kmst_status kmst_gen_rhg_device(kmst_context* ctx, uint64_t n, double avg_degree, double gamma,
                                uint32_t max_weight, uint64_t seed, uint32_t* src, uint32_t* dst,
                                uint32_t* weight, uint64_t capacity, uint64_t* m_directed) {
  if (!ctx || !kmst::geo_args_ok(n, avg_degree, max_weight, m_directed) || !(gamma > 2.0))
    return KMST_EINVAL;
  return kmst::guarded_call([&] {
    KMST_HIP(hipSetDevice(ctx->device));
    kmst::rhg_device(*ctx, n, avg_degree, gamma, max_weight, seed, src, dst, weight, capacity,
                     m_directed);
  });
}

}  // extern "C"
