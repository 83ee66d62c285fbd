// kmst_gen_geo.cpp -- deterministic geometric input generators on the host
// (the role KaGen's RGG_2D and RHG play for the reference driver,
// benchmarks/mst_benchmarks.hpp:116-199).  Output: symmetric records sorted
// by (src, dst), weights from a hash of (min, max) so both directions agree.
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
// (cosh(alpha R) - 1), angles uniform.  {u, v} is an edge iff their
// hyperbolic distance is at most R, i.e. (law of cosines in haversine form,
// well conditioned for the tiny angles of boundary pairs)
//   2 sin^2(dtheta / 2) sinh r_u sinh r_v <= cosh R - cosh(r_u - r_v),
// evaluated in double precision with the lower id first, so the test is the
// same from both sides.  Candidates come from radial bands sorted by angle;
// a band's angular window is taken at its smallest member radius, where it is
// widest (the bound above decreases in r_v), with a relative margin of 1e-9.
// Vertex ids follow the angular order (neighbours have nearby ids).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "../../include/kmst.h"

namespace {

using u32 = std::uint32_t;
using u64 = std::uint64_t;

inline u64 mix64(u64 x) {  // SplitMix64 finaliser
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
inline double unit(u64 h) { return double(h >> 11) * (1.0 / 9007199254740992.0); }  // [0,1)

struct Out {
  std::vector<u64> pairs;  // (u << 32 | v), both directions
};

kmst_status emit(const Out& o, u32 max_w, u64 seed, u32* src, u32* dst, u32* w, u64 cap,
                 u64* m_out) {
  *m_out = o.pairs.size();
  if (!src || cap < o.pairs.size()) return KMST_EINVAL;
  const u64 wseed = mix64(seed ^ 0x5745494748543A32ull);
  for (u64 i = 0; i < o.pairs.size(); ++i) {
    const u32 s = static_cast<u32>(o.pairs[i] >> 32), d = static_cast<u32>(o.pairs[i]);
    const u32 a = s < d ? s : d, b = s < d ? d : s;
    src[i] = s;
    dst[i] = d;
    w[i] = 1u + static_cast<u32>(mix64(wseed ^ ((u64(a) << 32) | b)) % max_w);
  }
  return KMST_OK;
}

Out rgg2d(u64 n, double avg_degree, u64 seed) {
  const double r = std::sqrt(avg_degree / (M_PI * double(n)));
  const double scale = 4294967296.0;  // coordinates in [0, 2^32)
  const u64 rr = static_cast<u64>(r * scale);
  const unsigned __int128 r2 = (unsigned __int128)rr * rr;
  u64 cells = std::max<u64>(1, static_cast<u64>(1.0 / r));
  if (cells > (u64{1} << 15)) cells = u64{1} << 15;
  std::vector<u32> x(n), y(n);
  const u64 sm = mix64(seed);
  for (u64 i = 0; i < n; ++i) {
    const u64 h = mix64(sm + i);
    x[i] = static_cast<u32>(h >> 32);
    y[i] = static_cast<u32>(h);
  }
  auto cell_of = [&](u32 c) { return (u64(c) * cells) >> 32; };
  std::vector<u64> start(cells * cells + 1, 0);
  for (u64 i = 0; i < n; ++i) start[cell_of(y[i]) * cells + cell_of(x[i]) + 1]++;
  for (u64 c = 0; c < cells * cells; ++c) start[c + 1] += start[c];
  // renumber: vertex id = position in cell order
  {
    std::vector<u64> fill(start.begin(), start.end() - 1);
    std::vector<u32> nx(n), ny(n);
    for (u64 i = 0; i < n; ++i) {
      const u64 p = fill[cell_of(y[i]) * cells + cell_of(x[i])]++;
      nx[p] = x[i];
      ny[p] = y[i];
    }
    x.swap(nx);
    y.swap(ny);
  }
  Out o;
  o.pairs.reserve(static_cast<size_t>(avg_degree * double(n) * 1.1));
  std::vector<u32> nb;
  for (u64 i = 0; i < n; ++i) {
    nb.clear();
    const long cx = static_cast<long>(cell_of(x[i])), cy = static_cast<long>(cell_of(y[i]));
    for (long dy = -1; dy <= 1; ++dy)
      for (long dx = -1; dx <= 1; ++dx) {
        const long gx = cx + dx, gy = cy + dy;
        if (gx < 0 || gy < 0 || gx >= long(cells) || gy >= long(cells)) continue;
        const u64 c = u64(gy) * cells + u64(gx);
        for (u64 k = start[c]; k < start[c + 1]; ++k) {
          const u32 j = static_cast<u32>(k);  // ids are cell-ordered
          if (j == i) continue;
          const u64 ddx = x[i] > x[j] ? x[i] - x[j] : x[j] - x[i];
          const u64 ddy = y[i] > y[j] ? y[i] - y[j] : y[j] - y[i];
          const unsigned __int128 d2 = (unsigned __int128)ddx * ddx + (unsigned __int128)ddy * ddy;
          if (d2 <= r2) nb.push_back(j);
        }
      }
    std::sort(nb.begin(), nb.end());
    for (u32 j : nb) o.pairs.push_back((i << 32) | j);
  }
  return o;
}

Out rhg(u64 n, double avg_degree, double gamma, u64 seed) {
  const double alpha = (gamma - 1.0) / 2.0;
  const double xi = alpha / (alpha - 0.5);
  const double R = 2.0 * std::log((2.0 / M_PI) * xi * xi * double(n) / avg_degree);
  const u64 sm = mix64(seed ^ 0x524847ull);
  struct P {
    double theta, r;
  };
  std::vector<P> pts(n);
  const double cr = std::cosh(alpha * R) - 1.0;
  for (u64 i = 0; i < n; ++i) {
    const u64 h1 = mix64(sm + 2 * i), h2 = mix64(sm + 2 * i + 1);
    pts[i].theta = 2.0 * M_PI * unit(h1);
    pts[i].r = std::acosh(1.0 + unit(h2) * cr) / alpha;
  }
  std::sort(pts.begin(), pts.end(), [](const P& a, const P& b) {
    return a.theta < b.theta || (a.theta == b.theta && a.r < b.r);
  });
  std::vector<double> shr(n);
  for (u64 i = 0; i < n; ++i) shr[i] = std::sinh(pts[i].r);
  const double coshR = std::cosh(R);
  auto adjacent = [&](u64 a, u64 b) {  // a < b: same expression from both sides
    const double h = std::sin(0.5 * (pts[a].theta - pts[b].theta));
    return 2.0 * h * h * shr[a] * shr[b] <= coshR - std::cosh(pts[a].r - pts[b].r);
  };
  // radial bands of width about 1
  const int nb = std::max(1, static_cast<int>(std::ceil(R)));
  std::vector<std::vector<u32>> band(nb);  // ids sorted by angle (ids are angle-ordered)
  std::vector<double> band_rmin(nb, R);
  for (u64 i = 0; i < n; ++i) {
    int b = static_cast<int>(pts[i].r / R * nb);
    if (b >= nb) b = nb - 1;
    if (b < 0) b = 0;
    band[b].push_back(static_cast<u32>(i));
    band_rmin[b] = std::min(band_rmin[b], pts[i].r);
  }
  Out o;
  o.pairs.reserve(static_cast<size_t>(avg_degree * double(n) * 1.2));
  std::vector<u32> nbrs;
  for (u64 i = 0; i < n; ++i) {
    nbrs.clear();
    const double ti = pts[i].theta;
    for (int b = 0; b < nb; ++b) {
      if (band[b].empty()) continue;
      const double rb = band_rmin[b];  // smallest radius: the widest angular window
      double dmax = M_PI;
      const double den = 2.0 * shr[i] * std::sinh(rb);
      if (den > 0.0) {
        const double x = (coshR - std::cosh(pts[i].r - rb)) / den;  // sin^2(dmax / 2)
        if (x < 1.0) dmax = std::min(M_PI, 2.0 * std::asin(std::sqrt(std::max(0.0, x))) *
                                               (1.0 + 1e-9) + 1e-15);
      }
      const auto& ids = band[b];
      if (dmax >= M_PI) {
        for (u32 j : ids)
          if (j != i && adjacent(std::min<u64>(i, j), std::max<u64>(i, j))) nbrs.push_back(j);
        continue;
      }
      // ids are sorted by angle; scan [ti - dmax, ti + dmax] with wrap-around
      auto scan = [&](double lo, double hi) {
        auto it = std::lower_bound(ids.begin(), ids.end(), lo,
                                   [&](u32 id, double v) { return pts[id].theta < v; });
        for (; it != ids.end() && pts[*it].theta <= hi; ++it) {
          const u32 j = *it;
          if (j != i && adjacent(std::min<u64>(i, j), std::max<u64>(i, j))) nbrs.push_back(j);
        }
      };
      const double lo = ti - dmax, hi = ti + dmax;
      if (lo < 0.0) {
        scan(0.0, hi);
        scan(lo + 2.0 * M_PI, 2.0 * M_PI);
      } else if (hi >= 2.0 * M_PI) {
        scan(lo, 2.0 * M_PI);
        scan(0.0, hi - 2.0 * M_PI);
      } else {
        scan(lo, hi);
      }
    }
    std::sort(nbrs.begin(), nbrs.end());
    nbrs.erase(std::unique(nbrs.begin(), nbrs.end()), nbrs.end());
    for (u32 j : nbrs) o.pairs.push_back((i << 32) | j);
  }
  return o;
}

}  // namespace

extern "C" {

kmst_status kmst_gen_rgg2d_host(uint64_t n, double avg_degree, uint32_t max_weight, uint64_t seed,
                                uint32_t* src, uint32_t* dst, uint32_t* weight, uint64_t capacity,
                                uint64_t* m_directed) {
  if (!m_directed || n < 2 || n >= (u64{1} << 31) || max_weight == 0 || !(avg_degree > 0.0))
    return KMST_EINVAL;
  try {
    return emit(rgg2d(n, avg_degree, seed), max_weight, seed, src, dst, weight, capacity,
                m_directed);
  } catch (const std::bad_alloc&) {
    return KMST_ENOMEM;
  }
}

kmst_status kmst_gen_rhg_host(uint64_t n, double avg_degree, double gamma, uint32_t max_weight,
                              uint64_t seed, uint32_t* src, uint32_t* dst, uint32_t* weight,
                              uint64_t capacity, uint64_t* m_directed) {
  if (!m_directed || n < 2 || n >= (u64{1} << 31) || max_weight == 0 || !(avg_degree > 0.0) ||
      !(gamma > 2.0))
    return KMST_EINVAL;
  try {
    return emit(rhg(n, avg_degree, gamma, seed), max_weight, seed, src, dst, weight, capacity,
                m_directed);
  } catch (const std::bad_alloc&) {
    return KMST_ENOMEM;
  }
}

}  // extern "C"
