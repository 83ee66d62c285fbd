// Host-side code under the sanitizers: the library's host-only entry points
// (generators, difference encoder, argument checks) and the oracle.  Built
// with AddressSanitizer + UndefinedBehaviorSanitizer, or with ThreadSanitizer
// (mode "threads": the threaded host generators, the oracle on one thread),
// by tests/test_sanitizers.py.  Exits non-zero on a failed check; the
// sanitizers abort on their own findings.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <tuple>
#include <vector>

#include "kmst.h"

struct WEdge {
  uint64_t src, dst;
  uint32_t weight, pad;
};

extern "C" {
int orc_kruskal(const WEdge* in, uint64_t N, WEdge* out, uint64_t* out_n, uint64_t* total);
int orc_boruvka(const WEdge* in, uint64_t N, int local_preprocessing, int threads, WEdge* out,
                uint64_t* out_n, uint64_t* total);
int orc_filter_boruvka(const WEdge* in, uint64_t N, int local_preprocessing,
                       uint64_t filter_threshold, int threads, WEdge* out, uint64_t* out_n,
                       uint64_t* total);
void orc_wedges_from_soa(const uint32_t* src, const uint32_t* dst, const uint32_t* weight,
                         uint64_t N, WEdge* out);
}

namespace {

int failures = 0;

void expect(bool ok, const std::string& what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what.c_str());
    ++failures;
  }
}

struct Graph {
  std::vector<uint32_t> s, d, w;
};

// sorted by (src, dst), no self-loops, every record's reverse twin present
void check_records(const Graph& g, uint64_t n, const std::string& name) {
  const size_t m = g.s.size();
  bool sorted = true, ids = true, loops = false;
  for (size_t i = 0; i < m; ++i) {
    ids &= g.s[i] < n && g.d[i] < n;
    loops |= g.s[i] == g.d[i];
    if (i && std::make_pair(g.s[i - 1], g.d[i - 1]) > std::make_pair(g.s[i], g.d[i])) sorted = false;
  }
  std::vector<std::tuple<uint32_t, uint32_t, uint32_t>> fwd, rev;
  for (size_t i = 0; i < m; ++i) {
    if (g.s[i] < g.d[i]) fwd.emplace_back(g.s[i], g.d[i], g.w[i]);
    else rev.emplace_back(g.d[i], g.s[i], g.w[i]);
  }
  std::sort(fwd.begin(), fwd.end());
  std::sort(rev.begin(), rev.end());
  expect(sorted, name + ": records sorted by (src, dst)");
  expect(ids && !loops, name + ": ids in range, no self-loops");
  expect(fwd == rev, name + ": symmetric");
}

std::vector<WEdge> wedges(const Graph& g) {
  std::vector<WEdge> r(g.s.size());
  orc_wedges_from_soa(g.s.data(), g.d.data(), g.w.data(), g.s.size(), r.data());
  return r;
}

// Kruskal, both Boruvka forms and filter-Boruvka at several thresholds agree
void check_oracle(const std::vector<WEdge>& rec, int threads, const std::string& name) {
  const uint64_t N = rec.size();
  std::vector<WEdge> a(N / 2 + 1), b(N / 2 + 1);
  uint64_t na = 0, nb = 0, ta = 0, tb = 0;
  expect(orc_kruskal(rec.data(), N, a.data(), &na, &ta) == 0, name + ": kruskal");
  auto same = [&](const std::string& what) {
    bool eq = na == nb && ta == tb;
    for (uint64_t i = 0; eq && i < na; ++i)
      eq = a[i].src == b[i].src && a[i].dst == b[i].dst && a[i].weight == b[i].weight;
    expect(eq, name + ": " + what + " == kruskal");
  };
  for (int lp = 0; lp <= 1; ++lp) {
    expect(orc_boruvka(rec.data(), N, lp, threads, b.data(), &nb, &tb) == 0, name + ": boruvka");
    same("boruvka lp=" + std::to_string(lp));
  }
  for (uint64_t thr : {0, 1, 2, 4, 8}) {
    expect(orc_filter_boruvka(rec.data(), N, 0, thr, threads, b.data(), &nb, &tb) == 0,
           name + ": filter_boruvka");
    same("filter_boruvka threshold " + std::to_string(thr));
  }
}

Graph gnm(uint64_t n, uint64_t m, uint32_t mw, uint64_t seed) {
  Graph g;
  g.s.resize(2 * m);
  g.d.resize(2 * m);
  g.w.resize(2 * m);
  expect(kmst_gen_gnm_host(n, m, mw, seed, g.s.data(), g.d.data(), g.w.data()) == KMST_OK,
         "gnm host");
  return g;
}

template <typename Gen>
Graph sized(Gen gen, const std::string& name) {
  uint64_t need = 0;
  Graph g;
  // a capacity too small writes nothing and reports the count
  expect(gen(nullptr, nullptr, nullptr, 0, &need) == KMST_EINVAL, name + ": capacity refusal");
  g.s.resize(need);
  g.d.resize(need);
  g.w.resize(need);
  uint64_t got = 0;
  expect(gen(g.s.data(), g.d.data(), g.w.data(), need, &got) == KMST_OK && got == need,
         name + ": generation");
  return g;
}

// 7-bit varints of the difference encoding, read back
uint32_t varint(const uint8_t* b, uint64_t& p) {
  uint32_t x = 0;
  for (int shift = 0;; shift += 7) {
    const uint8_t c = b[p++];
    x |= uint32_t(c & 0x7f) << shift;
    if (!(c & 0x80)) return x;
  }
}

void check_codec(const Graph& g, uint32_t chunk) {
  const uint64_t m = g.s.size();
  uint64_t used = 0;
  expect(kmst_encode_host(g.s.data(), g.d.data(), g.w.data(), m, chunk, nullptr, 0, nullptr,
                          nullptr, &used) == KMST_OK,
         "encode size query");
  const uint64_t chunks = (m + chunk - 1) / chunk;
  std::vector<uint8_t> bytes(used + 1);  // never empty: a null array asks for the size
  std::vector<uint64_t> off(chunks + 1);
  std::vector<uint32_t> first(3 * chunks);
  uint64_t used2 = 0;
  expect(kmst_encode_host(g.s.data(), g.d.data(), g.w.data(), m, chunk, bytes.data(), used,
                          off.data(), first.data(), &used2) == KMST_OK &&
             used2 == used,
         "encode");
  bool ok = off[chunks] == used;
  for (uint64_t k = 0; ok && k < chunks; ++k) {
    const uint64_t i0 = k * chunk, i1 = std::min<uint64_t>(m, i0 + chunk);
    uint32_t ps = first[3 * k];
    ok = ps == g.s[i0] && first[3 * k + 1] == g.d[i0] && first[3 * k + 2] == g.w[i0];
    uint64_t p = off[k];
    for (uint64_t i = i0 + 1; ok && i < i1; ++i) {
      const uint32_t s = ps + varint(bytes.data(), p);
      const uint32_t d = varint(bytes.data(), p);
      const uint32_t w = varint(bytes.data(), p);
      ok = s == g.s[i] && d == g.d[i] && w == g.w[i];
      ps = s;
    }
    ok = ok && p == off[k + 1];
  }
  expect(ok, "difference encoding round trip, chunk " + std::to_string(chunk));
}

void check_argument_errors() {
  kmst_config cfg;
  kmst_config_default(&cfg);
  expect(cfg.filter_threshold == 4, "default filter threshold");
  kmst_tuning t;
  kmst_tuning_default(&t);
  expect(t.sort_threads == 0 && t.balanced_digits == 1, "default tuning");
  uint64_t n = 0, w = 0;
  kmst_wedge out[4];
  expect(kmst_boruvka(nullptr, 2, nullptr, out, &n, &w) == KMST_EINVAL, "null config refused");
  expect(std::strlen(kmst_last_error()) > 0, "last error names the refusal");
  uint64_t used = 0;
  expect(kmst_encode_host(nullptr, nullptr, nullptr, 4, 16, nullptr, 0, nullptr, nullptr, &used) ==
             KMST_EINVAL,
         "encode refuses null arrays");
  uint64_t need = 0;
  expect(kmst_gen_rhg_host(1024, 8.0, 1.5, 254, 1, nullptr, nullptr, nullptr, 0, &need) ==
             KMST_EINVAL,
         "RHG refuses gamma <= 2");
}

// oracle error statuses on broken inputs (unsorted, asymmetric)
void check_oracle_errors(const Graph& g) {
  std::vector<WEdge> rec = wedges(g);
  std::vector<WEdge> out(rec.size());
  uint64_t no = 0, t = 0;
  std::vector<WEdge> bad = rec;
  std::swap(bad[10], bad[11]);
  bad[10].weight ^= 1;
  expect(orc_kruskal(bad.data(), bad.size(), out.data(), &no, &t) != 0,
         "oracle refuses a broken input");
}

}  // namespace

int main(int argc, char** argv) {
  const bool threads_mode = argc > 1 && std::strcmp(argv[1], "threads") == 0;
  const int oracle_threads = threads_mode ? 1 : 4;
  check_argument_errors();

  Graph g1 = gnm(1 << 10, 1 << 13, 254, 7);
  check_records(g1, 1 << 10, "gnm");
  check_oracle(wedges(g1), oracle_threads, "gnm");
  Graph g2 = gnm(70000, 1 << 17, (1u << 31) - 1, 8);  // 31-bit weights
  check_records(g2, 70000, "gnm wide");
  check_oracle(wedges(g2), oracle_threads, "gnm wide");
  check_oracle_errors(g1);

  Graph r = sized(
      [](uint32_t* s, uint32_t* d, uint32_t* w, uint64_t cap, uint64_t* m) {
        return kmst_gen_rgg2d_host(1 << 14, 16.0, (1u << 31) - 1, 3, s, d, w, cap, m);
      },
      "rgg2d");
  check_records(r, 1 << 14, "rgg2d");
  Graph h = sized(
      [](uint32_t* s, uint32_t* d, uint32_t* w, uint64_t cap, uint64_t* m) {
        return kmst_gen_rhg_host(1 << 14, 32.0, 2.8, 254, 4, s, d, w, cap, m);
      },
      "rhg");
  check_records(h, 1 << 14, "rhg");
  if (!threads_mode) {
    check_oracle(wedges(r), oracle_threads, "rgg2d");
    check_oracle(wedges(h), oracle_threads, "rhg");
  }
  for (uint32_t chunk : {1u, 7u, 64u, 4096u}) check_codec(g1, chunk);
  check_codec(g2, 256);

  if (failures) {
    std::fprintf(stderr, "%d check(s) failed\n", failures);
    return 1;
  }
  std::printf("ok (%s)\n", threads_mode ? "threads" : "address+undefined");
  return 0;
}
