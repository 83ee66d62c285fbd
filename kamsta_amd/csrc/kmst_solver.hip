// kmst_solver.hip -- host orchestration of the Boruvka / filter-Boruvka round
// loop on one MI355X and the C ABI declared in include/kmst.h.
//
// Plain Boruvka (hybridMST::boruvka, include/algorithms/hybrid_boruvka.hpp:263-301):
//   setup    validate contract, CSR offsets, strict keys   (UncompressedGraph ids)
//   round    K1 segmented argmin       (MinimumEdgeOpenMP, compute_min_edges)
//            K2 hook + MST append      (IsRepresentative / determine_mst_edges)
//            K3 label doubling         (ComputeRepresentative / update_parents)
//            renumber roots            (MakeVerticesConsecutive)
//            K4+K5 relabel + compact   (EdgeRenamer + remove_self_loops)
//            K6 radix sort (src, dst)  (RedistributeViaSelectedSorter)
//            K7 keep lightest copy     (EdgeProcessor::remove_duplicates)
//   finish   K9 gather canonical edges (GetMstEdge)
// Filter-Boruvka (hybridMST::filter_boruvka,
// include/algorithms/hybrid_boruvka_edgefilter.hpp:227-330) wraps the same
// rounds in the pivot recursion: light/heavy stream split, recurse on light,
// relabel heavy by the parent array, drop same-component edges, recurse.
#include <cstdlib>
#include <algorithm>
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <sstream>
#include <vector>

#include "kmst_kernels.hpp"
#include "kmst_solver.hpp"

namespace kmst {

void gen_gnm_device(kmst_context& C, u64 n, u64 m, u32 max_w, u64 seed, u32* out_s, u32* out_d,
                    u32* out_w);
void gen_gnm_host(u64 n, u64 m, u32 max_w, u64 seed, u32* out_s, u32* out_d, u32* out_w);

thread_local std::string g_last_error;
thread_local KernelTimer* g_ktimer = nullptr;

// ------------------------------------------------------------ kernel timer
void KernelTimer::begin(hipStream_t st, int kind, double bytes) {
  if (depth > 0) {  // inside an open scope: its launches count for that scope
    ++depth;
    return;
  }
  if (used == pool.size()) {
    hipEvent_t a, b;
    KMST_HIP(hipEventCreate(&a));
    KMST_HIP(hipEventCreate(&b));
    pool.push_back({a, b});
  }
  KMST_HIP(hipEventRecord(pool[used].first, st));
  marks.push_back({kind, bytes});
  depth = 1;
}
void KernelTimer::end(hipStream_t st) noexcept {
  if (depth == 0 || --depth > 0) return;
  (void)hipEventRecord(pool[used].second, st);
  ++used;
}
void KernelTimer::collect() {
  for (size_t i = 0; i < used; ++i) {
    float ms = 0.f;
    KMST_HIP(hipEventElapsedTime(&ms, pool[i].first, pool[i].second));
    auto& acc = totals[marks[i].kind];
    acc.launches += 1;
    acc.ms += ms;
    acc.bytes += marks[i].bytes;
  }
  used = 0;
  depth = 0;
  marks.clear();
}
KernelTimer::~KernelTimer() {
  for (auto& e : pool) {
    (void)hipEventDestroy(e.first);
    (void)hipEventDestroy(e.second);
  }
}

const char* kernel_kind_name(int k) {
  static const char* names[] = {"fill", "scan_reduce", "scan_down", "validate", "offsets",
                                "keys", "argmin", "argmin_fixup", "hook", "jump", "label",
                                "gpar", "relabel_count", "relabel_scatter", "sort_hist",
                                "sort_scatter", "dedup", "sample", "split_count",
                                "split_scatter", "gather", "gen", "persistent_rounds",
                                "other"};
  return (k >= 0 && k < K_COUNT) ? names[k] : "?";
}

// ------------------------------------------------------------ buffers
void* DevBuf::get(size_t need) {
  if (need == 0) need = 1;
  if (arena && need > bytes) {
    const size_t at = (arena->used + 255) & ~size_t{255};
    if (arena->frozen || at > arena->size || need > arena->size - at)
      throw Error(KMST_EINVAL, "a workspace of " + std::to_string(need) +
                                   " bytes exceeds the fixed scratch");
    p = arena->base + at;
    bytes = need;
    arena->used = at + need;
    return p;
  }
  if (need > bytes) {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    // grow with headroom so that a stream of slightly larger requests does
    // not reallocate (hipFree synchronises the device) on every call
    if (need < (size_t{1} << 40)) need += need / 4;
    if (hipMalloc(&p, need) != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      throw Error(KMST_ENOMEM, "hipMalloc of " + std::to_string(need) + " bytes failed");
    }
    bytes = need;
  }
  return p;
}
void DevBuf::release() {
  if (arena) return;  // carved storage is never freed
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

}  // namespace kmst

using namespace kmst;

// ------------------------------------------------------------ context
kmst_context::kmst_context(int dev) : device(dev) {
  KMST_HIP(hipSetDevice(device));
  KMST_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  // coherent pinned words, which the device writes directly; where the
  // system offers none (or KMST_READBACK_COPY=1 asks for it), default pinned
  // words read back by copies
  const char* copy_env = std::getenv("KMST_READBACK_COPY");
  const bool publish = !(copy_env && copy_env[0] == '1');
  if (publish && hipHostMalloc(reinterpret_cast<void**>(&host_counters), HOST_COUNTER_BYTES,
                               hipHostMallocCoherent) == hipSuccess) {
    host_counters[HOST_SEQ_WORD] = 0;
    host_mapped = true;
  } else {
    (void)hipGetLastError();
    host_counters = nullptr;
    KMST_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_counters), HOST_COUNTER_BYTES,
                           hipHostMallocDefault));
  }
}
kmst_context::kmst_context(int dev, kmst::u32* host_words, kmst::Arena* arena)
    : device(dev), owns_resources(false), host_counters(host_words) {
  for (DevBuf* b : all_bufs()) b->arena = arena;
}
kmst_context::~kmst_context() {
  if (!owns_resources) return;
  (void)hipSetDevice(device);
  (void)hipStreamSynchronize(stream);
  for (DevBuf* b : all_bufs()) b->release();
  if (host_counters) (void)hipHostFree(host_counters);
  (void)hipStreamDestroy(stream);
}
std::vector<DevBuf*> kmst_context::all_bufs() {
  return {&xs,       &xd,         &xv,      &ys,        &yd,    &yv,      &val0,
          &off,      &best_val,   &best_dst, &parent,   &flag,  &newid,   &label,
          &rep,      &rep_next,   &gpar,    &seg_start, &seg_end, &mst,   &pieces,
          &scan_scratch, &hist,  &counters,  &lookback, &in_s, &in_d,
          &in_w,     &out_s,      &out_d,   &out_w,     &samples, &gen_a, &gen_b,
          &gen_c,    &gen_d,      &tune,    &codec_bytes, &codec_off, &codec_first,
          &persist_log, &part_giant};
}

namespace kmst {

// Device counter slots (u32 words; u64 values at even word offsets >= 16).
enum Ctr : int {
  C_ERR = 0, C_FWD = 1, C_PIECES = 3, C_NNEXT = 4, C_MNEXT = 5, C_LIGHT = 6, C_MAXW = 7,
  C_MSTLIST = 8, C_MSTLONG = 9,  // output order: listed long groups, a group too long
  C_RUNS = 10,                    // run starts among the next round's sampled records
  C_ROOTS = 11,                   // parent-array roots with records (after a flatten)
  C_TOTAL64 = 16, C_HSUM64 = 18, C_ROUND64 = 22,
  C_CLASS64 = 24,   // MAX_CLASSES u64 class counts (words 24..41)
  C_OUTSIDE = 60    // outside_list: count (60), u64 total range length (62..63)
};

// A range of directed records; keys are stored in v, or (the read-only input
// in packed-key mode, v == nullptr) computed from the weights w.
struct RecView {
  u32* s = nullptr;
  u32* d = nullptr;
  u64* v = nullptr;
  u64 m = 0;
  const u32* w = nullptr;
  RecView at(u64 off, u64 cnt) const { return RecView{s + off, d + off, v + off, cnt, nullptr}; }
};

class Solver {
 public:
  // part (count + 1 device words, optional): ranges [part[r], part[r + 1]) of
  // vertex ids that no record joins
  Solver(kmst_context& c, const u32* is, const u32* id, const u32* iw, u64 m, u64 n,
         const u32* part_ranges = nullptr, u32 part_ranges_count = 0)
      : C(c), st(c.stream), in_s(is), in_d(id), in_w(iw), m0(m), n0(static_cast<u32>(n)),
        part(part_ranges), part_count(part_ranges_count) {}

  void run(const kmst_config& cfg, u32* out_s, u32* out_d, u32* out_w, u64 out_cap, u64* out_n,
           u64* total);
  // sizes every workspace for solves of up to m0 records and n0 vertices
  void reserve() {
    allocate();
    C.val0.get((m0 ? m0 : 1) * 8);
  }

 private:
  kmst_context& C;
  hipStream_t st;
  const u32 *in_s, *in_d, *in_w;
  u64 m0;
  u32 n0;
  const u32* part = nullptr;
  u32 part_count = 0;
  u32* part_giant = nullptr;  // largest component of every range (part_count words)
  u32* ctr = nullptr;
  u64 filter_threshold = 4;
  bool packed = true;   // strict key packed as (w, min, max), else (w, uid)
  u32 vbits = 0;
  RecView X, Y;
  u64* val0 = nullptr;
  u64* best_val = nullptr;
  u32 *best_dst = nullptr, *parent = nullptr, *label = nullptr;
  u64 *state = nullptr, *pos = nullptr;  // per-round vertex state and its scan
  u32 *flag = nullptr, *newid = nullptr;  // u32 views of the same memory (compaction)
  u32 *rep = nullptr, *rep_next = nullptr, *gpar = nullptr;
  // input record range of every vertex (filter-Boruvka); seg_end[v] == 0
  // marks vertices without records
  u32 *seg_start = nullptr, *seg_end = nullptr;
  u64* mst = nullptr;
  u64 mst_count = 0;
  u32* scan_scratch = nullptr;
  ArgPiece* pieces = nullptr;
  u32* samples = nullptr;

  WeightPivots pivots{};  // weight classes of the input (filter-Boruvka)

  void allocate();
  // validates the input; for filter-Boruvka also marks vertices with records
  // and their record ranges and, with plan_classes, counts the weight classes
  void setup(bool filter, bool plan_classes);
  // packed keys when bits(max weight) + 2 bits(n) <= 64 (from the counters)
  void set_key_format();
  // errors and the symmetry test of a finished input check
  void check_input();
  // set by setup when the class scatter validates the input
  bool check_in_scatter = false;
  // most sampled input records join ids FAR_SPAN or more apart (random
  // graphs): their relabels gather all over the label table
  bool far_input = false;
  void plan_pivots();
  void filter_classes();
  void check_errors(const char* where);
  void read_counters();
  void fetch(const u32* extra, u32 words);
  u64 counter64(int word) const {
    u64 x;
    std::memcpy(&x, C.host_counters + word, 8);
    return x;
  }
  // rounds of plain Boruvka on records grouped by src with ids < n_start
  // (default n0); A is writable scratch disjoint from recs, B writable and may
  // alias recs.  With track_gpar every round updates the parent array.
  void rounds(RecView recs, RecView A, RecView B, bool track_gpar, u64 n_start = 0,
              bool rep_given = false);
  // true when the remaining rounds of a graph with m live records and ids
  // below n run as one persistent launch
  bool use_persistent(u64 m, u64 n) const;
  bool skip_sort(u64 m, u64 n, u32 far_starts) const;
  // every remaining round of `cur` in the persistent kernel (same contract
  // as one pass of the rounds loop, run to the end)
  void persistent(RecView cur, RecView A, RecView B, u64 n, bool track_gpar, bool first);
  u64* persist_log = nullptr;
  // base case: compact sparse vertex ids first, then run the rounds
  void base_case(RecView P, RecView Q);
  void filter_rec(RecView P, RecView Q, bool stop, int depth);
  // heavy may be the read-only input; scratch and alt are writable, disjoint
  // from each other and from heavy, each with room for heavy.m records
  // With from_input, heavy is the read-only input (sorted by (src, dst)) and
  // the records come from the ranges of the vertices outside the largest
  // component when those are few.
  void filter_heavy(RecView heavy, RecView scratch, RecView alt, int depth, u64 key_min = 0,
                    u64 key_max = ~0ull, bool from_input = false);
  // true when the parent array has at most one root with edges; also picks
  // the most frequent root of a sample of vertices (giant_root) and its
  // sample share
  bool one_component();
  // ctr + C_ROOTS holds the parent array's roots with records: the last
  // rounds flattened it and nothing changed it since
  bool roots_counted = false;
  u32 giant_root = 0;
  double giant_share = 0.0;
  u32 pick_pivot(RecView r);
  u64 partition(RecView in, RecView out, u32 pivot, bool write_heavy = true);
  // keys of r: stored, packed from the weights, or (positional mode, r the
  // input without stored keys) made by the kernels that read the input, with
  // the pair_position search in the validated record ranges
  KeySrc keys_of(const RecView& r) const {
    KeySrc k{r.v, r.w, vbits};
    if (!r.v && !packed) {
      k.twin_lo = seg_start;
      k.twin_hi = seg_end;
    }
    return k;
  }
  // positional keys of the whole input (val0), built once when a relabel
  // reads the input
  u64* input_keys();
  // positional keys not stored for the input: the class scatter and the
  // outside gather make them for the records they write
  bool lazy_keys = false;
  // records grouped by src are all the argmin needs; where parallel records
  // abound, sort by (src, dst) and keep only the lightest of each pair
  RecView sort_dedup(RecView cur, RecView alt, u32 bits);
  RecView sort_by_src(RecView cur, RecView alt, u32 bits);
};

void Solver::allocate() {
  const u64 m = m0 ? m0 : 1;
  const u64 n = u64(n0) + 1;
  X = RecView{static_cast<u32*>(C.xs.get(m * 4)), static_cast<u32*>(C.xd.get(m * 4)),
              static_cast<u64*>(C.xv.get(m * 8)), m};
  Y = RecView{static_cast<u32*>(C.ys.get(m * 4)), static_cast<u32*>(C.yd.get(m * 4)),
              static_cast<u64*>(C.yv.get(m * 8)), m};
  best_val = static_cast<u64*>(C.best_val.get(n * 8));
  best_dst = static_cast<u32*>(C.best_dst.get(n * 4));
  parent = static_cast<u32*>(C.parent.get(n * 4));
  state = static_cast<u64*>(C.flag.get(n * 8));
  pos = static_cast<u64*>(C.newid.get(n * 8));
  flag = reinterpret_cast<u32*>(state);
  newid = reinterpret_cast<u32*>(pos);
  label = static_cast<u32*>(C.label.get(n * 4));
  rep = static_cast<u32*>(C.rep.get(n * 4));
  rep_next = static_cast<u32*>(C.rep_next.get(n * 4));
  gpar = static_cast<u32*>(C.gpar.get(n * 4));
  seg_start = static_cast<u32*>(C.seg_start.get(n * 4));
  seg_end = static_cast<u32*>(C.seg_end.get(n * 4));
  mst = static_cast<u64*>(C.mst.get(n * 8));
  const u64 arg_tiles = (m + ARG_TILE - 1) / ARG_TILE;
  pieces = static_cast<ArgPiece*>(C.pieces.get(2 * arg_tiles * sizeof(ArgPiece) + 64));
  C.lb_work.bind(st, static_cast<u64*>(C.lookback.get(lookback_elems(m) * 8)), m);
  const u64 sort_m = std::max(m, n);
  C.sort_work.bind(st, static_cast<u32*>(C.hist.get(sort_work_words(sort_m) * 4)), sort_m);
  apply_tuning(C);
  const u64 scan_max = std::max<u64>({n, (m + CMP_TILE - 1) / CMP_TILE + 1, record_units(m)});
  // u64 scans reuse the same scratch, so size it in u64 units
  scan_scratch = static_cast<u32*>(C.scan_scratch.get(scan_scratch_elems(scan_max) * 8 + 64));
  ctr = static_cast<u32*>(C.counters.get(256));
  samples = static_cast<u32*>(C.samples.get(4096 * 4));
  persist_log = static_cast<u64*>(
      C.persist_log.get(PERSIST_CTL_WORDS * 4 + PERSIST_LOG_TOTAL * 8 + PERSIST_GROUP_WORDS * 4));
  part_giant = part_count ? static_cast<u32*>(C.part_giant.get(u64(part_count) * 4 + 4)) : nullptr;
}

void apply_tuning(kmst_context& C) {
  C.sort_work.threads = C.tuning.sort_threads;
  C.sort_work.balanced = C.tuning.balanced_digits != 0;
  C.sort_work.wide_min = C.tuning.sort_wide_min;
}

void Solver::read_counters() { fetch(nullptr, 0); }

// The counters, and `words` words of `extra` into the staging words, on the
// host once all work queued so far has finished.  Contexts with coherent
// pinned words publish them with one small kernel and spin on its sequence
// word (about 10 us less idle GPU per read-back than a copy and a stream
// synchronisation); the stream is queried now and then, so that a failed
// launch or a drained stream without the publication ends the wait.
void Solver::fetch(const u32* extra, u32 words) {
  if (!C.host_mapped) {
    KMST_HIP(hipMemcpyAsync(C.host_counters, ctr, 256, hipMemcpyDeviceToHost, st));
    if (words)
      KMST_HIP(hipMemcpyAsync(C.host_counters + HOST_SAMPLE_WORD, extra, words * 4ull,
                              hipMemcpyDeviceToHost, st));
    KMST_HIP(hipStreamSynchronize(st));
    return;
  }
  const u32 seq = ++C.publish_seq;
  publish_words(st, ctr, 64, extra, words, C.host_counters, HOST_SAMPLE_WORD, HOST_SEQ_WORD, seq);
  u32* flag = C.host_counters + HOST_SEQ_WORD;
  for (u32 spins = 1;; ++spins) {
    if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == seq) return;
#if defined(__x86_64__)
    __builtin_ia32_pause();
#endif
    if ((spins & 1023) == 0) {
      const hipError_t q = hipStreamQuery(st);
      if (q == hipErrorNotReady) continue;
      if (q != hipSuccess)
        throw Error(KMST_EDEVICE, std::string("read-back: ") + hipGetErrorString(q));
      // the stream has drained: the publication is complete and visible
      if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == seq) return;
      throw Error(KMST_EDEVICE, "read-back: published counters not visible to the host");
    }
  }
}

void Solver::check_errors(const char* where) {
  read_counters();
  const u32 e = C.host_counters[C_ERR];
  if (e & ERR_SPIN) throw Error(KMST_EDEVICE, std::string(where) + ": bounded wait expired");
  if (e & ERR_BOUNDS) throw Error(KMST_EDEVICE, std::string(where) + ": output position out of range");
  if (e & ERR_INVALID)
    throw Error(KMST_EINVAL, std::string(where) + ": self-loop or vertex id out of range");
  if (e & ERR_UNSORTED) throw Error(KMST_EUNSORTED, std::string(where) + ": records not sorted by (src, dst)");
  if (e & ERR_ASYMMETRIC)
    throw Error(KMST_EASYMMETRIC, std::string(where) + ": a record has no reverse twin");
}

void Solver::setup(bool filter, bool plan_classes) {
  fill_u32(st, ctr, 64, 0);
  pivots = WeightPivots{};
  check_in_scatter = false;
  if (plan_classes) plan_pivots();
  // init_parent_array (hybrid_boruvka_edgefilter.hpp:332-352): vertices with
  // records, marked during validation
  if (filter) fill_u32(st, seg_end, n0, 0);
  vbits = bits_for(n0);
  u64* class64 = reinterpret_cast<u64*>(ctr + C_CLASS64);
  u64* hsum = reinterpret_cast<u64*>(ctr + C_HSUM64);
  if (plan_classes && pivots.K > 0) {
    // weights-only pass: the largest weight (the key format), the class
    // counts and the per-unit counts of the two lightest classes (into X.s /
    // X.d, free until the class scatter)
    const u32 q0 = pivots.q[0], q1 = pivots.K >= 2 ? pivots.q[1] : q0;
    weight_classes(st, in_w, m0, pivots, q0, q1, class64, ctr + C_MAXW, X.s, X.d);
    read_counters();
    set_key_format();
    if (packed) {
      // the class scatter reads every record anyway: it checks the input too
      check_in_scatter = true;
      return;
    }
    // positional keys: the class scatter validates the input and makes the
    // keys of the records it writes, by a pair_position search in the record ranges
    // (clamped to the arrays while the input is unchecked), recorded first
    record_ranges(st, in_s, m0, n0, seg_start, seg_end);
    check_in_scatter = true;
    lazy_keys = true;
    C.stats.lazy_keys = true;
    return;
  } else {
    validate_records(st, in_s, in_d, in_w, m0, n0, ctr + C_ERR, ctr + C_FWD, ctr + C_MAXW, hsum,
                     pivots, class64, filter ? seg_start : nullptr, filter ? seg_end : nullptr);
    check_input();
    set_key_format();
  }
  if (!packed) {
    val0 = static_cast<u64*>(C.val0.get(m0 * 8));
    build_offsets(st, in_s, m0, n0, label /* reused as offsets, n0+1 */);
    build_keys(st, in_s, in_d, in_w, m0, label, label + 1, val0, ctr + C_ERR);
    check_errors("input");
  }
}

void Solver::set_key_format() {
  C.stats.max_weight = C.host_counters[C_MAXW];
  const u32 wbits = bits_for(u64(C.host_counters[C_MAXW]) + 1);
  packed = wbits + 2 * vbits <= 64;
  C.stats.packed_keys = packed;
}

void Solver::check_input() {
  check_errors("input");
  if (u64(C.host_counters[C_FWD]) * 2 != m0 || counter64(C_HSUM64) != counter64(C_HSUM64 + 2))
    throw Error(KMST_EASYMMETRIC, "input: records are not symmetric (src,dst,w) <-> (dst,src,w)");
}

RecView Solver::sort_by_src(RecView cur, RecView alt, u32 bits) {
  if (cur.m > 1 && bits > 0 &&
      radix_sort(st, cur.s, cur.d, cur.v, alt.s, alt.d, alt.v, cur.m, 0, bits, C.sort_work,
                 ctr + C_ERR))
    cur = RecView{alt.s, alt.d, alt.v, cur.m};
  return cur;
}

RecView Solver::sort_dedup(RecView cur, RecView alt, u32 bits) {
  if (cur.m > 1 && bits > 0) {
    // the last scatter pass keeps the lightest copy of each (src, dst) run
    if (radix_sort(st, cur.s, cur.d, cur.v, alt.s, alt.d, alt.v, cur.m, bits, 2 * bits,
                   C.sort_work, ctr + C_ERR, /*dedup_last=*/true))
      cur = RecView{alt.s, alt.d, alt.v, cur.m};
  } else {
    dedup_mark(st, cur.s, cur.d, cur.v, cur.m);
  }
  return cur;
}

void Solver::rounds(RecView recs, RecView A, RecView B, bool track_gpar, u64 n_start,
                    bool rep_given) {
  RecView cur = recs;
  u64 m = recs.m;
  u64 n = n_start ? n_start : n0;  // by default the first round runs in the original id space
  bool first = !rep_given;
  bool deduped = false;  // cur holds at most n^2 live records, the rest marked dead
  bool grouped = true;   // cur is grouped by src (the input, or a sort's output)
  while (m > 0 && n > 1) {
    if (use_persistent(deduped ? std::min(m, n * n) : m, n)) {
      persistent(cur, A, B, n, track_gpar, first);
      break;
    }
    RoundStat rs;
    rs.n = n;
    rs.m = m;
    if (grouped)
      segmented_argmin(st, cur.s, cur.d, keys_of(cur), m, n, best_val, best_dst, pieces,
                       ctr + C_PIECES);
    else
      unordered_argmin(st, cur.s, cur.d, keys_of(cur), m, n, best_val, best_dst);
    hook(st, n, best_val, best_dst, parent, state);
    // the new ids of the roots: tagged into the roots' own parent entries
    // (below 2^31 vertices), else into best_dst, free after the hook
    const bool tagged = n <= ROOT_TAG;
    scan_exclusive_u64(st, state, pos, n, reinterpret_cast<u64*>(scan_scratch),
                       reinterpret_cast<u64*>(ctr + C_ROUND64), tagged ? parent : best_dst,
                       tagged);
    jump_and_label(st, n, parent, tagged ? nullptr : best_dst, state, pos, best_val, label, mst,
                   mst_count,
                   (track_gpar && !first) ? rep : nullptr, track_gpar ? gpar : nullptr,
                   rep_next, ctr + C_ERR);
    if (track_gpar) std::swap(rep, rep_next);
    RecView out = (cur.s == A.s) ? B : A;
    // a first round of far-apart ids on a label table past an L2 relabels
    // only the src < dst half and writes both directions (half the random
    // gathers); its output is no longer grouped, which costs nothing when the
    // next round sorts -- few records per vertex, where it does
    const bool twins = first && grouped && far_input && n > TWINS_MIN_N && m < UNORDERED_DEGREE * n;
    relabel_compact(st, cur.s, cur.d, keys_of(cur), 0, m, label, out.s, out.d, out.v, C.lb_work,
                    ctr + C_MNEXT, ctr + C_ERR, nullptr, 0, ~0ull, twins);
    sample_run_starts(st, out.s, ctr + C_MNEXT, ctr + C_RUNS);
    check_errors("round");
    if (g_ktimer) g_ktimer->adjust_last(K_RELABEL_SCATTER, 16.0 * C.host_counters[C_MNEXT]);
    const u64 round_total = counter64(C_ROUND64);
    const u64 n_next = round_total & 0xFFFFFFFFull;
    mst_count += round_total >> 32;
    const u64 m_next = C.host_counters[C_MNEXT];
    rs.n_next = n_next;
    rs.m_next = m_next;
    rs.twins = twins;
    rs.run_starts = m_next ? (C.host_counters[C_RUNS] & 0xFFFFu) : 0;
    rs.far_starts = m_next ? (C.host_counters[C_RUNS] >> 16) : 0;
    C.stats.rounds.push_back(rs);
    m = m_next;
    n = n_next;
    cur = RecView{out.s, out.d, out.v, m};
    if (m == 0 || n <= 1) break;
    RecView alt = (cur.s == A.s) ? B : A;
    // rounds with many records per vertex, or whose runs of equal source
    // mostly continue near the previous run, skip the sort and take their
    // minima by atomics over the compaction order; a sort by (src, dst) that
    // keeps the lightest parallel record pays only where it removes most
    // records (a multigraph with m near n^2 records keeps about 2/3 of them)
    // (from 2^28 vertices on, 8 n^2 exceeds every 32-bit record count)
    deduped = n < (u64{1} << 28) && m > 8 * n * n;
    if (deduped) {
      cur = sort_dedup(cur, alt, bits_for(n));
      grouped = true;
    } else if (skip_sort(m, n, rs.far_starts)) {
      grouped = false;  // the relabel kept the order: no sort
    } else {
      cur = sort_by_src(cur, alt, bits_for(n));
      grouped = true;
    }
    first = false;
  }
  if (track_gpar) {
    // flattened, the parent array's roots with records are counted in the
    // same pass (the one-component test before a heavy class reads them)
    pointer_jump(st, n0, gpar, ctr + C_ERR, seg_end, ctr + C_ROOTS);
    roots_counted = true;
  }
}

// A round on records in compaction order skips the sort when it has at
// least UNORDERED_DEGREE records per vertex (plain Boruvka at GNM: runs of
// about 32 records of one original vertex), or at least
// UNORDERED_MIN_DEGREE and its atomic minima mostly stay near one another:
// at most 2/5 of the sampled records start a run in another 128-byte line of
// the minima (geometric and hyperbolic graphs, whose neighbouring vertices
// merge into neighbouring ids), or the minima fit in an XCD's L2.  A GNM
// base case after its first round (more than half of the runs start far
// away, on 4·10^6 vertices) sorts.
bool Solver::skip_sort(u64 m, u64 n, u32 far_starts) const {
  if (m >= UNORDERED_DEGREE * n) return true;
  if (m < UNORDERED_MIN_DEGREE * n) return false;
  return n <= UNORDERED_SMALL_N || u64(far_starts) * 5 <= u64(RUN_SAMPLES) * 2;
}

bool Solver::use_persistent(u64 m, u64 n) const {
  return C.tuning.persistent_max != 0 && m <= C.tuning.persistent_max && n <= 2 * m + 1024 &&
         (m <= PERSIST_MAX_DEGREE * n || m <= PERSIST_SMALL) &&
         persistent_blocks_max(C.device) != 0;
}

void Solver::persistent(RecView cur, RecView A, RecView B, u64 n, bool track_gpar, bool first) {
  // outputs alternate between the two buffers that are not `cur` (B may be
  // the records of the first round, which are consumed by then)
  const RecView o0 = (cur.s == A.s) ? B : A;
  const RecView o1 = (o0.s == A.s) ? B : A;
  PersistArgs a;
  a.s0 = cur.s;
  a.d0 = cur.d;
  a.k0 = keys_of(cur);
  a.m0 = cur.m;
  a.s[0] = o0.s;
  a.d[0] = o0.d;
  a.v[0] = o0.v;
  a.s[1] = o1.s;
  a.d[1] = o1.d;
  a.v[1] = o1.v;
  a.cap = cur.m;
  a.n = n;
  a.best = best_val;
  a.bdst = best_dst;
  a.parent = parent;
  a.mst = mst;
  a.mst_base = mst_count;
  a.mst_cap = u64(n0) + 1;
  a.rep = (track_gpar && !first) ? rep : nullptr;
  a.gpar = track_gpar ? gpar : nullptr;
  a.ctl = reinterpret_cast<u32*>(persist_log);
  a.log = persist_log + PERSIST_CTL_WORDS / 2;
  a.groups = reinterpret_cast<u32*>(a.log + PERSIST_LOG_TOTAL);
  a.err = ctr + C_ERR;
  // about four records or vertices per thread (one chunk per wave), at most
  // one workgroup per CU: fewer workgroups make each grid barrier cheaper,
  // and small graphs run in one workgroup without grid barriers
  const u64 per_block = 4 * u64(persistent_threads());
  const u64 size = std::max(cur.m, n);
  const u64 want = size <= PERSIST_ONE_BLOCK ? 1 : (size + per_block - 1) / per_block;
  const u32 blocks =
      static_cast<u32>(std::max<u64>(1, std::min<u64>(want, persistent_blocks_max(C.device))));
  persistent_rounds(st, a, blocks);
  // control words and log in one copy, into the host staging words
  fetch(reinterpret_cast<const u32*>(persist_log), PERSIST_CTL_WORDS + PERSIST_LOG_TOTAL * 2);
  const u32* pc = C.host_counters + HOST_SAMPLE_WORD;
  const u64* plog = reinterpret_cast<const u64*>(pc + PERSIST_CTL_WORDS);
  if (C.host_counters[C_ERR]) {
    // which guard stopped the launch (cycle 1, rounds 2, barrier timeout 4,
    // a peer's error 8, unbound minimum 16, output full 32, error on entry
    // 64) and in which round
    std::ostringstream w;
    w << "persistent rounds (stop bits " << pc[6] << " round " << pc[7] << ", n " << n
      << ", m " << cur.m << ", blocks " << blocks << ", records per round";
    for (u32 i = 0; i < std::min<u32>(pc[5] ? pc[5] : PERSIST_MAX_ROUNDS, 8); ++i) {
      const u64* lg = plog + PERSIST_LOG_WORDS * i;
      w << " " << lg[0] << "/" << (lg[1] & 0xFFFFFFFFull) << "/" << (lg[1] >> 32);
    }
    w << ")";
    check_errors(w.str().c_str());
  }
  const u64 added = pc[4];
  const u32 r = std::min<u32>(pc[5], PERSIST_MAX_ROUNDS);
  const double tpu = persistent_ticks_per_us(C.device);
  double bytes = 0;
  for (u32 i = 0; i < r; ++i) {
    const u64* lg = plog + PERSIST_LOG_WORDS * i;
    RoundStat rs;
    rs.m = lg[0];
    rs.n = lg[1] & 0xFFFFFFFFull;
    rs.n_next = rs.n - (lg[1] >> 32);
    rs.m_next = (i + 1 < r) ? lg[PERSIST_LOG_WORDS] : 0;
    rs.persistent = true;
    for (int p = 0; p < 4; ++p) rs.phase_us[p] = double(lg[3 + p] - lg[2 + p]) / tpu;
    C.stats.rounds.push_back(rs);
    // per round: bind (16 B per record + its source's 8-byte minimum),
    // relabel (8 B of ids, two 4-byte label gathers, the 8-byte key; 16 B per
    // kept record written), hook and jump (about 30 B per vertex)
    bytes += 24.0 * double(rs.m) + 24.0 * double(rs.m) + 16.0 * double(rs.m_next) +
             30.0 * double(n);
  }
  if (g_ktimer) g_ktimer->adjust_last(K_PERSIST, bytes);
  mst_count += added;
}

void Solver::base_case(RecView P, RecView Q) {
  // a base case with few records touches few vertices: renumber them to
  // [0, n_c) (monotone, so the records stay grouped) instead of running the
  // first round over all n0 vertex slots.  P is sorted by src and symmetric
  // (every part the filter hands down is), so its vertices are its sources
  // and the runs of equal src number them, in passes over P only; Q.s holds
  // the run marks and their scan until the relabel writes Q
  if (P.m < u64(n0)) {
    run_start_flags(st, P.s, P.m, Q.s);
    scan_exclusive_u32(st, Q.s, Q.s, P.m, scan_scratch, ctr + C_NNEXT);
    number_runs(st, P.s, P.m, Q.s, newid, rep);
    relabel_compact(st, P.s, P.d, keys_of(P), 0, P.m, newid, Q.s, Q.d, Q.v, C.lb_work, ctr + C_MNEXT,
                    ctr + C_ERR);
    check_errors("compact");
    if (g_ktimer) g_ktimer->adjust_last(K_RELABEL_SCATTER, 16.0 * C.host_counters[C_MNEXT]);
    const u64 nc = C.host_counters[C_NNEXT];
    const u64 mc = C.host_counters[C_MNEXT];
    rounds(Q.at(0, mc), P.at(0, mc), Q.at(0, mc), true, nc, true);
    return;
  }
  rounds(P, Q, P, true);
}

u32 Solver::pick_pivot(RecView r) {
  const u32 k = static_cast<u32>(std::min<u64>(LEVEL_SAMPLES, r.m));
  sample_weights(st, r.v, r.w, r.m, k, packed ? 2 * vbits : 32, samples);
  // through the pinned staging words
  fetch(samples, k);
  std::vector<u32> h(C.host_counters + HOST_SAMPLE_WORD, C.host_counters + HOST_SAMPLE_WORD + k);
  h.erase(std::remove(h.begin(), h.end(), NONE32), h.end());
  if (h.empty()) return NONE32;
  std::sort(h.begin(), h.end());
  return h[h.size() / 2];  // median sampled weight (select_pivot, distributed_partitioning.hpp:92-120)
}

// All pivots of the top levels of the recursion from one sample of the input
// weights: level l splits the light part of level l-1 at its median
// (select_pivot, distributed_partitioning.hpp:92-120), so the pivots are
// sample quantiles near 1/2, 1/4, 1/8, ...  The chain stops where the
// reference would stop splitting: the light part's average degree is at most
// the filter threshold, it has at most 1000 records, or its share of the part
// exceeds 0.85 (hybrid_boruvka_edgefilter.hpp:233-240, filter_helpers.hpp:130-132).
void Solver::plan_pivots() {
  const u32 k = static_cast<u32>(std::min<u64>(PIVOT_SAMPLES, m0));
  sample_weights(st, nullptr, in_w, m0, k, 32, samples);
  // the same records' id distances decide where relabels gather far apart
  sample_far(st, in_s, in_d, m0, k, FAR_SPAN, samples + k);
  fetch(samples, k + 1);
  std::vector<u32> h(C.host_counters + HOST_SAMPLE_WORD,
                     C.host_counters + HOST_SAMPLE_WORD + k + 1);
  far_input = 2 * u64(h[k]) > k;
  h.pop_back();
  std::sort(h.begin(), h.end());
  std::vector<u32> chain;  // descending
  size_t cnt = h.size();
  double est_m = double(m0);
  while (chain.size() < MAX_PIVOTS && cnt > 0) {
    if (est_m / double(n0) <= double(filter_threshold) || est_m <= 1000.0) break;
    const u32 p = h[cnt / 2];
    const size_t light = std::upper_bound(h.begin(), h.begin() + cnt, p) - h.begin();
    chain.push_back(p);
    if (double(light) / double(cnt) > 0.85 || light == 0) break;
    est_m = est_m * double(light) / double(cnt);
    cnt = light;
  }
  pivots.K = static_cast<u32>(chain.size());
  for (u32 j = 0; j < pivots.K; ++j) pivots.q[j] = chain[pivots.K - 1 - j];
}

// The top K levels of the filter recursion in one pass over the input: the
// class scatter writes weight classes 0..K-1 (class 0 sorted) into Y, class K
// (above every pivot) stays in the input.  Then, in the recursion's order:
// class 0 recurses as the lightest part; class c = 1..K-1 is the heavy half of
// the level with pivot q[c-1]; class K is the top level's heavy half, streamed
// from the input again with key >= (q[K-1] + 1) << wshift.
void Solver::filter_classes() {
  const u32 K = pivots.K;
  u64 cnt[MAX_CLASSES] = {};
  for (u32 c = 0; c <= K; ++c) cnt[c] = counter64(C_CLASS64 + 2 * c);
  // only the two lightest classes are written out; the heavier ones are
  // taken from the input when their turn comes, by then mostly through the
  // ranges of the few vertices outside the largest component
  const u32 J = std::min<u32>(K, 2);
  const u32 q0 = pivots.q[0], q1 = J >= 2 ? pivots.q[1] : q0;
  const u64 c1 = J >= 2 ? cnt[1] : 0, written = cnt[0] + c1;
  const u32 wshift = packed ? 2 * vbits : 32;
  const RecView input{const_cast<u32*>(in_s), const_cast<u32*>(in_d), packed ? nullptr : val0, m0,
                      packed || lazy_keys ? in_w : nullptr};
  // validation left the per-unit class counts in X.s / X.d: their exclusive
  // scans place every unit's records
  const u64 units = record_units(m0);
  scan_exclusive_u32(st, X.s, X.s, units, scan_scratch, nullptr);
  scan_exclusive_u32(st, X.d, X.d, units, scan_scratch, nullptr);
  InputCheck chk;
  chk.n = n0;
  chk.fwd_count = ctr + C_FWD;
  chk.hsum = reinterpret_cast<u64*>(ctr + C_HSUM64);
  chk.seg_start = lazy_keys ? nullptr : seg_start;  // lazy keys: recorded already
  chk.seg_end = lazy_keys ? nullptr : seg_end;
  class_scatter(st, in_s, in_d, keys_of(input), m0, wshift, q0, q1, X.s, X.d, cnt[0], written,
                Y.s, Y.d, Y.v, ctr + C_ERR, check_in_scatter ? &chk : nullptr);
  if (check_in_scatter)
    check_input();
  else
    check_errors("split");
  const u64 base1 = cnt[0];
  // level l (0 = top) splits classes 0..K-l at pivot q[K-1-l]
  std::vector<FilterStat> level(K);
  u64 part = m0;
  for (u32 l = 0; l < K; ++l) {
    const u32 c = K - l;  // heavy class of level l
    level[l] = FilterStat{static_cast<int>(l), part, part - cnt[c], pivots.q[K - 1 - l], false, 0};
    C.stats.filter.push_back(level[l]);
    part -= cnt[c];
  }
  const bool stop = double(cnt[0]) / double(cnt[0] + cnt[1]) > 0.85;
  filter_rec(Y.at(0, cnt[0]), X.at(0, cnt[0]), stop, static_cast<int>(K));
  if (J >= 2 && cnt[1] > 0) {
    C.stats.filter.push_back(level[K - 1]);
    RecView cls = Y.at(base1, cnt[1]);
    filter_heavy(cls, X.at(0, cnt[1]), cls, static_cast<int>(K - 1));
  }
  // class c holds keys in [(q[c-1] + 1) << wshift, ((q[c] + 1) << wshift) - 1]
  auto key_upper = [&](u32 q) {
    return (u64(q) + 1) >= (u64{1} << (64 - wshift)) ? ~0ull : ((u64(q) + 1) << wshift) - 1;
  };
  for (u32 c = J; c <= K; ++c) {
    if (cnt[c] == 0) continue;
    C.stats.filter.push_back(level[K - c]);
    const u64 kmin = (u64(pivots.q[c - 1]) + 1) << wshift;
    const u64 kmax = c < K ? key_upper(pivots.q[c]) : ~0ull;
    filter_heavy(input, X, Y, static_cast<int>(K - c), kmin, kmax, /*from_input=*/true);
  }
}

u64* Solver::input_keys() {
  if (!val0) {
    val0 = static_cast<u64*>(C.val0.get(m0 * 8));
    build_keys(st, in_s, in_d, in_w, m0, seg_start, seg_end, val0, ctr + C_ERR);
  }
  return val0;
}

u64 Solver::partition(RecView in, RecView out, u32 pivot, bool write_heavy) {
  split_light_heavy(st, in.s, in.d, keys_of(in), in.m, pivot, packed ? 2 * vbits : 32,
                    write_heavy, out.s, out.d, out.v, C.lb_work, ctr + C_LIGHT, ctr + C_ERR);
  check_errors("split");
  const u64 L = C.host_counters[C_LIGHT];
  if (g_ktimer) g_ktimer->adjust_last(K_SPLIT_SCATTER, 16.0 * (write_heavy ? in.m : L));
  return L;
}

// Heavy half of one recursion level: relabel by the parent array, drop
// same-component edges, sort + keep the lightest copies, recurse
// (boruvka_recursive_setup, hybrid_boruvka_edgefilter.hpp:173-225).
bool Solver::one_component() {
  if (!roots_counted) count_roots(st, gpar, seg_end, n0, ctr + C_ROOTS);
  const u32 k = std::min<u32>(GPAR_SAMPLES, n0);
  sample_weights(st, nullptr, gpar, n0, k, 32, samples);
  fetch(samples, k);
  std::vector<u32> h(C.host_counters + HOST_SAMPLE_WORD, C.host_counters + HOST_SAMPLE_WORD + k);
  std::sort(h.begin(), h.end());
  size_t best = 0;
  for (size_t i = 0; i < h.size();) {
    size_t j = i;
    while (j < h.size() && h[j] == h[i]) ++j;
    if (j - i > best) {
      best = j - i;
      giant_root = h[i];
    }
    i = j;
  }
  giant_share = k ? double(best) / double(k) : 0.0;
  return C.host_counters[C_ROOTS] <= 1;
}

void Solver::filter_heavy(RecView heavy, RecView scratch, RecView alt, int depth, u64 key_min,
                          u64 key_max, bool from_input) {
  if (heavy.m == 0) return;
  // the light edges left at most one component with edges: every heavy
  // edge is a self-loop (hybrid_boruvka_edgefilter.hpp:289-296)
  if (one_component()) {
    C.stats.filter.back().skipped = true;
    return;
  }
  bool gathered = false;
  // a partitioned solve (ranges of vertex ids that no record joins, e.g. a
  // batch of requests) marks the largest component of every range in a bitmap
  // (best_val is free here)
  const u64* inside = nullptr;
  if (part_count > 1) {
    partition_giant_bits(st, gpar, seg_end, n0, part, part_count, part_giant, best_val);
    inside = best_val;
  }
  if (from_input && (inside || giant_share >= 0.25)) {
    // few vertices outside the largest component(s): scan only their ranges
    // (the per-vertex arrays of the rounds are free here)
    outside_list(st, gpar, seg_start, seg_end, n0, giant_root, inside, best_dst, parent, label,
                 ctr + C_OUTSIDE);
    read_counters();
    const u64 n_list = C.host_counters[C_OUTSIDE];
    const u64 total = counter64(C_OUTSIDE + 2);
    if (n_list == 0) {
      C.stats.filter.back().skipped = true;
      return;
    }
    if (3 * total <= m0 && 2 * total <= scratch.m) {
      outside_gather(st, best_dst, parent, label, static_cast<u32>(n_list), total, in_s, in_d,
                     keys_of(heavy), key_min, key_max, gpar, giant_root, inside, scratch.s,
                     scratch.d, scratch.v, scratch.m, ctr + C_MNEXT, ctr + C_ERR);
      gathered = true;
    }
  }
  if (!gathered) {
    // most vertices usually sit in one component by now: mark it in a bitmap
    // (n/8 bytes, cache resident) so that records inside it are dropped
    // without the key read and the two parent-array gathers (best_val is free
    // here)
    const u64* giant = inside;
    if (!giant && giant_share >= 0.25) {
      component_bits(st, gpar, n0, giant_root, best_val);
      giant = best_val;
    }
    if (from_input && !heavy.v && !packed) heavy.v = input_keys();
    // the heavy records are symmetric and sorted right after: relabel the
    // src < dst half and write both directions (half the gathers)
    relabel_compact(st, heavy.s, heavy.d, keys_of(heavy), key_min, heavy.m, gpar, scratch.s,
                    scratch.d, scratch.v, C.lb_work, ctr + C_MNEXT, ctr + C_ERR, giant, n0, key_max,
                    /*twins=*/true);
  }
  check_errors("filter");
  const u64 Hk = C.host_counters[C_MNEXT];
  if (g_ktimer) g_ktimer->adjust_last(K_RELABEL_SCATTER, 16.0 * Hk);
  C.stats.filter.back().heavy_kept = Hk;
  if (Hk == 0) return;
  RecView kept = scratch.at(0, Hk);
  RecView sorted = sort_by_src(kept, alt.at(0, Hk), vbits);
  RecView other = (sorted.s == kept.s) ? alt.at(0, Hk) : kept;
  filter_rec(sorted, other, false, depth + 1);
}

// boruvka_recursive (include/algorithms/hybrid_boruvka_edgefilter.hpp:227-330)
// on records P (grouped by src, ids are parent-array representatives);
// Q is scratch of the same extent.
void Solver::filter_rec(RecView P, RecView Q, bool stop, int depth) {
  if (P.m == 0) return;
  const double avg_degree = double(P.m) / double(n0);
  if (avg_degree <= double(filter_threshold) || stop || P.m <= 1000 || depth > 40) {
    C.stats.filter.push_back(FilterStat{depth, P.m, 0, 0, true, 0});
    base_case(P, Q);
    return;
  }
  const u32 pivot = pick_pivot(P);
  const u64 L = partition(P, Q, pivot);
  const bool stop_rec = double(L) / double(P.m) > 0.85;  // filter_helpers.hpp:130-132
  C.stats.filter.push_back(FilterStat{depth, P.m, L, pivot, false, 0});
  const size_t me = C.stats.filter.size() - 1;
  filter_rec(Q.at(0, L), P.at(0, L), stop_rec, depth + 1);
  // heavy records sit in Q[L, m); P[L, m) is free scratch
  const u64 H = P.m - L;
  if (H == 0) return;
  C.stats.filter.push_back(C.stats.filter[me]);
  // (heavy records are consumed by the relabel, so their range is the alt)
  filter_heavy(Q.at(L, H), P.at(L, H), Q.at(L, H), depth);
}

void Solver::run(const kmst_config& cfg, u32* out_s, u32* out_d, u32* out_w, u64 out_cap,
                 u64* out_n, u64* total) {
  C.stats = SolveStats{};
  C.stats.n = n0;
  C.stats.m = m0;
  C.stats.algorithm = cfg.algorithm;
  filter_threshold = cfg.filter_threshold;
  allocate();
  const double avg_degree = double(m0) / double(std::max<u32>(n0, 1));
  const bool filter = cfg.algorithm == KMST_ALGO_FILTER_BORUVKA;
  const bool filter_base = avg_degree <= double(filter_threshold) || m0 <= 1000;
  setup(filter, filter && !filter_base);
  const RecView input{const_cast<u32*>(in_s), const_cast<u32*>(in_d), packed ? nullptr : val0, m0,
                      packed ? in_w : nullptr};
  if (filter) {
    iota_u32(st, gpar, n0);
    if (filter_base || pivots.K == 0) {
      C.stats.filter.push_back(FilterStat{0, m0, 0, 0, true, 0});
      rounds(input, X, Y, true);
    } else {
      filter_classes();
    }
  } else {
    rounds(input, X, Y, false);
  }
  // finish: the output is in (min, max) order -- the low 2 bits(n) bits of a
  // packed key, the twin position (monotone in (min, max)) of a positional
  // one.  The keys are sorted by group only (the min field; the position
  // without its low MST_RANK_BITS bits, at most bits(m0) bits), and one pass
  // orders each group while it decodes.
  const u64 cnt = mst_count;
  if (cnt > out_cap) throw Error(KMST_EINVAL, "output capacity too small");
  const u32 fbits = packed ? 2 * vbits : bits_for(m0);
  const u32 lbits = std::min(fbits, packed ? vbits : MST_RANK_BITS);
  u64* keys = mst;
  if (cnt > 1 && radix_sort_u64(st, mst, X.v, cnt, fbits - lbits, C.sort_work, ctr + C_ERR, lbits))
    keys = X.v;
  u64* total_dev = reinterpret_cast<u64*>(ctr + C_TOTAL64);
  fill_u64(st, total_dev, 1, 0);
  fill_u32(st, ctr + C_MSTLIST, 2, 0);
  // parent (n0 words) is free here: it lists the long groups' first keys
  mst_order(st, keys, cnt, vbits, packed, lbits, in_s, in_d, out_s, out_d, out_w, total_dev,
            parent, ctr + C_MSTLIST, ctr + C_MSTLONG);
  check_errors("finish");
  C.stats.mst_long_groups = C.host_counters[C_MSTLIST];
  C.stats.mst_resorted = C.host_counters[C_MSTLONG] != 0;
  if (C.host_counters[C_MSTLONG]) {
    // a vertex is the smaller endpoint of more than MST_GROUP_MAX MST edges
    // (a star): sort the whole order field and write everything again
    u64* alt = keys == mst ? X.v : mst;
    if (radix_sort_u64(st, keys, alt, cnt, fbits, C.sort_work, ctr + C_ERR)) keys = alt;
    fill_u64(st, total_dev, 1, 0);
    mst_output(st, keys, cnt, vbits, packed, in_s, in_d, out_s, out_d, out_w, total_dev);
    check_errors("finish");
  }
  *out_n = cnt;
  *total = counter64(C_TOTAL64);
  C.stats.mst_edges = cnt;
  C.stats.total_weight = *total;
}

}  // namespace kmst

// =================================================================== C ABI
namespace {

// The default contexts of the host API (one per device) live until the
// process ends and are never destroyed: a static destructor would make HIP
// calls (stream synchronise, frees) after the runtime may already have shut
// down during exit.  The driver reclaims their memory with the process.
std::mutex g_default_mu;
std::map<int, kmst_context*>& default_contexts() {
  static auto* m = new std::map<int, kmst_context*>();
  return *m;
}

template <typename F>
kmst_status guarded(F&& f) {
  return kmst::guarded_call(std::forward<F>(f));
}

kmst_context* default_context(int device) {
  auto& slot = default_contexts()[device];
  if (!slot) slot = new kmst_context(device);
  return slot;
}

void check_config(const kmst_config* cfg) {
  if (!cfg) throw kmst::Error(KMST_EINVAL, "null config");
  if (cfg->algorithm != KMST_ALGO_BORUVKA && cfg->algorithm != KMST_ALGO_FILTER_BORUVKA)
    throw kmst::Error(KMST_EINVAL, "unknown algorithm");
  if (cfg->compression != 0 && cfg->compression != 1)
    throw kmst::Error(KMST_EINVAL, "compression must be 0 (none) or 1 (difference codec)");
}

void solve_device_impl(kmst_context* ctx, const kmst_dev_edges* in, const kmst_config* cfg,
                       kmst_dev_mst* out, uint64_t* out_n, uint64_t* total_weight,
                       const kmst::u32* part = nullptr, kmst::u32 part_count = 0) {
  check_config(cfg);
  if (!ctx || !in || !out || !out_n || !total_weight)
    throw kmst::Error(KMST_EINVAL, "null argument");
  if (in->m_directed >= (uint64_t{1} << 32) - 1)
    throw kmst::Error(KMST_EINVAL, "more than 2^32-2 records");
  if (in->n >= (uint64_t{1} << 32) - 1) throw kmst::Error(KMST_EINVAL, "n >= 2^32-1");
  if (in->m_directed && (!in->src || !in->dst || !in->weight))
    throw kmst::Error(KMST_EINVAL, "null input array");
  KMST_HIP(hipSetDevice(ctx->device));
  kmst::KernelTimer* prev = kmst::g_ktimer;
  kmst::g_ktimer = ctx->timing ? &ctx->timer : nullptr;
  try {
    if (in->m_directed == 0 || in->n <= 1) {
      if (in->m_directed != 0) throw kmst::Error(KMST_EINVAL, "records with n <= 1");
      *out_n = 0;
      *total_weight = 0;
    } else {
      kmst::Solver s(*ctx, in->src, in->dst, in->weight, in->m_directed, in->n, part, part_count);
      s.run(*cfg, out->src, out->dst, out->weight, out->capacity, out_n, total_weight);
    }
    if (ctx->timing) ctx->timer.collect();
  } catch (...) {
    kmst::g_ktimer = prev;
    (void)hipStreamSynchronize(ctx->stream);
    if (ctx->timing) ctx->timer.collect();
    throw;
  }
  kmst::g_ktimer = prev;
}

kmst_status host_solve(const kmst_wedge* in, uint64_t m, const kmst_config* cfg,
                       int32_t algorithm, kmst_wedge* out, uint64_t* out_n,
                       uint64_t* total_weight) {
  return guarded([&] {
    check_config(cfg);
    if ((!in && m) || !out_n || !total_weight) throw kmst::Error(KMST_EINVAL, "null argument");
    if (m >= (uint64_t{1} << 32) - 1) throw kmst::Error(KMST_EINVAL, "more than 2^32-2 records");
    kmst_config c = *cfg;
    c.algorithm = algorithm;
    std::vector<uint32_t> s(m), d(m), w(m);
    uint64_t vmax = 0;
    for (uint64_t i = 0; i < m; ++i) {
      if (in[i].src >= 0xFFFFFFFFull || in[i].dst >= 0xFFFFFFFFull)
        throw kmst::Error(KMST_EINVAL, "vertex id >= 2^32-1");
      s[i] = static_cast<uint32_t>(in[i].src);
      d[i] = static_cast<uint32_t>(in[i].dst);
      w[i] = in[i].weight;
      vmax = std::max<uint64_t>(vmax, std::max(in[i].src, in[i].dst));
    }
    const uint64_t n = m ? vmax + 1 : 0;
    std::lock_guard<std::mutex> lk(g_default_mu);
    kmst_context* ctx = default_context(c.device);
    KMST_HIP(hipSetDevice(ctx->device));
    const uint64_t mm = m ? m : 1;
    uint32_t* ds = static_cast<uint32_t*>(ctx->in_s.get(mm * 4));
    uint32_t* dd = static_cast<uint32_t*>(ctx->in_d.get(mm * 4));
    uint32_t* dw = static_cast<uint32_t*>(ctx->in_w.get(mm * 4));
    const uint64_t cap = n ? n : 1;
    kmst_dev_mst o{static_cast<uint32_t*>(ctx->out_s.get(cap * 4)),
                   static_cast<uint32_t*>(ctx->out_d.get(cap * 4)),
                   static_cast<uint32_t*>(ctx->out_w.get(cap * 4)), cap};
    if (m && c.compression == 1) {
      // compression level 1: the records travel difference-encoded and are
      // decoded on the device (DifferenceEncodedGraph::get_WEdges)
      std::vector<uint8_t> bytes;
      std::vector<uint64_t> offset;
      std::vector<uint32_t> first;
      kmst::encode_host(s.data(), d.data(), w.data(), m, kmst::CODEC_CHUNK, bytes, offset, first);
      auto* db = static_cast<uint8_t*>(ctx->codec_bytes.get(bytes.size() + 16));
      auto* dof = static_cast<uint64_t*>(ctx->codec_off.get(offset.size() * 8));
      auto* dfi = static_cast<uint32_t*>(ctx->codec_first.get(first.size() * 4));
      auto* derr = static_cast<uint32_t*>(ctx->tune.get(64));
      kmst::fill_u32(ctx->stream, derr, 1, 0);
      if (!bytes.empty())
        KMST_HIP(hipMemcpyAsync(db, bytes.data(), bytes.size(), hipMemcpyHostToDevice, ctx->stream));
      KMST_HIP(hipMemcpyAsync(dof, offset.data(), offset.size() * 8, hipMemcpyHostToDevice,
                              ctx->stream));
      KMST_HIP(hipMemcpyAsync(dfi, first.data(), first.size() * 4, hipMemcpyHostToDevice,
                              ctx->stream));
      kmst::decode_device(ctx->stream, db, dof, dfi, m, kmst::CODEC_CHUNK, ds, dd, dw, derr);
      uint32_t err = 0;
      KMST_HIP(hipMemcpyAsync(&err, derr, 4, hipMemcpyDeviceToHost, ctx->stream));
      KMST_HIP(hipStreamSynchronize(ctx->stream));
      if (err) throw kmst::Error(KMST_EDEVICE, "difference-coded input failed to decode");
    } else if (m) {
      KMST_HIP(hipMemcpyAsync(ds, s.data(), m * 4, hipMemcpyHostToDevice, ctx->stream));
      KMST_HIP(hipMemcpyAsync(dd, d.data(), m * 4, hipMemcpyHostToDevice, ctx->stream));
      KMST_HIP(hipMemcpyAsync(dw, w.data(), m * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    kmst_dev_edges e{ds, dd, dw, m, n};
    solve_device_impl(ctx, &e, &c, &o, out_n, total_weight);
    const uint64_t k = *out_n;
    std::vector<uint32_t> hs(k), hd(k), hw(k);
    if (k) {
      KMST_HIP(hipMemcpyAsync(hs.data(), o.src, k * 4, hipMemcpyDeviceToHost, ctx->stream));
      KMST_HIP(hipMemcpyAsync(hd.data(), o.dst, k * 4, hipMemcpyDeviceToHost, ctx->stream));
      KMST_HIP(hipMemcpyAsync(hw.data(), o.weight, k * 4, hipMemcpyDeviceToHost, ctx->stream));
    }
    KMST_HIP(hipStreamSynchronize(ctx->stream));
    for (uint64_t i = 0; i < k; ++i) out[i] = kmst_wedge{hs[i], hd[i], hw[i], 0};
  });
}

struct TuneResult {
  uint64_t n = 0, weight = 0, fingerprint = 0;
  bool operator!=(const TuneResult& o) const {
    return n != o.n || weight != o.weight || fingerprint != o.fingerprint;
  }
};

// Coordinate search over the filter threshold and the sort launch shapes;
// each candidate is checked against the first one's MST.
void autotune_impl(kmst_context* ctx, const kmst_dev_edges* in, kmst_config* cfg,
                   int32_t repeats, kmst_tuning* tuning, double* best_ms, uint32_t* candidates) {
  check_config(cfg);
  if (!ctx || !in || !tuning || !best_ms || !candidates)
    throw kmst::Error(KMST_EINVAL, "null argument");
  KMST_HIP(hipSetDevice(ctx->device));
  const int reps = repeats < 1 ? 1 : repeats;
  const uint64_t cap = in->n ? in->n : 1;
  kmst_dev_mst out{static_cast<uint32_t*>(ctx->out_s.get(cap * 4)),
                   static_cast<uint32_t*>(ctx->out_d.get(cap * 4)),
                   static_cast<uint32_t*>(ctx->out_w.get(cap * 4)), cap};
  auto* fp = static_cast<uint64_t*>(ctx->tune.get(64));
  const kmst_tuning entry = ctx->tuning;
  auto solve_once = [&](const kmst_config& c, const kmst_tuning& t, TuneResult* r) {
    ctx->tuning = t;
    const auto t0 = std::chrono::steady_clock::now();
    uint64_t on = 0, tw = 0;
    solve_device_impl(ctx, in, &c, &out, &on, &tw);
    const double ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (r) {
      kmst::fill_u64(ctx->stream, fp, 1, 0);
      kmst::edge_fingerprint(ctx->stream, out.src, out.dst, out.weight, on, fp);
      KMST_HIP(hipMemcpyAsync(&r->fingerprint, fp, 8, hipMemcpyDeviceToHost, ctx->stream));
      KMST_HIP(hipStreamSynchronize(ctx->stream));
      r->n = on;
      r->weight = tw;
    }
    return ms;
  };
  TuneResult ref;
  bool have_ref = false;
  uint32_t count = 0;
  auto measure = [&](const kmst_config& c, const kmst_tuning& t) {
    TuneResult r;
    solve_once(c, t, &r);  // warm-up and result check
    if (!have_ref) {
      ref = r;
      have_ref = true;
    } else if (r != ref) {
      throw kmst::Error(KMST_EDEVICE, "autotune: candidates returned different MSTs");
    }
    std::vector<double> ts;
    for (int i = 0; i < reps; ++i) ts.push_back(solve_once(c, t, nullptr));
    std::sort(ts.begin(), ts.end());
    ++count;
    return ts[ts.size() / 2];
  };
  try {
    kmst_config best_cfg = *cfg;
    kmst_tuning best_t = entry;
    double best = measure(best_cfg, best_t);
    if (cfg->algorithm == KMST_ALGO_FILTER_BORUVKA) {
      for (uint64_t thr : {2, 3, 4, 6}) {
        if (thr == cfg->filter_threshold) continue;  // the entry point, timed above
        kmst_config c = best_cfg;
        c.filter_threshold = thr;
        const double ms = measure(c, best_t);
        if (ms < best) best = ms, best_cfg = c;
      }
    }
    for (uint32_t threads : {0u, 256u, 512u}) {
      if (threads == entry.sort_threads) continue;
      kmst_tuning t = best_t;
      t.sort_threads = threads;
      const double ms = measure(best_cfg, t);
      if (ms < best) best = ms, best_t = t;
    }
    {
      kmst_tuning t = best_t;
      t.balanced_digits = best_t.balanced_digits ? 0 : 1;
      const double ms = measure(best_cfg, t);
      if (ms < best) best = ms, best_t = t;
    }
    for (uint64_t pm : {uint64_t{0}, uint64_t{1} << 16, uint64_t{1} << 18, uint64_t{1} << 20}) {
      if (pm == entry.persistent_max) continue;
      kmst_tuning t = best_t;
      t.persistent_max = pm;
      const double ms = measure(best_cfg, t);
      if (ms < best) best = ms, best_t = t;
    }
    ctx->tuning = best_t;
    *tuning = best_t;
    cfg->filter_threshold = best_cfg.filter_threshold;
    *best_ms = best;
    *candidates = count;
  } catch (...) {
    ctx->tuning = entry;
    throw;
  }
}

}  // namespace

extern "C" {

void kmst_config_default(kmst_config* cfg) {
  if (!cfg) return;
  cfg->algorithm = KMST_ALGO_BORUVKA;
  cfg->local_preprocessing = 0;
  cfg->compression = 0;
  cfg->device = 0;
  cfg->filter_threshold = 4;
}

kmst_status kmst_context_create(int device, kmst_context** out) {
  return guarded([&] {
    if (!out) throw kmst::Error(KMST_EINVAL, "null out");
    *out = new kmst_context(device);
  });
}

void kmst_context_destroy(kmst_context* ctx) { delete ctx; }

void kmst_context_trim(kmst_context* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (kmst::DevBuf* b : ctx->all_bufs())
    if (b != &ctx->counters) b->release();
  ctx->sort_work = kmst::SortWork{};  // its status words are gone with `hist`
  ctx->lb_work = kmst::LookbackWork{};  // and these with `lookback`
}

kmst_status kmst_boruvka(const kmst_wedge* in, uint64_t m, const kmst_config* cfg, kmst_wedge* out,
                         uint64_t* out_n, uint64_t* total_weight) {
  return host_solve(in, m, cfg, KMST_ALGO_BORUVKA, out, out_n, total_weight);
}

kmst_status kmst_filter_boruvka(const kmst_wedge* in, uint64_t m, const kmst_config* cfg,
                                kmst_wedge* out, uint64_t* out_n, uint64_t* total_weight) {
  return host_solve(in, m, cfg, KMST_ALGO_FILTER_BORUVKA, out, out_n, total_weight);
}

kmst_status kmst_solve_device(kmst_context* ctx, const kmst_dev_edges* in, const kmst_config* cfg,
                              kmst_dev_mst* out, uint64_t* out_n, uint64_t* total_weight) {
  return guarded([&] { solve_device_impl(ctx, in, cfg, out, out_n, total_weight); });
}

const char* kmst_last_error(void) { return kmst::g_last_error.c_str(); }

kmst_status kmst_context_reserve(kmst_context* ctx, uint64_t m_directed, uint64_t n) {
  return guarded([&] {
    if (!ctx) throw kmst::Error(KMST_EINVAL, "null context");
    if (m_directed >= (uint64_t{1} << 32) - 1 || n >= (uint64_t{1} << 32) - 1)
      throw kmst::Error(KMST_EINVAL, "reservation beyond 32-bit ids");
    KMST_HIP(hipSetDevice(ctx->device));
    kmst::Solver s(*ctx, nullptr, nullptr, nullptr, m_directed, n);
    s.reserve();
  });
}

void kmst_context_set_timing(kmst_context* ctx, int on) {
  if (!ctx) return;
  ctx->timing = on != 0;
  ctx->timer.reset();
}

kmst_status kmst_encode_host(const uint32_t* src, const uint32_t* dst, const uint32_t* weight,
                             uint64_t m, uint32_t chunk_records, uint8_t* bytes, uint64_t bytes_cap,
                             uint64_t* offsets, uint32_t* first, uint64_t* bytes_used) {
  return guarded([&] {
    if (!bytes_used || chunk_records == 0 || (m && (!src || !dst || !weight)))
      throw kmst::Error(KMST_EINVAL, "null argument or zero chunk length");
    std::vector<uint8_t> b;
    std::vector<uint64_t> off;
    std::vector<uint32_t> fi;
    kmst::encode_host(src, dst, weight, m, chunk_records, b, off, fi);
    *bytes_used = b.size();
    if (!bytes) return;  // size query
    if (bytes_cap < b.size() || !offsets || !first)
      throw kmst::Error(KMST_EINVAL, "output buffers too small");
    std::copy(b.begin(), b.end(), bytes);
    std::copy(off.begin(), off.end(), offsets);
    std::copy(fi.begin(), fi.end(), first);
  });
}

void kmst_tuning_default(kmst_tuning* t) {
  if (!t) return;
  t->sort_threads = 0;
  t->balanced_digits = 1;
  t->sort_wide_min = kmst::SORT_WIDE_MIN;
  t->persistent_max = kmst::PERSIST_MAX_DEFAULT;
}

kmst_status kmst_context_set_tuning(kmst_context* ctx, const kmst_tuning* t) {
  return guarded([&] {
    if (!ctx || !t) throw kmst::Error(KMST_EINVAL, "null argument");
    if (t->sort_threads != 0 && t->sort_threads != 256 && t->sort_threads != 512)
      throw kmst::Error(KMST_EINVAL, "sort_threads must be 0, 256 or 512");
    if (t->balanced_digits > 1) throw kmst::Error(KMST_EINVAL, "balanced_digits must be 0 or 1");
    if (t->sort_wide_min == 0) throw kmst::Error(KMST_EINVAL, "sort_wide_min must be positive");
    if (t->persistent_max >= (uint64_t{1} << 32))
      throw kmst::Error(KMST_EINVAL, "persistent_max must be below 2^32");
    ctx->tuning = *t;
  });
}

kmst_status kmst_context_get_tuning(const kmst_context* ctx, kmst_tuning* t) {
  return guarded([&] {
    if (!ctx || !t) throw kmst::Error(KMST_EINVAL, "null argument");
    *t = ctx->tuning;
  });
}

kmst_status kmst_autotune(kmst_context* ctx, const kmst_dev_edges* in, kmst_config* cfg,
                          int32_t repeats, kmst_tuning* tuning, double* best_ms,
                          uint32_t* candidates) {
  return guarded([&] { autotune_impl(ctx, in, cfg, repeats, tuning, best_ms, candidates); });
}

kmst_status kmst_gen_gnm_device(kmst_context* ctx, uint64_t n, uint64_t m, uint32_t max_weight,
                                uint64_t seed, uint32_t* src, uint32_t* dst, uint32_t* weight) {
  return guarded([&] {
    if (!ctx || (m && (!src || !dst || !weight))) throw kmst::Error(KMST_EINVAL, "null argument");
    KMST_HIP(hipSetDevice(ctx->device));
    kmst::gen_gnm_device(*ctx, n, m, max_weight, seed, src, dst, weight);
  });
}

kmst_status kmst_gen_gnm_host(uint64_t n, uint64_t m, uint32_t max_weight, uint64_t seed,
                              uint32_t* src, uint32_t* dst, uint32_t* weight) {
  return guarded([&] {
    if (m && (!src || !dst || !weight)) throw kmst::Error(KMST_EINVAL, "null argument");
    kmst::gen_gnm_host(n, m, max_weight, seed, src, dst, weight);
  });
}

kmst_status kmst_k_segmented_argmin(kmst_context* ctx, const uint32_t* src, const uint32_t* dst,
                                    const uint64_t* val, uint64_t m, uint64_t n,
                                    uint64_t* best_val, uint32_t* best_dst) {
  return guarded([&] {
    if (!ctx) throw kmst::Error(KMST_EINVAL, "null context");
    KMST_HIP(hipSetDevice(ctx->device));
    const uint64_t tiles = (m + kmst::ARG_TILE - 1) / kmst::ARG_TILE;
    auto* pieces = static_cast<kmst::ArgPiece*>(
        ctx->pieces.get(2 * tiles * sizeof(kmst::ArgPiece) + 64));
    auto* ctr = static_cast<uint32_t*>(ctx->counters.get(256));
    kmst::segmented_argmin(ctx->stream, src, dst, kmst::KeySrc{val, nullptr, 0}, m, n, best_val,
                           best_dst, pieces,
                           ctr + kmst::C_PIECES);
    KMST_HIP(hipStreamSynchronize(ctx->stream));
  });
}

kmst_status kmst_k_segmented_argmin_packed(kmst_context* ctx, const uint32_t* src,
                                           const uint32_t* dst, const uint32_t* weight,
                                           uint32_t vbits, uint64_t m, uint64_t n,
                                           uint64_t* best_val, uint32_t* best_dst) {
  return guarded([&] {
    if (!ctx) throw kmst::Error(KMST_EINVAL, "null context");
    if (vbits == 0 || vbits > 31) throw kmst::Error(KMST_EINVAL, "vbits outside [1, 31]");
    if (n > (uint64_t{1} << vbits)) throw kmst::Error(KMST_EINVAL, "n exceeds 2^vbits");
    KMST_HIP(hipSetDevice(ctx->device));
    const uint64_t tiles = (m + kmst::ARG_TILE - 1) / kmst::ARG_TILE;
    auto* pieces = static_cast<kmst::ArgPiece*>(
        ctx->pieces.get(2 * tiles * sizeof(kmst::ArgPiece) + 64));
    auto* ctr = static_cast<uint32_t*>(ctx->counters.get(256));
    kmst::segmented_argmin(ctx->stream, src, dst, kmst::KeySrc{nullptr, weight, vbits}, m, n,
                           best_val, best_dst, pieces, ctr + kmst::C_PIECES);
    KMST_HIP(hipStreamSynchronize(ctx->stream));
  });
}

kmst_status kmst_k_sort_pairs(kmst_context* ctx, uint32_t* src, uint32_t* dst, uint64_t* val,
                              uint64_t m, uint32_t bits) {
  return guarded([&] {
    if (!ctx) throw kmst::Error(KMST_EINVAL, "null context");
    if (bits > 32) throw kmst::Error(KMST_EINVAL, "bits > 32");
    KMST_HIP(hipSetDevice(ctx->device));
    const uint64_t mm = m ? m : 1;
    auto* as = static_cast<uint32_t*>(ctx->xs.get(mm * 4));
    auto* ad = static_cast<uint32_t*>(ctx->xd.get(mm * 4));
    auto* av = static_cast<uint64_t*>(ctx->xv.get(mm * 8));
    ctx->sort_work.bind(ctx->stream,
                        static_cast<uint32_t*>(ctx->hist.get(kmst::sort_work_words(mm) * 4)), mm);
    kmst::apply_tuning(*ctx);
    auto* ctr = static_cast<uint32_t*>(ctx->counters.get(256));
    kmst::fill_u32(ctx->stream, ctr, 64, 0);
    if (kmst::radix_sort(ctx->stream, src, dst, val, as, ad, av, m, bits, 2 * bits,
                         ctx->sort_work, ctr + kmst::C_ERR)) {
      KMST_HIP(hipMemcpyAsync(src, as, m * 4, hipMemcpyDeviceToDevice, ctx->stream));
      KMST_HIP(hipMemcpyAsync(dst, ad, m * 4, hipMemcpyDeviceToDevice, ctx->stream));
      KMST_HIP(hipMemcpyAsync(val, av, m * 8, hipMemcpyDeviceToDevice, ctx->stream));
    }
    uint32_t err = 0;
    KMST_HIP(hipMemcpyAsync(&err, ctr + kmst::C_ERR, 4, hipMemcpyDeviceToHost, ctx->stream));
    KMST_HIP(hipStreamSynchronize(ctx->stream));
    if (err) throw kmst::Error(KMST_EDEVICE, "sort: bounded wait expired");
  });
}

kmst_status kmst_k_pointer_jump(kmst_context* ctx, uint32_t* parent, uint64_t n) {
  return guarded([&] {
    if (!ctx) throw kmst::Error(KMST_EINVAL, "null context");
    KMST_HIP(hipSetDevice(ctx->device));
    auto* ctr = static_cast<uint32_t*>(ctx->counters.get(256));
    kmst::fill_u32(ctx->stream, ctr, 64, 0);
    kmst::pointer_jump(ctx->stream, n, parent, ctr + kmst::C_ERR);
    uint32_t err = 0;
    KMST_HIP(hipMemcpyAsync(&err, ctr, 4, hipMemcpyDeviceToHost, ctx->stream));
    KMST_HIP(hipStreamSynchronize(ctx->stream));
    if (err) throw kmst::Error(KMST_EDEVICE, "pointer jumping did not converge");
  });
}

size_t kmst_context_stats_json(const kmst_context* ctx, char* buf, size_t cap) {
  if (!ctx) return 0;
  std::ostringstream o;
  const kmst::SolveStats& s = ctx->stats;
  o << "{\"n\":" << s.n << ",\"m\":" << s.m << ",\"algorithm\":" << s.algorithm
    << ",\"packed_keys\":" << (s.packed_keys ? "true" : "false")
    << ",\"lazy_keys\":" << (s.lazy_keys ? "true" : "false")
    << ",\"readback\":\"" << (ctx->host_mapped ? "publish" : "copy") << "\""
    << ",\"mst_edges\":" << s.mst_edges << ",\"total_weight\":" << s.total_weight
    << ",\"mst_long_groups\":" << s.mst_long_groups
    << ",\"mst_resorted\":" << (s.mst_resorted ? "true" : "false") << ",\"rounds\":[";
  for (size_t i = 0; i < s.rounds.size(); ++i) {
    const auto& r = s.rounds[i];
    o << (i ? "," : "") << "{\"n\":" << r.n << ",\"m\":" << r.m << ",\"n_next\":" << r.n_next
      << ",\"m_next\":" << r.m_next << ",\"run_starts\":" << r.run_starts
      << ",\"far_starts\":" << r.far_starts
      << ",\"persistent\":" << (r.persistent ? "true" : "false")
      << ",\"twins\":" << (r.twins ? "true" : "false");
    if (r.persistent)
      o << ",\"phase_us\":[" << r.phase_us[0] << "," << r.phase_us[1] << "," << r.phase_us[2]
        << "," << r.phase_us[3] << "]";
    o << "}";
  }
  o << "],\"filter\":[";
  for (size_t i = 0; i < s.filter.size(); ++i) {
    const auto& f = s.filter[i];
    o << (i ? "," : "") << "{\"depth\":" << f.depth << ",\"m\":" << f.m << ",\"light\":" << f.light
      << ",\"pivot\":" << f.pivot << ",\"base\":" << (f.base ? "true" : "false")
      << ",\"heavy_kept\":" << f.heavy_kept << ",\"heavy_skipped\":"
      << (f.skipped ? "true" : "false") << "}";
  }
  o << "],\"kernels\":{";
  bool firstk = true;
  for (const auto& kv : ctx->timer.totals) {
    o << (firstk ? "" : ",") << "\"" << kmst::kernel_kind_name(kv.first)
      << "\":{\"launches\":" << kv.second.launches << ",\"ms\":" << kv.second.ms
      << ",\"bytes\":" << kv.second.bytes << "}";
    firstk = false;
  }
  o << "}}";
  const std::string str = o.str();
  if (buf && cap > str.size()) {
    std::memcpy(buf, str.c_str(), str.size() + 1);
    return str.size();
  }
  return str.size() + 1;
}

}  // extern "C"

namespace kmst {
kmst_status solve_device_partitioned(kmst_context* ctx, const kmst_dev_edges* in,
                                     const kmst_config* cfg, kmst_dev_mst* out, uint64_t* out_n,
                                     uint64_t* total_weight, const u32* part, u32 count) {
  return guarded([&] { solve_device_impl(ctx, in, cfg, out, out_n, total_weight, part, count); });
}
}  // namespace kmst
