// kmst_kernels.hpp -- host-side launchers of the libkmst HIP kernels.
// Every launcher enqueues work on `st` and returns immediately.
#pragma once

#include <vector>

#include "kmst_common.hpp"

namespace kmst {

// Tile shapes (elements per workgroup).
constexpr int SCAN_IPT = 16;
constexpr int SCAN_TILE = BLOCK * SCAN_IPT;      // 4096
constexpr int SORT_IPT = 16;
constexpr int SORT_TILE = BLOCK * SORT_IPT;      // 4096
constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;
constexpr int CMP_IPT = 16;
constexpr int CMP_TILE = BLOCK * CMP_IPT;        // 4096
constexpr int ARG_IPT = 16;
constexpr int ARG_TILE = BLOCK * ARG_IPT;        // 4096

// Boundary piece of the segmented argmin: a vertex whose records cross a tile
// boundary; the per-tile partial minimum is resolved in a fix-up pass.
struct ArgPiece {
  u64 val;
  u32 src;
  u32 idx;
};

void fill_u64(hipStream_t st, u64* p, u64 n, u64 value);
void fill_u32(hipStream_t st, u32* p, u64 n, u32 value);
// ctr[0, ctr_words) to host[0, ...), extra[0, extra_words) to host[extra_at, ...),
// then host[seq_at] = seq with a system-scope release (host: coherent pinned
// memory the device writes directly).
void publish_words(hipStream_t st, const u32* ctr, u32 ctr_words, const u32* extra,
                   u32 extra_words, u32* host, u32 extra_at, u32 seq_at, u32 seq);

// Exclusive prefix sum of in[0,n) into out (may alias in); *total_dev gets the
// sum when non-null.  `scratch` needs scan_scratch_elems(n) entries.
u64 scan_scratch_elems(u64 n);
void scan_exclusive_u32(hipStream_t st, const u32* in, u32* out, u64 n, u32* scratch,
                        u32* total_dev);
// With flagged_low (n entries) also flagged_low[i] = (u32) out[i] where bit 0
// of in[i] is set, NONE32 elsewhere; with tag_only the flagged entries get
// (u32) out[i] | ROOT_TAG and the others are left as they are.
constexpr u32 ROOT_TAG = 0x80000000u;
void scan_exclusive_u64(hipStream_t st, const u64* in, u64* out, u64 n, u64* scratch,
                        u64* total_dev, u32* flagged_low = nullptr, bool tag_only = false);

// Setup: contract checks (hsum[0..1] = forward / reverse multiset hashes),
// CSR offsets, strict keys (packed when they fit 64 bits, positional else).
// Ascending weight pivots q[0..K); weight class of w = #{j : w > q[j]}.
constexpr u32 MAX_PIVOTS = 8;
constexpr u32 MAX_CLASSES = MAX_PIVOTS + 1;
struct WeightPivots {
  u32 q[MAX_PIVOTS];
  u32 K;
};
// With P.K > 0 it also counts the records of every weight class into
// class_counts[0..K] (zeroed by the caller); with seg_start / seg_end (n
// entries each, seg_end zeroed by the caller) it records every vertex's record
// range [seg_start[v], seg_end[v]) (seg_end stays 0 for vertices without
// records).  With unit_c0 / unit_c1 (record_units(m) entries each) it writes
// how many records of every 256-record unit fall in weight class 0 (w <= q0)
// and class 1 (q0 < w <= q1), when the arrays are 16-byte aligned; it returns
// whether it did (else unit_class_counts does it from the weights).
u64 record_units(u64 m);
void unit_class_counts(hipStream_t st, const u32* w, u64 m, u32 q0, u32 q1, u32* unit_c0,
                       u32* unit_c1);
bool validate_records(hipStream_t st, const u32* src, const u32* dst, const u32* w, u64 m,
                      u64 n, u32* err, u32* fwd_count, u32* max_w, u64* hsum,
                      const WeightPivots& P, u64* class_counts, u32* seg_start = nullptr,
                      u32* seg_end = nullptr, u32 q0 = 0, u32 q1 = 0, u32* unit_c0 = nullptr,
                      u32* unit_c1 = nullptr);
void build_offsets(hipStream_t st, const u32* src, u64 m, u64 n, u32* off);
// Record ranges [seg_start[v], seg_end[v]) of records sorted by src, written
// for the vertices with records only; bounded work on an unchecked input.
void record_ranges(hipStream_t st, const u32* src, u64 m, u64 n, u32* seg_start, u32* seg_end);
// Positional keys of a validated input (pair_position search in the record ranges
// [row_lo[v], row_hi[v]): the offsets off / off + 1, or seg_start / seg_end).
void build_keys(hipStream_t st, const u32* src, const u32* dst, const u32* w, u64 m,
                const u32* row_lo, const u32* row_hi, u64* val, u32* err);
void build_keys_packed(hipStream_t st, const u32* src, const u32* dst, const u32* w, u64 m,
                       u32 vbits, u64* val);

// K1 segmented argmin over records grouped by src.
void segmented_argmin(hipStream_t st, const u32* src, const u32* dst, const KeySrc& keys,
                      u64 m, u64 n, u64* best_val, u32* best_dst, ArgPiece* pieces,
                      u32* piece_count);

// K1 over records in any order (atomic minima; see the kernels): best_val is
// filled here, best_dst set for every vertex with records.
void unordered_argmin(hipStream_t st, const u32* src, const u32* dst, const KeySrc& keys, u64 m,
                      u64 n, u64* best_val, u32* best_dst);

// K2 hook (state[v]: bit 0 = root with an edge, bit 32 = emits its edge);
// K3 pointer jumping; renumbering + MST emission from the scanned state.
void hook(hipStream_t st, u64 n, const u64* best_val, const u32* best_dst, u32* parent,
          u64* state);
// With roots, also counts the roots v with seg_end[v] != 0 into *roots.
void pointer_jump(hipStream_t st, u64 n, u32* parent, u32* err, const u32* seg_end = nullptr,
                  u32* roots = nullptr);
// Pointer jumping, labels, MST emission and (with gpar) the parent-array
// update of one round in one pass; the state scan must have run.  Without
// root_id (n < 2^31) the scan has written every surviving root's new id into
// its own parent entry, tagged with ROOT_TAG.
void jump_and_label(hipStream_t st, u64 n, u32* parent, const u32* root_id, const u64* state,
                    const u64* pos, const u64* best_val, u32* label, u64* mst, u64 mst_base,
                    const u32* rep, u32* gpar, u32* rep_next, u32* err);
// root_id[r] = new id of root r (NONE32 for other vertices), from the scan
void make_labels(hipStream_t st, u64 n, const u32* parent, const u32* root_id, const u64* state,
                 const u64* pos, const u64* best_val, u32* label, u64* mst, u64 mst_base);

// Look-back state of the single-pass compactions (relabel, split): one
// status word per tile and a tile counter (lookback_elems(m) u64 words).
// Status words carry a call epoch, so they are cleared only when the memory
// is (re)bound; bind() with the buffer for the largest m before the calls.
inline u64 lookback_elems(u64 m) { return (m + CMP_TILE - 1) / CMP_TILE + 1; }
struct LookbackWork {
  u64* words = nullptr;
  u64 tiles_cap = 0;     // tiles the status area holds
  u32 epoch = 0;         // epoch of the last call
  u64 tiles_issued = 0;  // value of the tile counter
  void bind(hipStream_t st, u64* p, u64 m);
  u32 next_call(hipStream_t st, u32 tiles);  // epoch for a call over `tiles` tiles
};

// K4+K5 relabel, drop self-loops and dead records, order-preserving compaction
// in one pass; *count_dev gets the kept count.  Tiles of 512 threads from
// RELABEL_WIDE_MIN records on, 256 below.
constexpr u64 RELABEL_WIDE_MIN = u64{1} << 20;
// Records whose key is below key_min are dropped too (0 keeps all).  With
// giant (a bitmap over the n vertex ids of one component of label, see
// component_bits) records with both ends inside it are dropped without reading
// their key or labels.  With twins (symmetric records) only the src < dst half
// is relabelled and each kept record leaves in both directions: half the
// label gathers, and the output is no longer grouped by src.
void relabel_compact(hipStream_t st, const u32* s_in, const u32* d_in, const KeySrc& keys,
                     u64 key_min, u64 m, const u32* label, u32* s_out, u32* d_out, u64* v_out,
                     LookbackWork& lb, u32* count_dev, u32* err, const u64* giant = nullptr,
                     u64 n = 0, u64 key_max = ~0ull, bool twins = false);
// bits[v / 64] bit v % 64 = (gpar[v] == root) for v < n; bits needs
// ceil(n / 64) entries.
void component_bits(hipStream_t st, const u32* gpar, u64 n, u32 root, u64* bits);
// The same for a partition of the vertices into disjoint ranges [part[r],
// part[r + 1]) (r < count) that no record joins: the bit is set when v's root
// is the most frequent root of a sample of its range's vertices with records
// (giant[r], count entries of scratch).
void partition_giant_bits(hipStream_t st, const u32* gpar, const u32* seg_end, u64 n,
                          const u32* part, u32 count, u32* giant, u64* bits);

// Heavy records by the ranges of the vertices outside the largest component:
// that of `root`, or, with the bitmap `inside` (partition_giant_bits), the
// largest of every partition range.  Every record whose ends lie in different
// components has an end outside, and in a symmetric (src, dst)-sorted input
// that end's range [seg_start, seg_end) holds it.  outside_list writes the
// vertices with records outside (list, seg_lo, seg_hi: up to n entries
// each); counters[0] = count, counters[2..3] = u64 total range length.
// outside_gather then emits, for keys in [kmin, kmax], the relabelled
// cross-component records in both directions (see the kernel), at most
// 2 * total of them.  With lazy positional keys (keys.twin_lo / twin_hi the
// validated record ranges) [kmin, kmax] bound the weight << 32 and the keys
// of the emitted records are made by their pair_position search.
void outside_list(hipStream_t st, const u32* gpar, const u32* seg_start, const u32* seg_end,
                  u64 n, u32 root, const u64* inside, u32* list, u32* seg_lo, u32* seg_hi,
                  u32* counters);
void outside_gather(hipStream_t st, const u32* list, const u32* seg_lo, const u32* seg_hi,
                    u32 n_list, u64 total, const u32* in_s, const u32* in_d, const KeySrc& keys,
                    u64 kmin, u64 kmax, const u32* gpar, u32 root, const u64* inside, u32* s_out,
                    u32* d_out, u64* v_out, u64 cap, u32* count, u32* err);

// K6 stable LSD radix sort by key (s << dbits) | d over total_bits bits; with
// dbits == 0 the key is s alone and d is carried along.  Returns true when the
// sorted data ended in the *_alt arrays.  d / v (and their alt) may be null.
// A bounded look-back that expires raises ERR_SPIN in *err.
//
// Workspace: sort_work_words(m) u32 words (8-byte aligned) holding the
// per-(tile, digit) look-back status words, a tile counter and the digit
// histograms.  Status words carry a pass epoch, so they are cleared only when
// the memory is (re)bound; bind() before every sort with the buffer for the
// largest m it will see.
u64 sort_work_words(u64 m);
struct SortWork {
  u32* words = nullptr;
  u64 tiles_cap = 0;     // tiles the status area holds
  u32 epoch = 0;         // epoch of the last pass
  u64 tiles_issued = 0;  // value of the tile counter
  // threads per scatter tile (256 or 512; tiles of threads * SORT_IPT
  // records), 0 = by size (512 from SORT_WIDE_MIN records on)
  u32 threads = 0;
  // digit widths: balanced over the ceil(bits / 8) passes, else 8 bits
  // (the last pass takes the rest)
  bool balanced = true;
  u64 wide_min = u64{1} << 20;  // with threads 0: 512-thread tiles from here on
  void bind(hipStream_t st, u32* p, u64 m);
};
constexpr u64 SORT_WIDE_MIN = u64{1} << 20;
u32 sort_threads(const SortWork& w, u64 m);

// *out += sum over i < count of a hash of (s[i], d[i], w[i]) (order-free
// fingerprint of an MST).
void edge_fingerprint(hipStream_t st, const u32* s, const u32* d, const u32* w, u64 count,
                      u64* out);

// Per-pass digits of a radix sort: pass p sorts by key bits
// [shift[p], shift[p] + width[p]), width[p] <= RADIX_BITS.
constexpr u32 MAX_PASSES = 8;
struct DigitPlan {
  u32 passes;
  u8 shift[MAX_PASSES];
  u8 width[MAX_PASSES];
};
DigitPlan plan_digits(u32 total_bits, bool balanced);
// With dedup_last (needs d, v and dbits > 0) the last pass also keeps only
// the lightest record of each run of equal (s, d) inside an aligned group of
// 64 positions; the rest are marked dead.
bool radix_sort(hipStream_t st, u32* s, u32* d, u64* v, u32* s_alt, u32* d_alt, u64* v_alt,
                u64 m, u32 dbits, u32 total_bits, SortWork& work, u32* err,
                bool dedup_last = false);

// K7 keep the lightest record of each (s, d) run (others marked dead).
void dedup_mark(hipStream_t st, const u32* s, const u32* d, u64* v, u64 m);

// The same sort for one u64 array by its key bits [lo, lo + bits).
bool radix_sort_u64(hipStream_t st, u64* v, u64* v_alt, u64 m, u32 bits, SortWork& work,
                    u32* err, u32 lo = 0);

// K9 MST keys (sorted by their low bits: (min, max) for packed keys, the twin
// position for positional ones) -> canonical (min, max, w) output + u64 weight.
// A positional key's ends come from the input record at its position, its
// weight from the key.
void mst_output(hipStream_t st, const u64* keys, u64 count, u32 vbits, bool packed,
                const u32* in_src, const u32* in_dst, u32* out_s, u32* out_d, u32* out_w,
                u64* total_dev);

// K9 without the full sort.  The order field of an MST key is (min, max) for a
// packed key (its low 2 vbits bits), the twin position for a positional one
// (its low 32 bits); its group is the field without the low `lbits` bits
// (min, or the position / 64).  Given keys sorted by group, writes every key
// at its place in field order, decoded, and adds the weights to *total_dev.
// Groups longer than MST_GROUP_MAX are left unwritten and set *too_long.
// `list` holds at least count / MST_GROUP_SMALL + 1 words; *list_count and
// *too_long start at 0.
constexpr u32 MST_RANK_BITS = 6;  // positional keys: positions / 64 form the groups
constexpr u32 MST_GROUP_SMALL = 64;
constexpr u32 MST_CHUNK = 1024;
constexpr u32 MST_GROUP_MAX = 8 * MST_CHUNK;  // one group sorted in 64 KiB of LDS
void mst_order(hipStream_t st, const u64* keys, u64 count, u32 vbits, bool packed, u32 lbits,
               const u32* in_src, const u32* in_dst, u32* out_s, u32* out_d, u32* out_w,
               u64* total_dev, u32* list, u32* list_count, u32* too_long);

// How many of RUN_SAMPLES evenly spread records of s[0..*m_dev) start a run
// of equal s (differ from the record before them; record 0 counts), in the
// low 16 bits of *count, and how many of those start far from the record
// before (s / 16 differs), in the high 16 bits: estimates of how many
// atomic minima a round without a sort takes and how many of them hit a
// line the previous run did not.
constexpr u32 RUN_SAMPLES = 4096;
constexpr u64 UNORDERED_DEGREE = 8;
constexpr u64 UNORDERED_MIN_DEGREE = 2;
constexpr u64 UNORDERED_SMALL_N = u64{1} << 19;  // 4 MiB of u64 minima
// Relabels on more vertex ids (a u32 label table past one XCD's 4 MiB L2) of
// inputs whose sampled records mostly join ids FAR_SPAN or more apart gather
// the src < dst half only (relabel_compact's twins).
constexpr u64 TWINS_MIN_N = u64{1} << 20;
constexpr u32 FAR_SPAN = 1u << 16;
void sample_run_starts(hipStream_t st, const u32* s, const u32* m_dev, u32* count);

// Filter-Boruvka helpers.
void iota_u32(hipStream_t st, u32* p, u64 n);
void update_parent_array(hipStream_t st, u64 n, const u32* rep, const u32* parent,
                         const u64* state, const u64* pos, u32* gpar, u32* rep_next);
// weight of a key = key >> wshift (32 for positional keys, 2 bits(n) for packed)
void sample_weights(hipStream_t st, const u64* v, const u32* wt, u64 m, u32 k, u32 wshift,
                    u32* out);
// *out = how many of the same k sampled records join ids at least span apart
void sample_far(hipStream_t st, const u32* s, const u32* d, u64 m, u32 k, u32 span, u32* out);
void split_light_heavy(hipStream_t st, const u32* s_in, const u32* d_in, const KeySrc& keys,
                       u64 m, u32 pivot, u32 wshift, bool write_heavy, u32* s_out, u32* d_out,
                       u64* v_out, LookbackWork& lb, u32* light_count_dev, u32* err);
void count_roots(hipStream_t st, const u32* gpar, const u32* seg_end, u64 n, u32* count_dev);

// Vertex compaction of records sorted by src (symmetric, so the sources are
// all the vertices): flag[i] = 1 where record i starts a run of equal src;
// with id the exclusive scan of the flags, new_id[src] = id and rep[id] = src
// at every run start.
void run_start_flags(hipStream_t st, const u32* s, u64 m, u32* flag);
void number_runs(hipStream_t st, const u32* s, u64 m, const u32* id, u32* new_id, u32* rep);

// Weights-only pass over the input: *max_w, the record count of every weight
// class into class_counts[0..P.K] (zeroed by the caller), and the class 0
// (w <= q0) / class 1 (q0 < w <= q1) counts of every 256-record unit into
// unit_c0 / unit_c1 (record_units(m) entries each).
void weight_classes(hipStream_t st, const u32* w, u64 m, const WeightPivots& P, u32 q0, u32 q1,
                    u64* class_counts, u32* max_w, u32* unit_c0, u32* unit_c1);

// The input checks of validate_records, done inside the class scatter: ids
// below n and no self-loops, (src, dst) order, forward count and the two
// symmetry hashes (hsum[0..1]), and every vertex's record range.
struct InputCheck {
  u64 n = 0;
  u32* fwd_count = nullptr;
  u64* hsum = nullptr;
  u32* seg_start = nullptr;
  u32* seg_end = nullptr;
};

// Weight classes 0 (w <= q0) and 1 (q0 < w <= q1) of the input, in input
// order, to out[0, c0) and out[base1, base1 + c1); off0 / off1 are the
// exclusive scans of the unit counts.  written = c0 + c1; positions at or past
// it raise ERR_BOUNDS and are not written.  With check (packed keys, or lazy
// positional keys: keys.twin_lo / twin_hi the record ranges of the unchecked
// input, from record_ranges) the pass also validates the input, and records
// the ranges unless check->seg_end is null.  Lazy keys are made for the
// written records only and need the check.
void class_scatter(hipStream_t st, const u32* s_in, const u32* d_in, const KeySrc& keys, u64 m,
                   u32 wshift, u32 q0, u32 q1, const u32* off0, const u32* off1, u64 base1,
                   u64 written, u32* s_out, u32* d_out, u64* v_out, u32* err,
                   const InputCheck* check = nullptr);

// Difference codec of compression level 1 (kmst_codec.hip): chunks of
// chunk_records records, the first raw, the rest as 7-bit varints of
// (src delta, dst, weight).
constexpr u32 CODEC_CHUNK = 32;
u64 encoded_chunks(u64 m, u32 chunk_records);
void encode_host(const u32* src, const u32* dst, const u32* w, u64 m, u32 chunk_records,
                 std::vector<u8>& bytes, std::vector<u64>& offset, std::vector<u32>& first);
// Decodes into src / dst / w (m entries each); a malformed stream raises
// ERR_INVALID in *err.
void decode_device(hipStream_t st, const u8* bytes, const u64* offset, const u32* first, u64 m,
                   u32 chunk_records, u32* src, u32* dst, u32* w, u32* err);

// Every remaining Boruvka round of a small graph in one cooperative launch
// (kmst_persist.hip).  Round 0 reads (s0, d0, k0); round r writes its kept,
// relabelled records to s[r & 1], d[r & 1], v[r & 1] (cap records each) and
// the next round reads them.  Ids keep their values: a vertex whose
// component merged points at the component's root in parent[], live roots
// keep their own id.  The MST keys go to mst[mst_base + slot].  With gpar,
// gpar[rep[v]] (gpar[v] without rep) takes the name of v's root every round
// v has records.  ctl (PERSIST_CTL_WORDS, directly followed by log; both
// zeroed by the launcher) ends with the MST edges added in ctl[4] and the
// rounds in ctl[5]; round r's PERSIST_LOG_WORDS log words hold its records,
// its live vertices (low half) and MST edges (high half), and the wall clock
// (s_memrealtime ticks) of workgroup 0 at the start of the round and at the end
// of its bind, hook, jump and relabel phases.  One workgroup (graphs up to
// PERSIST_ONE_BLOCK records and vertices) synchronises with its own barrier.
constexpr u32 PERSIST_CTL_WORDS = 8;
constexpr u64 PERSIST_MAX_DEFAULT = u64{1} << 18;
// The persistent rounds keep parallel records (no dedup), so they are used
// only up to this average degree, or on rounds this small; denser rounds
// shrink faster by sorting.
constexpr u64 PERSIST_MAX_DEGREE = 32;
constexpr u64 PERSIST_SMALL = u64{1} << 14;
constexpr u64 PERSIST_ONE_BLOCK = u64{1} << 12;
constexpr u32 PERSIST_MAX_ROUNDS = 64;
constexpr u32 PERSIST_LOG_WORDS = 7;
constexpr u32 PERSIST_LOG_TOTAL = PERSIST_LOG_WORDS * PERSIST_MAX_ROUNDS;
// grid-barrier group counters (u32 words) that directly follow log
constexpr u32 PERSIST_GROUP_WORDS = 16 * 32;
// ticks of the wall clock the persistent kernel stamps, per microsecond
double persistent_ticks_per_us(int device);
struct PersistArgs {
  const u32* s0 = nullptr;
  const u32* d0 = nullptr;
  KeySrc k0{};
  u64 m0 = 0;
  u32* s[2] = {nullptr, nullptr};
  u32* d[2] = {nullptr, nullptr};
  u64* v[2] = {nullptr, nullptr};
  u64 cap = 0;
  u64 n = 0;
  u64* best = nullptr;
  u32* bdst = nullptr;
  u32* parent = nullptr;
  u64* mst = nullptr;
  u64 mst_base = 0;
  u64 mst_cap = 0;
  const u32* rep = nullptr;
  u32* gpar = nullptr;
  u32* ctl = nullptr;
  u64* log = nullptr;
  u32* groups = nullptr;
  u32* err = nullptr;
};
// Workgroups the persistent launch uses on `device` (one per CU), 0 where
// cooperative launches are not supported.
u32 persistent_blocks_max(int device);
u32 persistent_threads();
void persistent_rounds(hipStream_t st, const PersistArgs& a, u32 blocks);

}  // namespace kmst
