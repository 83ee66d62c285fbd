// kmst_solver.hpp -- the context behind the C ABI handle: stream, cached
// device workspaces, per-solve statistics and the kernel timer.
#pragma once

#include <new>
#include <string>
#include <vector>

#include "kmst_common.hpp"
#include "kmst_kernels.hpp"

namespace kmst {

// A fixed region of device memory that buffers carve their storage from
// instead of allocating (the context of the no-allocation entry point).
// While open, a buffer that must grow takes the next 256-byte aligned bytes;
// once frozen, a request beyond a buffer's carved size is an error.
struct Arena {
  char* base = nullptr;
  size_t size = 0, used = 0;
  bool frozen = false;
};

// A device allocation that only grows; reused across solves so steady-state
// calls do not allocate.  With an arena it never allocates (see Arena).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  Arena* arena = nullptr;
  void* get(size_t need);
  void release();
};

struct RoundStat {
  u64 n = 0, m = 0, n_next = 0, m_next = 0;
  u32 run_starts = 0;       // of RUN_SAMPLES sampled next-round records (0: not sampled)
  u32 far_starts = 0;       // ... whose source is 16 or more ids from the record before
  bool persistent = false;  // ran inside the persistent launch
  bool twins = false;       // relabelled the src < dst half, both directions written
  double phase_us[4] = {0, 0, 0, 0};  // persistent: bind, hook, jump, relabel
};
struct FilterStat {
  int depth = 0;
  u64 m = 0;
  u64 light = 0;
  u32 pivot = 0;
  bool base = false;
  u64 heavy_kept = 0;
  bool skipped = false;
};
struct SolveStats {
  u64 n = 0, m = 0, mst_edges = 0, total_weight = 0;
  int algorithm = 0;
  bool packed_keys = true;
  bool lazy_keys = false;  // positional keys made by the class scatter, not stored
  u64 max_weight = 0;
  // output order: groups of one smaller endpoint (or position range) longer
  // than MST_GROUP_SMALL keys, and whether one exceeded MST_GROUP_MAX (the
  // whole order field was then sorted)
  u64 mst_long_groups = 0;
  bool mst_resorted = false;
  std::vector<RoundStat> rounds;
  std::vector<FilterStat> filter;
};

extern thread_local std::string g_last_error;

// Runs f, mapping kmst::Error and host allocation failures to a status and
// the thread's last-error message (the error convention of every entry point).
template <typename F>
kmst_status guarded_call(F&& f) {
  try {
    f();
    g_last_error.clear();
    return KMST_OK;
  } catch (const Error& e) {
    g_last_error = e.what();
    return e.status;
  } catch (const std::bad_alloc&) {
    g_last_error = "host allocation failed";
    return KMST_ENOMEM;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return KMST_EDEVICE;
  }
}

// pinned host words: counters in [0, 64), then HOST_STAGE_WORDS of staging
// (vertex samples, or the persistent rounds' control words and log), then
// the sequence word of the last publication (contexts with coherent pinned
// words, which the device writes directly)
constexpr u32 GPAR_SAMPLES = 256;
constexpr u32 PIVOT_SAMPLES = 4095;  // input weights sampled for the top-level pivots
constexpr u32 LEVEL_SAMPLES = 1023;  // weights sampled for a deeper level's pivot
constexpr u32 HOST_SAMPLE_WORD = 64;
constexpr u32 HOST_STAGE_WORDS = 4096;
static_assert(GPAR_SAMPLES <= HOST_STAGE_WORDS && LEVEL_SAMPLES <= HOST_STAGE_WORDS &&
                  PIVOT_SAMPLES + 1 <= HOST_STAGE_WORDS,
              "sample staging");
static_assert(PERSIST_CTL_WORDS + 2 * PERSIST_LOG_TOTAL <= HOST_STAGE_WORDS,
              "persistent-round staging");
constexpr u32 HOST_SEQ_WORD = HOST_SAMPLE_WORD + HOST_STAGE_WORDS;
constexpr size_t HOST_COUNTER_BYTES = (HOST_SEQ_WORD + 64) * 4;

}  // namespace kmst

struct kmst_context {
  explicit kmst_context(int device);
  // A context that allocates nothing: no stream (the caller sets `stream`),
  // counters in the caller's host words (HOST_COUNTER_BYTES), workspaces
  // carved from `arena`.
  kmst_context(int device, kmst::u32* host_words, kmst::Arena* arena);
  ~kmst_context();
  kmst_context(const kmst_context&) = delete;
  kmst_context& operator=(const kmst_context&) = delete;

  int device = 0;
  bool owns_resources = true;  // created the stream and the pinned counters
  hipStream_t stream = nullptr;
  kmst::u32* host_counters = nullptr;  // pinned, HOST_COUNTER_BYTES
  // host_counters is coherent pinned memory the device writes directly
  // (read-backs by publication); otherwise by copy and synchronisation
  bool host_mapped = false;
  kmst::u32 publish_seq = 0;
  bool timing = false;
  kmst::KernelTimer timer;
  kmst::SolveStats stats;

  // record buffers (X, Y ping-pong), strict keys of the input
  kmst::DevBuf xs, xd, xv, ys, yd, yv, val0;
  // vertex arrays
  kmst::DevBuf off, best_val, best_dst, parent, flag, newid, label, rep, rep_next, gpar, seg_start, seg_end,
      mst;
  // kernel scratch
  kmst::DevBuf pieces, scan_scratch, hist, counters, samples, lookback, persist_log,
      part_giant;
  kmst::SortWork sort_work;  // radix sort state over `hist`
  kmst::LookbackWork lb_work;  // compaction look-back state over `lookback`
  kmst_tuning tuning{0, 1, kmst::SORT_WIDE_MIN, kmst::PERSIST_MAX_DEFAULT};  // launch shapes of later solves
  // host-API staging
  kmst::DevBuf in_s, in_d, in_w, out_s, out_d, out_w;
  // generator scratch
  kmst::DevBuf gen_a, gen_b, gen_c, gen_d;
  // autotuner result fingerprint (and the codec's error word)
  kmst::DevBuf tune;
  // difference-coded host input (compression level 1)
  kmst::DevBuf codec_bytes, codec_off, codec_first;

  std::vector<kmst::DevBuf*> all_bufs();
};

namespace kmst {
// The context's tuning onto its sort state (before every solve and sort).
void apply_tuning(kmst_context& C);

// kmst_solve_device for a solve whose vertex ids form ranges [part[r],
// part[r + 1]) (r < count; part is a device array of count + 1 words) that no
// record joins, e.g. a batch of requests: the heavy weight classes then skip
// the records inside the largest component of every range.
kmst_status solve_device_partitioned(kmst_context* ctx, const kmst_dev_edges* in,
                                     const kmst_config* cfg, kmst_dev_mst* out, uint64_t* out_n,
                                     uint64_t* total_weight, const u32* part, u32 count);
}  // namespace kmst
