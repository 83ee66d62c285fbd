/* kmst.h -- C ABI of libkmst, the MI355X-native Boruvka / filter-Boruvka
 * minimum spanning forest library.
 *
 * The two solve entry points replace KaMSTa's algorithm templates
 *   hybridMST::boruvka(InputEdges, const AlgorithmConfig&)
 *       include/algorithms/hybrid_boruvka.hpp:263-301
 *   hybridMST::filter_boruvka(InputEdges, const AlgorithmConfig&)
 *       include/algorithms/hybrid_boruvka_edgefilter.hpp:420-462
 * which the driver calls at benchmarks/mst.cpp:215-228.  Inputs follow their
 * contract: every undirected edge stored in both directions, records sorted
 * by (src, dst), no self-loops (include/util/utils.hpp:391-395).  Unlike the
 * reference, the contract is checked and violations are reported as status
 * codes instead of undefined behaviour or std::abort.
 *
 * Results are canonical: each MST edge once as (min, max, weight), sorted by
 * (min, max).  Ties are broken by the strict key (weight, position of the
 * (min, max) record), which equals the reference check's Kruskal order
 * (weight, src, dst) (include/util/utils.hpp:81-86); the forest is therefore
 * unique and identical for both algorithms and any filter threshold.
 *
 * Every call is blocking.  A context owns device workspaces and is not
 * thread-safe; calls without a context use one default context per device,
 * serialised by a lock.
 */
#ifndef KMST_H
#define KMST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* == hybridMST::WEdge (include/definitions.hpp:66-84): 24 bytes. */
typedef struct kmst_wedge {
  uint64_t src;
  uint64_t dst;
  uint32_t weight;
  uint32_t _pad;
} kmst_wedge;

/* Mirrors AlgorithmConfig (include/algorithms/algorithm_configurations.hpp:28-34)
 * plus the choice of entry point and device. */
typedef enum kmst_algorithm {
  KMST_ALGO_BORUVKA = 0,        /* hybridMST::boruvka         */
  KMST_ALGO_FILTER_BORUVKA = 1  /* hybridMST::filter_boruvka  */
} kmst_algorithm;

typedef struct kmst_config {
  int32_t algorithm;            /* kmst_algorithm */
  int32_t local_preprocessing;  /* LocalPreprocessing: accepted, no effect on one GPU
                                   (the whole graph is local) */
  int32_t compression;          /* Compression: 0 none; 1 the difference codec
                                   (DifferenceEncodedGraph): the host API sends the
                                   records difference-encoded and decodes them on the
                                   device; no effect on kmst_solve_device */
  int32_t device;               /* HIP device ordinal */
  uint64_t filter_threshold;    /* filter-Boruvka base case when m/n0 <= this (default 4) */
} kmst_config;

typedef enum kmst_status {
  KMST_OK = 0,
  KMST_EINVAL = 1,       /* bad argument, self-loop, id >= 2^32-1, too many records */
  KMST_EUNSORTED = 2,    /* records not sorted by (src, dst) */
  KMST_EASYMMETRIC = 3,  /* some record has no reverse twin of equal weight */
  KMST_ENOMEM = 4,       /* device allocation failed / admission refused */
  KMST_EDEVICE = 5,      /* HIP runtime error or a kernel's bounded wait expired */
  KMST_EBUSY = 6         /* serving: queue full */
} kmst_status;

/* Device-resident input, structure of arrays (12 bytes per directed record). */
typedef struct kmst_dev_edges {
  const uint32_t* src;
  const uint32_t* dst;
  const uint32_t* weight;
  uint64_t m_directed;   /* number of records (2 x undirected edges) */
  uint64_t n;            /* vertex ids are < n */
} kmst_dev_edges;

/* Device-resident output, capacity >= n - 1 entries each. */
typedef struct kmst_dev_mst {
  uint32_t* src;         /* min endpoint */
  uint32_t* dst;         /* max endpoint */
  uint32_t* weight;
  uint64_t capacity;
} kmst_dev_mst;

typedef struct kmst_context kmst_context;

/* Fills cfg with the reference defaults (AlgorithmConfig{}, threshold 4). */
void kmst_config_default(kmst_config* cfg);

kmst_status kmst_context_create(int device, kmst_context** out);
void kmst_context_destroy(kmst_context* ctx);
/* Frees the context's cached device workspaces. */
void kmst_context_trim(kmst_context* ctx);
/* Sizes the context's workspaces for solves of up to m_directed records and
 * n vertices, so that such solves allocate no device memory. */
kmst_status kmst_context_reserve(kmst_context* ctx, uint64_t m_directed, uint64_t n);

/* Host edge list in, host MST out.  `out` must hold at least
 * (number of distinct vertices - 1) entries; m_directed / 2 always suffices.
 * Replaces hybridMST::boruvka / filter_boruvka (selected by cfg->algorithm). */
kmst_status kmst_boruvka(const kmst_wedge* in, uint64_t m_directed,
                         const kmst_config* cfg, kmst_wedge* out,
                         uint64_t* out_n, uint64_t* total_weight);
kmst_status kmst_filter_boruvka(const kmst_wedge* in, uint64_t m_directed,
                                const kmst_config* cfg, kmst_wedge* out,
                                uint64_t* out_n, uint64_t* total_weight);

/* Device edge list in, device MST out; the headline timing path.  The input
 * must stay valid for the call; nothing is copied to the host except counts. */
kmst_status kmst_solve_device(kmst_context* ctx, const kmst_dev_edges* in,
                              const kmst_config* cfg, kmst_dev_mst* out,
                              uint64_t* out_n, uint64_t* total_weight);

/* kmst_solve_device without device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (serving loops and
 * harnesses that forbid hipMalloc).  The workspaces come from a fixed device
 * array inside the library, one per GPU, held until the process ends
 * (80 MiB); solves of more than 2^20 records or 2^19 vertices are refused
 * with KMST_EINVAL.  `stream` is the hipStream_t to run on (NULL: the default
 * stream); cfg->device selects the GPU.  The call makes no hipMalloc,
 * hipFree, pinned-memory or stream allocation.  Every call on a GPU shares
 * its array, which keeps the last call's contents; calls are blocking and
 * serialised per GPU (overlapping calls on one GPU would race on the array). */
kmst_status kmst_solve_device_noalloc(const kmst_dev_edges* in, const kmst_config* cfg,
                                      kmst_dev_mst* out, uint64_t* out_n,
                                      uint64_t* total_weight, void* stream);

/* Per-kernel HIP-event timing of later solves on this context (measurement
 * hook; off by default).  Totals appear in kmst_context_stats_json. */
void kmst_context_set_timing(kmst_context* ctx, int on);

/* Difference encoding of compression level 1 on the host
 * (DifferenceEncodedGraph::encode_edges / decode_edge,
 * include/datastructures/compression/difference_encoded_graph.hpp:166-212, with the
 * varints of include/datastructures/variable_length_encoding.hpp:6-160): records in
 * chunks of chunk_records; chunk k keeps its first record raw in
 * first[3k..3k+2] = (src, dst, weight), the others as 7-bit varints of
 * (src - previous src mod 2^32, dst, weight); offsets[k] is where chunk k's
 * varints start, offsets[chunks] the length.  With bytes == NULL only
 * *bytes_used is set. */
kmst_status kmst_encode_host(const uint32_t* src, const uint32_t* dst, const uint32_t* weight,
                             uint64_t m, uint32_t chunk_records, uint8_t* bytes, uint64_t bytes_cap,
                             uint64_t* offsets, uint32_t* first, uint64_t* bytes_used);

/* ---- Tuning (no counterpart in the reference, whose only knob on this path
 * is AlgorithmConfig::filter_threshold). ------------------------------------
 * Launch-shape choices of a context's solves.  They change speed only: every
 * tuning returns the same MST. */
typedef struct kmst_tuning {
  uint32_t sort_threads;     /* threads per radix scatter tile (tiles of 16 records per
                                thread): 256, 512, or 0 = by size (see sort_wide_min) */
  uint32_t balanced_digits;  /* 1: a b-bit key sorts in ceil(b/8) passes of nearly equal
                                digit widths; 0: 8-bit digits, the last pass takes the rest */
  uint64_t sort_wide_min;    /* with sort_threads 0: sorts of at least this many records
                                use 512-thread tiles */
  uint64_t persistent_max;   /* rounds on at most this many live records (with at most
                                twice as many vertex ids, plus 1024, and at most 32 records
                                per vertex id or at most 2^14 records) run to the end in
                                one persistent cooperative launch; 0 = never */
} kmst_tuning;

/* The defaults (0, 1, 2^20, 2^18). */
void kmst_tuning_default(kmst_tuning* t);
kmst_status kmst_context_set_tuning(kmst_context* ctx, const kmst_tuning* t);
kmst_status kmst_context_get_tuning(const kmst_context* ctx, kmst_tuning* t);

/* Autotuner: solves the device-resident input `in` with cfg->algorithm under
 * candidate settings -- filter thresholds {2, 3, 4, 6} (filter-Boruvka only),
 * then sort tile threads {by size, 256, 512}, then digit widths {balanced,
 * 8-bit}, then persistent_max {0, 2^16, 2^18, 2^20}, each stage keeping the
 * best so far -- timing each candidate by the
 * median of `repeats` solves (at least 1) after one warm-up solve.  Every
 * candidate must return the same MST (count, total weight and an
 * order-independent hash of the edges), else KMST_EDEVICE.  Leaves the
 * fastest tuning set on the context and copied to *tuning, its filter
 * threshold in cfg->filter_threshold, its median time (ms) in *best_ms and
 * the number of candidates timed in *candidates.  Uses the context's own
 * output staging; `in` must satisfy the kmst_solve_device contract. */
kmst_status kmst_autotune(kmst_context* ctx, const kmst_dev_edges* in, kmst_config* cfg,
                          int32_t repeats, kmst_tuning* tuning, double* best_ms,
                          uint32_t* candidates);

/* Thread-local message for the last non-OK status. */
const char* kmst_last_error(void);

/* Per-phase statistics of the last solve on a context (rounds, records,
 * vertices, algorithmic bytes per kernel class).  Returns the number of
 * bytes written to buf (a JSON object), or the size needed if buf is small. */
size_t kmst_context_stats_json(const kmst_context* ctx, char* buf, size_t cap);

/* ---- Input side: deterministic generators (KaGen replacement). ----------
 * GNM: m distinct undirected pairs on n vertices, the first m distinct
 * pairs of a counter-based hash sequence; weight = 1 + H(seed, min, max) mod
 * max_weight.  Output is symmetric and sorted by (src, dst).  The device
 * and host versions produce identical arrays. */
kmst_status kmst_gen_gnm_device(kmst_context* ctx, uint64_t n, uint64_t m,
                                uint32_t max_weight, uint64_t seed,
                                uint32_t* src, uint32_t* dst, uint32_t* weight);
kmst_status kmst_gen_gnm_host(uint64_t n, uint64_t m, uint32_t max_weight,
                              uint64_t seed, uint32_t* src, uint32_t* dst,
                              uint32_t* weight);

/* RGG2D: n points uniform in the unit square (32-bit fixed point), edge iff
 * the integer squared distance <= r^2, r = sqrt(avg_degree / (pi n)).
 * RHG: threshold random hyperbolic graph with power-law exponent gamma.
 * Both write symmetric records sorted by (src, dst), weights
 * 1 + H(seed, min, max) mod max_weight.  The record count is known only after
 * generation: *m_directed always receives it, and KMST_EINVAL is returned
 * (nothing written) when capacity is smaller. */
kmst_status kmst_gen_rgg2d_host(uint64_t n, double avg_degree, uint32_t max_weight,
                                uint64_t seed, uint32_t* src, uint32_t* dst, uint32_t* weight,
                                uint64_t capacity, uint64_t* m_directed);
kmst_status kmst_gen_rhg_host(uint64_t n, double avg_degree, double gamma, uint32_t max_weight,
                              uint64_t seed, uint32_t* src, uint32_t* dst, uint32_t* weight,
                              uint64_t capacity, uint64_t* m_directed);
/* The same graphs generated on the device into device arrays, identical to
 * the host versions record for record (RHG's floating point is built from
 * IEEE add, multiply, divide and square root only, evaluated alike on both
 * sides).  Same capacity rule: *m_directed always receives the record
 * count, and KMST_EINVAL is returned (nothing written) when capacity is
 * smaller.  Replaces the in-process KaGen generation of the reference driver
 * (benchmarks/mst_benchmarks.hpp:116-199). */
kmst_status kmst_gen_rgg2d_device(kmst_context* ctx, uint64_t n, double avg_degree,
                                  uint32_t max_weight, uint64_t seed, uint32_t* src,
                                  uint32_t* dst, uint32_t* weight, uint64_t capacity,
                                  uint64_t* m_directed);
kmst_status kmst_gen_rhg_device(kmst_context* ctx, uint64_t n, double avg_degree, double gamma,
                                uint32_t max_weight, uint64_t seed, uint32_t* src, uint32_t* dst,
                                uint32_t* weight, uint64_t capacity, uint64_t* m_directed);

/* ---- Batched serving ----------------------------------------------------
 * A persistent request loop on one GPU: requests (host SoA edge lists with
 * the same contract as kmst_boruvka) are queued by kmst_server_submit,
 * admitted into batches by device-memory footprint, solved as one segmented
 * device solve (vertex-id offsets keep the requests disjoint), and completed
 * asynchronously.  Each request's result is identical to solving it alone.
 * Input arrays must stay valid until the ticket completes; pinned host
 * memory (hipHostMalloc) uploads at full link speed.  Output arrays receive
 * (min, max, weight) records sorted by (min, max) in the request's own ids.
 * No counterpart in the reference, which solves one graph per call
 * (benchmarks/mst.cpp:198-275). */
typedef struct kmst_server kmst_server;

typedef struct kmst_server_config {
  int32_t device;
  int32_t algorithm;             /* kmst_algorithm used for every batch */
  uint64_t filter_threshold;
  uint64_t max_batch_records;    /* admission bound per batch (directed records; default 2^28) */
  uint64_t max_batch_requests;   /* 1 = one graph per launch */
  uint64_t hbm_budget_bytes;     /* 0: half of the free device memory */
  uint32_t queue_capacity;       /* queued requests before submit returns KMST_EBUSY */
  uint32_t buffer_sets;          /* batches in flight (upload / compute / delivery) */
  kmst_tuning tuning;            /* launch shapes of the server's solves (kmst_tuning rules) */
} kmst_server_config;

typedef struct kmst_request_result {
  uint64_t ticket;
  int32_t status;                /* kmst_status of this request */
  int32_t done;                  /* 1 once complete */
  uint64_t mst_edges;
  uint64_t total_weight;
  uint64_t batch_requests;       /* size of the batch it was solved in */
  uint64_t submit_ns, admit_ns, start_ns, solved_ns, done_ns;  /* steady clock */
} kmst_request_result;

void kmst_server_config_default(kmst_server_config* cfg);
kmst_status kmst_server_create(const kmst_server_config* cfg, kmst_server** out);
/* KMST_ENOMEM: the request alone exceeds the admission budget;
 * KMST_EBUSY: the queue is full. */
kmst_status kmst_server_submit(kmst_server* s, const uint32_t* src, const uint32_t* dst,
                               const uint32_t* weight, uint64_t m_directed, uint64_t n,
                               uint32_t* out_src, uint32_t* out_dst, uint32_t* out_weight,
                               uint64_t out_capacity, uint64_t* ticket);
/* Copies the ticket's state; a completed ticket is released by this call. */
kmst_status kmst_server_poll(kmst_server* s, uint64_t ticket, kmst_request_result* res);
kmst_status kmst_server_wait(kmst_server* s, uint64_t ticket, kmst_request_result* res);
kmst_status kmst_server_counters(kmst_server* s, uint64_t* batches, uint64_t* requests);
/* Waits for every submitted request to complete, then frees the server. */
void kmst_server_destroy(kmst_server* s);

/* ---- Kernel-level entry points (device pointers), for per-kernel tests. */
/* K1: per-vertex min of val over records grouped by src (sorted);
 * best_val[v] = UINT64_MAX and best_dst[v] untouched for empty vertices. */
kmst_status kmst_k_segmented_argmin(kmst_context* ctx, const uint32_t* src,
                                    const uint32_t* dst, const uint64_t* val,
                                    uint64_t m, uint64_t n, uint64_t* best_val,
                                    uint32_t* best_dst);
/* K1 with packed keys made from the weights: val = weight << 2 vbits |
 * min(src, dst) << vbits | max(src, dst) (vbits <= 31, weight < 2^(64 - 2 vbits));
 * otherwise as kmst_k_segmented_argmin. */
kmst_status kmst_k_segmented_argmin_packed(kmst_context* ctx, const uint32_t* src,
                                           const uint32_t* dst, const uint32_t* weight,
                                           uint32_t vbits, uint64_t m, uint64_t n,
                                           uint64_t* best_val, uint32_t* best_dst);
/* K6: stable LSD radix sort of (src, dst, val) by (src, dst) over
 * `bits` bits of each endpoint, in place (uses context scratch). */
kmst_status kmst_k_sort_pairs(kmst_context* ctx, uint32_t* src, uint32_t* dst,
                              uint64_t* val, uint64_t m, uint32_t bits);
/* K3: parent[v] <- root of v, in place. */
kmst_status kmst_k_pointer_jump(kmst_context* ctx, uint32_t* parent, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif /* KMST_H */
