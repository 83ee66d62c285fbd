// kmst_serve.hip -- batched MST serving: a persistent request loop that
// admits MST solves of different sizes as they arrive, batches them into one
// device-resident solve, and overlaps uploads, compute and result delivery.
//
// Batching: requests are concatenated with vertex-id offsets; because each
// request occupies its own id range, the concatenation stays sorted by
// (src, dst), no edge joins two requests, and the strict key order inside a
// request is unchanged (packed keys shift both endpoints by the same offset;
// positional keys are positions in the concatenation, monotone within a
// request).  The batch's minimum spanning forest is therefore the disjoint
// union of the requests' forests, edge for edge, and each request's result is
// identical alone and in any mix.
//
// Pipeline (three threads, rotating device buffer sets):
//   admission  pops queued requests while their footprint fits the batch
//              budget, uploads them (one async copy per array, pinned host
//              memory goes at link speed) on the copy stream;
//   compute    adds the vertex-id offsets and runs the solve;
//   delivery   splits the batch MST by request on the device (subtracting
//              offsets), writing it straight into output arrays the device
//              can write (device memory, pinned host memory) and copying the
//              rest, and completes the tickets.
// Requests that arrive while a batch runs are admitted into the next one as
// soon as a buffer set frees up (continuous batching at batch granularity).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "kmst_kernels.hpp"
#include "kmst_solver.hpp"

namespace kmst {

static u64 now_ns() {
  return static_cast<u64>(std::chrono::duration_cast<std::chrono::nanoseconds>(
                              std::chrono::steady_clock::now().time_since_epoch())
                              .count());
}

// The request of record i: the last r with r_off[r] <= i.
__device__ __forceinline__ u32 request_of(u64 i, const u64* __restrict__ r_off, u32 count) {
  u32 lo = 0, hi = count;
  while (hi - lo > 1) {
    const u32 mid = (lo + hi) >> 1;
    if (r_off[mid] <= i) lo = mid; else hi = mid;
  }
  return lo;
}

// The request of x for a whole wave (x wave-uniform): the last r < count with
// r_off[r] <= x.  Each step probes 64 evenly spaced entries of the current
// range at once, so 4096 requests take two dependent loads, not twelve.
__device__ __forceinline__ u32 wave_request_of(u64 x, const u64* __restrict__ r_off, u32 count) {
  const u32 lane = lane_id();
  u32 lo = 0, hi = count;
  while (hi - lo > 1) {
    const u32 step = (hi - lo + WAVE - 1) / WAVE;
    const u32 p = lo + lane * step;
    const u64 ok = __ballot(p < hi && r_off[p] <= x);  // lanes 0..k, r_off[lo] <= x
    const u32 k = 63u - static_cast<u32>(__clzll(static_cast<long long>(ok)));
    lo += k * step;
    hi = min(hi, lo + step);
  }
  return lo;
}

// Adds each request's vertex offset to its records; with gather (three
// device addresses per request: src, dst, weight; 0 when the request's
// records were copied into the batch arrays) it first reads the request's
// records from its own arrays and writes all three batch arrays.  A wave
// takes chunks of OFF_CHUNK records; a chunk inside one request (almost all
// of them) needs one request lookup for the whole wave, one that crosses a
// request boundary looks up every record.  With err, a record with an id at
// or past its request's vertex count (it would join another request's id
// range) raises ERR_INVALID; v_off then has count + 1 entries.
constexpr u32 OFF_CHUNK = WAVE * 4;
u32 offset_grid(u64 m) {
  const u64 per_block = u64(OFF_CHUNK) * (BLOCK / WAVE);
  return static_cast<u32>(std::max<u64>(1, std::min<u64>((m + per_block - 1) / per_block, 8192)));
}

__device__ __forceinline__ void offset_one(u64 j, u32 r, u32* __restrict__ src,
                                           u32* __restrict__ dst, u32* __restrict__ w,
                                           const u64* __restrict__ r_off,
                                           const u32* __restrict__ v_off,
                                           const u64* __restrict__ gather, u32* err) {
  const u32 o = v_off[r];
  u32 s, d;
  const u64* g = gather ? gather + 3 * u64(r) : nullptr;
  if (g && g[0]) {
    const u64 l = j - r_off[r];
    s = reinterpret_cast<const u32*>(g[0])[l];
    d = reinterpret_cast<const u32*>(g[1])[l];
    w[j] = reinterpret_cast<const u32*>(g[2])[l];
  } else {
    s = src[j];
    d = dst[j];
  }
  if (err && (s >= v_off[r + 1] - o || d >= v_off[r + 1] - o)) atomicOr(err, ERR_INVALID);
  src[j] = s + o;
  dst[j] = d + o;
}

__global__ __launch_bounds__(BLOCK) void batch_offset_kernel(
    u32* __restrict__ src, u32* __restrict__ dst, u32* __restrict__ w, u64 m,
    const u64* __restrict__ r_off, const u32* __restrict__ v_off, const u64* __restrict__ gather,
    u32 count, u32* err) {
  const u64 wave = (u64(blockIdx.x) * BLOCK + threadIdx.x) / WAVE;
  const u64 nwaves = u64(gridDim.x) * (BLOCK / WAVE);
  const u32 lane = lane_id();
  for (u64 c = wave; c * OFF_CHUNK < m; c += nwaves) {
    const u64 first = c * OFF_CHUNK;
    const u64 last = min(first + OFF_CHUNK, m) - 1;
    const u32 r0 = wave_request_of(first, r_off, count);
    const bool one = r0 + 1 >= count || r_off[r0 + 1] > last;  // chunk inside request r0
    const u64* g = gather ? gather + 3 * u64(r0) : nullptr;
    if (one && !(g && g[0])) {
      // in place: one 16-byte vector of src and of dst per lane when aligned
      const u64 i = first + u64(lane) * 4;
      const bool aligned =
          ((reinterpret_cast<uintptr_t>(src + i) | reinterpret_cast<uintptr_t>(dst + i)) & 15) == 0;
      if (i + 3 <= last && aligned) {
        const u32 o = v_off[r0];
        uint4 s4 = *reinterpret_cast<const uint4*>(src + i);
        uint4 d4 = *reinterpret_cast<const uint4*>(dst + i);
        if (err) {
          const u32 nr = v_off[r0 + 1] - o;
          if (max(max(max(s4.x, s4.y), max(s4.z, s4.w)), max(max(d4.x, d4.y), max(d4.z, d4.w))) >= nr)
            atomicOr(err, ERR_INVALID);
        }
        s4.x += o; s4.y += o; s4.z += o; s4.w += o;
        d4.x += o; d4.y += o; d4.z += o; d4.w += o;
        *reinterpret_cast<uint4*>(src + i) = s4;
        *reinterpret_cast<uint4*>(dst + i) = d4;
      } else {
        for (u64 j = i; j < i + 4 && j <= last; ++j)
          offset_one(j, r0, src, dst, w, r_off, v_off, gather, err);
      }
    } else {
      // gathered or crossing a request boundary: lane-interleaved records
      for (u64 j = first + lane; j <= last; j += WAVE)
        offset_one(j, one ? r0 : request_of(j, r_off, count), src, dst, w, r_off, v_off, gather,
                   err);
    }
  }
}

// bounds[r] = first MST edge of request r (edges sorted by min endpoint,
// request r owns vertex ids [v_off[r], v_off[r+1])); bounds[count] = k.
__global__ __launch_bounds__(BLOCK) void batch_bounds_kernel(const u32* __restrict__ out_s, u64 k,
                                                             const u32* __restrict__ v_off,
                                                             u32 count, u64* __restrict__ bounds) {
  const u32 r = blockIdx.x * BLOCK + threadIdx.x;
  if (r > count) return;
  if (r == count) {
    bounds[r] = k;
    return;
  }
  const u32 x = v_off[r];
  u64 lo = 0, hi = k;
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    if (out_s[mid] < x) lo = mid + 1; else hi = mid;
  }
  bounds[r] = lo;
}

// Back to request-local ids; per-request weight sums.  Block (r, j) of a
// (requests x split) grid takes every split-th 256-edge step of request r and
// adds its weight sum to totals[r] (zeroed by the caller).
// table holds four words per request: the device-usable addresses of its
// three output arrays (0: not device-usable) and their capacity.  A request
// whose arrays are device-usable and large enough gets its slice written there
// directly; the others are rewritten in place for the host to copy.
__global__ __launch_bounds__(BLOCK) void batch_unoffset_kernel(u32* __restrict__ out_s,
                                                               u32* __restrict__ out_d,
                                                               const u32* __restrict__ out_w,
                                                               const u64* __restrict__ bounds,
                                                               const u32* __restrict__ v_off,
                                                               const u64* __restrict__ table,
                                                               u64* __restrict__ totals) {
  __shared__ u64 lds[WAVES];
  const u32 r = blockIdx.x;
  const u64 b0 = bounds[r], b1 = bounds[r + 1];
  const u64 step = u64(gridDim.y) * BLOCK;
  const u32 o = v_off[r];
  const u64* t = table + 4 * u64(r);
  u32* ds = reinterpret_cast<u32*>(t[0]);
  u32* dd = reinterpret_cast<u32*>(t[1]);
  u32* dw = reinterpret_cast<u32*>(t[2]);
  const bool direct = ds && dd && dw && b1 - b0 <= t[3];
  u64 sum = 0;
  for (u64 i = b0 + u64(blockIdx.y) * BLOCK + threadIdx.x; i < b1; i += step) {
    const u32 w = out_w[i];
    if (direct) {
      ds[i - b0] = out_s[i] - o;
      dd[i - b0] = out_d[i] - o;
      dw[i - b0] = w;
    } else {
      out_s[i] -= o;
      out_d[i] -= o;
    }
    sum += w;
  }
  u64 total;
  block_exclusive_sum(sum, total, lds);
  if (threadIdx.x == 0 && total)
    atomicAdd(reinterpret_cast<unsigned long long*>(totals + r),
              static_cast<unsigned long long>(total));
}

struct Request {
  u64 ticket = 0;
  const u32 *src = nullptr, *dst = nullptr, *w = nullptr;
  u64 m = 0, n = 0;
  u32 *out_src = nullptr, *out_dst = nullptr, *out_w = nullptr;
  u64 out_cap = 0;
  bool out_direct = false;  // the output arrays are device-usable: written by the kernel
  bool in_device = false;   // the input arrays are this device's memory: read in place
  kmst_request_result res{};
};

struct BufferSet {
  DevBuf s, d, w, out_s, out_d, out_w, r_off, v_off, bounds, table, flag, gather;
  std::vector<u64> h_gather;  // src, dst, weight addresses per request (0: copied)
  bool any_gather = false;
  u32* h_flag = nullptr;    // pinned: the offset pass's error word
  std::vector<u64> h_r_off;
  std::vector<u32> h_v_off;
  u64* h_bounds = nullptr;  // pinned: 2 x (requests + 1) words (bounds, totals)
  u64 h_bounds_cap = 0;
  u64* h_table = nullptr;   // pinned: 4 words per request (output addresses, capacity)
  u64 h_table_cap = 0;
  hipEvent_t uploaded = nullptr;
  std::vector<std::shared_ptr<Request>> reqs;
  u64 m = 0, n = 0;
  u64 mst_n = 0, mst_w = 0;
  kmst_status status = KMST_OK;
  bool solved_individually = false;  // results already written per request
  std::string error;
};

}  // namespace kmst

using namespace kmst;

struct kmst_server {
  kmst_server_config cfg{};
  std::unique_ptr<kmst_context> ctx;       // compute context (its own stream)
  static constexpr int UPLOAD_STREAMS = 4;  // concurrent DMA engines for uploads
  hipStream_t copy_stream = nullptr;       // uploads (first of up_streams)
  hipStream_t up_streams[UPLOAD_STREAMS] = {};
  hipEvent_t up_events[UPLOAD_STREAMS] = {};
  hipStream_t down_stream = nullptr;       // result downloads
  std::vector<std::unique_ptr<BufferSet>> sets;
  std::mutex mu;
  std::condition_variable cv;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<u64, std::shared_ptr<Request>> pending;  // not yet collected
  std::deque<BufferSet*> free_sets, ready_sets, done_sets;
  u64 next_ticket = 1;
  bool stop = false;
  std::thread admit_th, compute_th, deliver_th;
  u64 batches = 0, batched_requests = 0;

  void admit_loop();
  void compute_loop();
  void deliver_loop();
  void solve_alone(BufferSet& b);
};

namespace {

constexpr u64 BYTES_PER_RECORD = 3 * 12 + 48;  // upload sets + solver workspace
constexpr u64 BYTES_PER_VERTEX = 64;

u64 footprint(u64 m, u64 n) { return m * BYTES_PER_RECORD + n * BYTES_PER_VERTEX; }

// True when kernels on `device` can write p: memory of that device, or pinned
// host memory the device maps (the same address under unified addressing).
// Pageable host memory, managed memory and other devices' memory are not.
// True when p is memory of `device` itself (kernels read it in place).
bool device_resident(const void* p, int device) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return !a.isManaged && a.type == hipMemoryTypeDevice && a.device == device;
}

bool device_writable(const void* p, int device) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory may report an error: clear it
    return false;
  }
  if (a.isManaged) return false;
  if (a.type == hipMemoryTypeDevice) return a.device == device;
  return a.type == hipMemoryTypeHost && a.devicePointer != nullptr;
}

template <typename F>
kmst_status guard(F&& f) {
  try {
    f();
    return KMST_OK;
  } catch (const kmst::Error& e) {
    kmst::g_last_error = e.what();
    return e.status;
  } catch (const std::exception& e) {
    kmst::g_last_error = e.what();
    return KMST_EDEVICE;
  }
}

}  // namespace

void kmst_server::admit_loop() {
  (void)hipSetDevice(cfg.device);
  while (true) {
    BufferSet* b = nullptr;
    std::vector<std::shared_ptr<Request>> take;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return stop || (!queue.empty() && !free_sets.empty()); });
      if (stop) return;  // set only after every request has completed
      b = free_sets.front();
      free_sets.pop_front();
      u64 m = 0, n = 0;
      while (!queue.empty() && take.size() < cfg.max_batch_requests) {
        auto& r = queue.front();
        if (!take.empty() && (m + r->m > cfg.max_batch_records ||
                              n + r->n >= (u64{1} << 31)))
          break;
        m += r->m;
        n += r->n;
        take.push_back(r);
        queue.pop_front();
      }
    }
    const u64 t_admit = now_ns();
    u64 m = 0, n = 0;
    b->h_r_off.assign(take.size() + 1, 0);
    b->h_v_off.assign(take.size() + 1, 0);
    for (size_t i = 0; i < take.size(); ++i) {
      b->h_r_off[i] = m;
      b->h_v_off[i] = static_cast<u32>(n);
      m += take[i]->m;
      n += take[i]->n;
      take[i]->res.admit_ns = t_admit;
    }
    b->h_r_off[take.size()] = m;
    b->h_v_off[take.size()] = static_cast<u32>(n);
    b->m = m;
    b->n = n;
    b->reqs = std::move(take);
    b->status = KMST_OK;
    b->solved_individually = false;
    b->error.clear();
    const kmst_status st = guard([&] {
      const u64 mm = m ? m : 1, nn = n ? n : 1;
      u32* ds = static_cast<u32*>(b->s.get(mm * 4));
      u32* dd = static_cast<u32*>(b->d.get(mm * 4));
      u32* dw = static_cast<u32*>(b->w.get(mm * 4));
      b->out_s.get(nn * 4);
      b->out_d.get(nn * 4);
      b->out_w.get(nn * 4);
      u64* dro = static_cast<u64*>(b->r_off.get((b->reqs.size() + 1) * 8));
      u32* dvo = static_cast<u32*>(b->v_off.get(b->reqs.size() * 4 + 4));
      b->bounds.get(2 * (b->reqs.size() + 1) * 8);
      if (b->h_bounds_cap < 2 * (b->reqs.size() + 1)) {
        if (b->h_bounds) (void)hipHostFree(b->h_bounds);
        b->h_bounds = nullptr;
        const u64 cap = 2 * (b->reqs.size() + 1) + 1024;
        KMST_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_bounds), cap * 8, hipHostMallocDefault));
        b->h_bounds_cap = cap;
      }
      // One copy per run of requests whose host arrays are adjacent in
      // memory (and so are their device slots), spread over several
      // streams so that more than one DMA engine moves data.
      int next_stream = 0;
      auto upload = [&](u32* dst_dev, const u32* src_host, u64 count) {
        if (count == 0) return;
        // host (pinned or pageable) or device-resident request arrays
        KMST_HIP(hipMemcpyAsync(dst_dev, src_host, count * 4, hipMemcpyDefault,
                                up_streams[next_stream]));
        next_stream = (next_stream + 1) % UPLOAD_STREAMS;
      };
      for (int a = 0; a < 3; ++a) {
        u32* dbase = a == 0 ? ds : (a == 1 ? dd : dw);
        const u32* run_src = nullptr;
        u64 run_off = 0, run_len = 0;
        for (size_t i = 0; i < b->reqs.size(); ++i) {
          const Request& r = *b->reqs[i];
          const u32* h = a == 0 ? r.src : (a == 1 ? r.dst : r.w);
          // device-resident requests of a multi-request batch are gathered by
          // the offset pass instead (a single request is solved in place)
          if (r.m == 0 || (r.in_device && b->reqs.size() > 1)) continue;
          if (run_src && h == run_src + run_len && b->h_r_off[i] == run_off + run_len &&
              run_len + r.m <= (u64{1} << 24)) {
            run_len += r.m;
          } else {
            upload(dbase + run_off, run_src, run_len);
            run_src = h;
            run_off = b->h_r_off[i];
            run_len = r.m;
          }
        }
        upload(dbase + run_off, run_src, run_len);
      }
      for (int q = 1; q < UPLOAD_STREAMS; ++q) {
        KMST_HIP(hipEventRecord(up_events[q], up_streams[q]));
        KMST_HIP(hipStreamWaitEvent(copy_stream, up_events[q], 0));
      }
      KMST_HIP(hipMemcpyAsync(dro, b->h_r_off.data(), b->h_r_off.size() * 8, hipMemcpyHostToDevice,
                              copy_stream));
      KMST_HIP(hipMemcpyAsync(dvo, b->h_v_off.data(), b->h_v_off.size() * 4, hipMemcpyHostToDevice,
                              copy_stream));
      b->any_gather = false;
      if (b->reqs.size() > 1) {
        b->h_gather.assign(3 * b->reqs.size(), 0);
        for (size_t i = 0; i < b->reqs.size(); ++i) {
          const Request& r = *b->reqs[i];
          if (!r.in_device || r.m == 0) continue;
          b->h_gather[3 * i] = reinterpret_cast<u64>(r.src);
          b->h_gather[3 * i + 1] = reinterpret_cast<u64>(r.dst);
          b->h_gather[3 * i + 2] = reinterpret_cast<u64>(r.w);
          b->any_gather = true;
        }
        if (b->any_gather) {
          u64* dg = static_cast<u64*>(b->gather.get(b->h_gather.size() * 8));
          KMST_HIP(hipMemcpyAsync(dg, b->h_gather.data(), b->h_gather.size() * 8,
                                  hipMemcpyHostToDevice, copy_stream));
        }
      }
      KMST_HIP(hipEventRecord(b->uploaded, copy_stream));
      // the host vectors above must outlive the copies
      KMST_HIP(hipEventSynchronize(b->uploaded));
    });
    if (st != KMST_OK) {
      b->status = st;
      b->error = kmst::g_last_error;
    }
    {
      std::lock_guard<std::mutex> lk(mu);
      ready_sets.push_back(b);
    }
    cv.notify_all();
  }
}

void kmst_server::solve_alone(BufferSet& b) {
  // a failed batch is re-solved request by request so that only the request
  // that breaks the input contract reports the error
  for (size_t i = 0; i < b.reqs.size(); ++i) {
    Request& r = *b.reqs[i];
    const u64 off = b.h_r_off[i];
    kmst_dev_edges e{static_cast<u32*>(b.s.p) + off, static_cast<u32*>(b.d.p) + off,
                     static_cast<u32*>(b.w.p) + off, r.m, r.n};
    kmst_dev_mst o{static_cast<u32*>(b.out_s.p), static_cast<u32*>(b.out_d.p),
                   static_cast<u32*>(b.out_w.p), std::max<u64>(r.n, 1)};
    kmst_config c;
    kmst_config_default(&c);
    c.algorithm = cfg.algorithm;
    c.filter_threshold = cfg.filter_threshold;
    c.device = cfg.device;
    // undo the batch offset of this request's records before solving alone
    const u32 vo = b.h_v_off[i];
    if (vo && r.m) {
      const u64 rr[2] = {0, r.m};
      const u32 neg = static_cast<u32>(0u - vo);
      u64* dro = static_cast<u64*>(b.r_off.p);
      u32* dvo = static_cast<u32*>(b.v_off.p);
      (void)hipMemcpyAsync(dro, rr, 16, hipMemcpyHostToDevice, ctx->stream);
      (void)hipMemcpyAsync(dvo, &neg, 4, hipMemcpyHostToDevice, ctx->stream);
      batch_offset_kernel<<<offset_grid(r.m), BLOCK, 0, ctx->stream>>>(
          const_cast<u32*>(e.src), const_cast<u32*>(e.dst), const_cast<u32*>(e.weight), r.m, dro,
          dvo, nullptr, 1, nullptr);
      (void)hipStreamSynchronize(ctx->stream);
    }
    u64 cnt = 0, tw = 0;
    r.res.status = kmst_solve_device(ctx.get(), &e, &c, &o, &cnt, &tw);
    if (r.res.status == KMST_OK) {
      r.res.mst_edges = cnt;
      r.res.total_weight = tw;
      if (cnt > r.out_cap) {
        r.res.status = KMST_EINVAL;
      } else if (cnt) {
        (void)hipMemcpy(r.out_src, o.src, cnt * 4, hipMemcpyDefault);
        (void)hipMemcpy(r.out_dst, o.dst, cnt * 4, hipMemcpyDefault);
        (void)hipMemcpy(r.out_w, o.weight, cnt * 4, hipMemcpyDefault);
      }
    }
  }
}

void kmst_server::compute_loop() {
  (void)hipSetDevice(cfg.device);
  while (true) {
    BufferSet* b = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return stop || !ready_sets.empty(); });
      if (ready_sets.empty()) return;
      b = ready_sets.front();
      ready_sets.pop_front();
    }
    const u64 t0 = now_ns();
    for (auto& r : b->reqs) r->res.start_ns = t0;
    if (b->status == KMST_OK && b->m > 0) {
      kmst_config c;
      kmst_config_default(&c);
      c.algorithm = cfg.algorithm;
      c.filter_threshold = cfg.filter_threshold;
      c.device = cfg.device;
      const kmst_status st = guard([&] {
        if (b->reqs.size() > 1) {
          // the offset pass also checks that no record leaves its request's id
          // range (the partitioned solve relies on it); its error word is
          // read back after the solve
          u32* dflag = static_cast<u32*>(b->flag.get(4));
          if (!b->h_flag)
            KMST_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_flag), 4, hipHostMallocDefault));
          fill_u32(ctx->stream, dflag, 1, 0);
          batch_offset_kernel<<<offset_grid(b->m), BLOCK, 0, ctx->stream>>>(
              static_cast<u32*>(b->s.p), static_cast<u32*>(b->d.p), static_cast<u32*>(b->w.p),
              b->m, static_cast<const u64*>(b->r_off.p), static_cast<const u32*>(b->v_off.p),
              b->any_gather ? static_cast<const u64*>(b->gather.p) : nullptr,
              static_cast<u32>(b->reqs.size()), dflag);
          KMST_HIP(hipGetLastError());
          KMST_HIP(hipMemcpyAsync(b->h_flag, dflag, 4, hipMemcpyDeviceToHost, ctx->stream));
        }
        kmst_dev_edges e{static_cast<u32*>(b->s.p), static_cast<u32*>(b->d.p),
                         static_cast<u32*>(b->w.p), b->m, b->n};
        if (b->reqs.size() == 1 && b->reqs[0]->in_device) {
          // one device-resident request: solved in place, nothing was copied
          const Request& r = *b->reqs[0];
          e = kmst_dev_edges{r.src, r.dst, r.w, b->m, b->n};
        }
        kmst_dev_mst o{static_cast<u32*>(b->out_s.p), static_cast<u32*>(b->out_d.p),
                       static_cast<u32*>(b->out_w.p), std::max<u64>(b->n, 1)};
        const kmst_status s2 =
            b->reqs.size() > 1
                ? solve_device_partitioned(ctx.get(), &e, &c, &o, &b->mst_n, &b->mst_w,
                                           static_cast<const u32*>(b->v_off.p),
                                           static_cast<u32>(b->reqs.size()))
                : kmst_solve_device(ctx.get(), &e, &c, &o, &b->mst_n, &b->mst_w);
        if (b->reqs.size() > 1) {
          KMST_HIP(hipStreamSynchronize(ctx->stream));
          if (*b->h_flag)
            throw kmst::Error(KMST_EINVAL, "a request has a vertex id at or past its n");
        }
        if (s2 != KMST_OK) throw kmst::Error(s2, kmst_last_error());
      });
      if (st != KMST_OK) {
        b->status = st;
        b->error = kmst::g_last_error;
      }
      if (b->status != KMST_OK && b->status != KMST_EDEVICE && b->reqs.size() > 1) {
        solve_alone(*b);
        b->solved_individually = true;
      }
    } else if (b->status == KMST_OK) {
      b->mst_n = 0;
      b->mst_w = 0;
    }
    const u64 t1 = now_ns();
    for (auto& r : b->reqs) r->res.solved_ns = t1;
    {
      std::lock_guard<std::mutex> lk(mu);
      done_sets.push_back(b);
      ++batches;
      batched_requests += b->reqs.size();
    }
    cv.notify_all();
  }
}

void kmst_server::deliver_loop() {
  (void)hipSetDevice(cfg.device);
  while (true) {
    BufferSet* b = nullptr;
    {
      std::unique_lock<std::mutex> lk(mu);
      cv.wait(lk, [&] { return stop || !done_sets.empty(); });
      if (done_sets.empty()) return;
      b = done_sets.front();
      done_sets.pop_front();
    }
    const u32 count = static_cast<u32>(b->reqs.size());
    if (b->status == KMST_OK && !b->solved_individually) {
      // split on the device: request ranges, local ids, weight sums; then
      // every request's slice goes straight into its output arrays
      const kmst_status st = guard([&] {
        const u64 k = b->mst_n;
        u64* dbounds = static_cast<u64*>(b->bounds.p);
        u64* dtotals = dbounds + (count + 1);
        batch_bounds_kernel<<<(count + BLOCK) / BLOCK, BLOCK, 0, down_stream>>>(
            static_cast<const u32*>(b->out_s.p), k, static_cast<const u32*>(b->v_off.p), count,
            dbounds);
        KMST_HIP(hipGetLastError());
        // output addresses for the requests whose arrays the kernel can write
        if (b->h_table_cap < 4 * u64(count)) {
          if (b->h_table) (void)hipHostFree(b->h_table);
          b->h_table = nullptr;
          const u64 cap = 4 * u64(count) + 4096;
          KMST_HIP(hipHostMalloc(reinterpret_cast<void**>(&b->h_table), cap * 8,
                                 hipHostMallocDefault));
          b->h_table_cap = cap;
        }
        for (u32 i = 0; i < count; ++i) {
          const Request& r = *b->reqs[i];
          u64* t = b->h_table + 4 * u64(i);
          t[0] = r.out_direct ? reinterpret_cast<u64>(r.out_src) : 0;
          t[1] = r.out_direct ? reinterpret_cast<u64>(r.out_dst) : 0;
          t[2] = r.out_direct ? reinterpret_cast<u64>(r.out_w) : 0;
          t[3] = r.out_cap;
        }
        u64* dtable = static_cast<u64*>(b->table.get(4 * u64(count) * 8));
        KMST_HIP(hipMemcpyAsync(dtable, b->h_table, 4 * u64(count) * 8, hipMemcpyHostToDevice,
                                down_stream));
        // several blocks per request, about 2048 in all
        fill_u64(down_stream, dtotals, count, 0);
        const u32 split = static_cast<u32>(std::min<u64>(64, std::max<u64>(1, 2048 / count)));
        batch_unoffset_kernel<<<dim3(count, split), BLOCK, 0, down_stream>>>(
            static_cast<u32*>(b->out_s.p), static_cast<u32*>(b->out_d.p),
            static_cast<const u32*>(b->out_w.p), dbounds, static_cast<const u32*>(b->v_off.p),
            dtable, dtotals);
        KMST_HIP(hipGetLastError());
        KMST_HIP(hipMemcpyAsync(b->h_bounds, dbounds, 2 * (count + 1) * 8, hipMemcpyDeviceToHost,
                                down_stream));
        KMST_HIP(hipStreamSynchronize(down_stream));
        const u64* hb = b->h_bounds;
        const u64* ht = b->h_bounds + (count + 1);
        for (u32 i = 0; i < count; ++i) {
          Request& r = *b->reqs[i];
          const u64 cnt = hb[i + 1] - hb[i];
          if (cnt > r.out_cap) {
            r.res.status = KMST_EINVAL;
            continue;
          }
          r.res.status = KMST_OK;
          r.res.mst_edges = cnt;
          r.res.total_weight = ht[i];
          if (cnt == 0 || r.out_direct) continue;  // direct: the kernel wrote it
          KMST_HIP(hipMemcpyAsync(r.out_src, static_cast<u32*>(b->out_s.p) + hb[i], cnt * 4,
                                  hipMemcpyDefault, down_stream));
          KMST_HIP(hipMemcpyAsync(r.out_dst, static_cast<u32*>(b->out_d.p) + hb[i], cnt * 4,
                                  hipMemcpyDefault, down_stream));
          KMST_HIP(hipMemcpyAsync(r.out_w, static_cast<u32*>(b->out_w.p) + hb[i], cnt * 4,
                                  hipMemcpyDefault, down_stream));
        }
        KMST_HIP(hipStreamSynchronize(down_stream));
      });
      if (st != KMST_OK) {
        for (auto& r : b->reqs) r->res.status = st;
      }
    } else if (!b->solved_individually) {
      for (auto& r : b->reqs) r->res.status = b->status;
    }
    for (auto& r : b->reqs) r->res.batch_requests = count;
    const u64 t_done = now_ns();
    {
      std::lock_guard<std::mutex> lk(mu);
      for (auto& r : b->reqs) {
        r->res.done_ns = t_done;
        r->res.done = 1;
      }
      b->reqs.clear();
      free_sets.push_back(b);
    }
    cv.notify_all();
  }
}

extern "C" {

void kmst_server_config_default(kmst_server_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->device = 0;
  c->algorithm = KMST_ALGO_FILTER_BORUVKA;
  c->filter_threshold = 4;
  c->max_batch_records = u64{1} << 28;
  c->max_batch_requests = 4096;
  c->hbm_budget_bytes = 0;
  c->queue_capacity = 1u << 20;
  c->buffer_sets = 3;
  kmst_tuning_default(&c->tuning);
}

kmst_status kmst_server_create(const kmst_server_config* cfg, kmst_server** out) {
  return guard([&] {
    if (!cfg || !out) throw kmst::Error(KMST_EINVAL, "null argument");
    if (cfg->max_batch_requests == 0 || cfg->max_batch_records == 0)
      throw kmst::Error(KMST_EINVAL, "empty batch limits");
    std::unique_ptr<kmst_server> s(new kmst_server);
    s->cfg = *cfg;
    if (s->cfg.buffer_sets < 2) s->cfg.buffer_sets = 2;
    KMST_HIP(hipSetDevice(cfg->device));
    if (s->cfg.hbm_budget_bytes == 0) {
      size_t free_b = 0, total_b = 0;
      KMST_HIP(hipMemGetInfo(&free_b, &total_b));
      s->cfg.hbm_budget_bytes = free_b / 2;
    }
    // admission bound: a batch must fit the budget together with the other
    // buffer sets' uploads
    const u64 max_rec = s->cfg.hbm_budget_bytes / BYTES_PER_RECORD;
    s->cfg.max_batch_records = std::min<u64>(s->cfg.max_batch_records, max_rec);
    s->ctx.reset(new kmst_context(cfg->device));
    if (kmst_context_set_tuning(s->ctx.get(), &cfg->tuning) != KMST_OK)
      throw kmst::Error(KMST_EINVAL, std::string("server tuning: ") + kmst_last_error());
    // reserve the solver workspaces and upload buffers for full batches up
    // front (admission by footprint: a batch never grows them)
    const u64 n_res = std::max<u64>(s->cfg.max_batch_records / 4, 1024);
    if (kmst_context_reserve(s->ctx.get(), s->cfg.max_batch_records, n_res) != KMST_OK)
      throw kmst::Error(KMST_ENOMEM, std::string("reserving batch workspaces: ") + kmst_last_error());
    for (int q = 0; q < kmst_server::UPLOAD_STREAMS; ++q) {
      KMST_HIP(hipStreamCreateWithFlags(&s->up_streams[q], hipStreamNonBlocking));
      KMST_HIP(hipEventCreateWithFlags(&s->up_events[q], hipEventDisableTiming));
    }
    s->copy_stream = s->up_streams[0];
    KMST_HIP(hipStreamCreateWithFlags(&s->down_stream, hipStreamNonBlocking));
    for (u32 i = 0; i < s->cfg.buffer_sets; ++i) {
      s->sets.emplace_back(new BufferSet);
      BufferSet& bs = *s->sets.back();
      for (DevBuf* d : {&bs.s, &bs.d, &bs.w}) d->get(s->cfg.max_batch_records * 4);
      for (DevBuf* d : {&bs.out_s, &bs.out_d, &bs.out_w}) d->get(n_res * 4);
      KMST_HIP(hipEventCreateWithFlags(&s->sets.back()->uploaded, hipEventDisableTiming));
      s->free_sets.push_back(s->sets.back().get());
    }
    kmst_server* raw = s.get();
    raw->admit_th = std::thread([raw] { raw->admit_loop(); });
    raw->compute_th = std::thread([raw] { raw->compute_loop(); });
    raw->deliver_th = std::thread([raw] { raw->deliver_loop(); });
    *out = s.release();
  });
}

kmst_status kmst_server_submit(kmst_server* s, const uint32_t* src, const uint32_t* dst,
                               const uint32_t* weight, uint64_t m_directed, uint64_t n,
                               uint32_t* out_src, uint32_t* out_dst, uint32_t* out_weight,
                               uint64_t out_capacity, uint64_t* ticket) {
  return guard([&] {
    if (!s || !ticket) throw kmst::Error(KMST_EINVAL, "null argument");
    if (m_directed && (!src || !dst || !weight)) throw kmst::Error(KMST_EINVAL, "null input");
    if (n >= (u64{1} << 31)) throw kmst::Error(KMST_EINVAL, "n >= 2^31");
    if (m_directed > s->cfg.max_batch_records || footprint(m_directed, n) > s->cfg.hbm_budget_bytes)
      throw kmst::Error(KMST_ENOMEM, "request exceeds the HBM admission budget");
    auto r = std::make_shared<Request>();
    r->src = src;
    r->dst = dst;
    r->w = weight;
    r->m = m_directed;
    r->n = n;
    r->out_src = out_src;
    r->out_dst = out_dst;
    r->out_w = out_weight;
    r->out_cap = out_capacity;
    r->in_device = device_resident(src, s->cfg.device) && device_resident(dst, s->cfg.device) &&
                   device_resident(weight, s->cfg.device);
    r->out_direct = device_writable(out_src, s->cfg.device) &&
                    device_writable(out_dst, s->cfg.device) &&
                    device_writable(out_weight, s->cfg.device);
    r->res.submit_ns = now_ns();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      if (s->queue.size() >= s->cfg.queue_capacity) throw kmst::Error(KMST_EBUSY, "queue full");
      r->ticket = s->next_ticket++;
      r->res.ticket = r->ticket;
      s->queue.push_back(r);
      s->pending[r->ticket] = r;
    }
    s->cv.notify_all();
    *ticket = r->ticket;
  });
}

kmst_status kmst_server_poll(kmst_server* s, uint64_t ticket, kmst_request_result* res) {
  return guard([&] {
    if (!s || !res) throw kmst::Error(KMST_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->pending.find(ticket);
    if (it == s->pending.end()) throw kmst::Error(KMST_EINVAL, "unknown ticket");
    *res = it->second->res;
    if (res->done) s->pending.erase(it);
  });
}

kmst_status kmst_server_wait(kmst_server* s, uint64_t ticket, kmst_request_result* res) {
  return guard([&] {
    if (!s || !res) throw kmst::Error(KMST_EINVAL, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->pending.find(ticket);
    if (it == s->pending.end()) throw kmst::Error(KMST_EINVAL, "unknown ticket");
    std::shared_ptr<Request> r = it->second;
    s->cv.wait(lk, [&] { return r->res.done != 0; });
    *res = r->res;
    s->pending.erase(ticket);
  });
}

kmst_status kmst_server_counters(kmst_server* s, uint64_t* batches, uint64_t* requests) {
  return guard([&] {
    if (!s) throw kmst::Error(KMST_EINVAL, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (batches) *batches = s->batches;
    if (requests) *requests = s->batched_requests;
  });
}

void kmst_server_destroy(kmst_server* s) {
  if (!s) return;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    // drain: wait until every admitted or queued request has completed
    s->cv.wait(lk, [&] {
      return s->queue.empty() && s->ready_sets.empty() && s->done_sets.empty() &&
             s->free_sets.size() == s->sets.size();
    });
    s->stop = true;
  }
  s->cv.notify_all();
  s->admit_th.join();
  s->compute_th.join();
  s->deliver_th.join();
  (void)hipSetDevice(s->cfg.device);
  for (auto& b : s->sets) {
    for (DevBuf* d : {&b->s, &b->d, &b->w, &b->out_s, &b->out_d, &b->out_w, &b->r_off, &b->v_off,
                      &b->bounds, &b->table, &b->flag, &b->gather})
      d->release();
    if (b->h_bounds) (void)hipHostFree(b->h_bounds);
    if (b->h_table) (void)hipHostFree(b->h_table);
    if (b->h_flag) (void)hipHostFree(b->h_flag);
    if (b->uploaded) (void)hipEventDestroy(b->uploaded);
  }
  for (int q = 0; q < kmst_server::UPLOAD_STREAMS; ++q) {
    if (s->up_streams[q]) (void)hipStreamDestroy(s->up_streams[q]);
    if (s->up_events[q]) (void)hipEventDestroy(s->up_events[q]);
  }
  if (s->down_stream) (void)hipStreamDestroy(s->down_stream);
  delete s;
}

}  // extern "C"
