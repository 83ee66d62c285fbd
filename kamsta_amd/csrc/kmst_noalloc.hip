// kmst_noalloc.hip -- kmst_solve_device without device allocation.
//
// kmst_solve_device_noalloc solves like kmst_solve_device, but its workspaces
// (record buffers, vertex arrays, sort and look-back status words, counters)
// come from `scratch`, a file-scope device array of fixed size, instead of a
// context's allocations; it runs on the caller's stream and keeps its host
// counters in static host words, so a call makes no hipMalloc, hipFree,
// hipHostMalloc or stream creation.  For callers that cannot allocate
// (serving loops that forbid it, code that must not touch the allocator).
//
// Set-up, once per device at its first call, with that device current:
// hipGetSymbolAddress gives this device's copy of `scratch`; a context whose
// buffers carve from it is reserved for the largest accepted solve (the same
// reservation kmst_context_reserve makes), after which the carving is frozen.
// Host code uses only that address, never the array's name.  Every call on a
// device shares that device's array, which keeps the last call's contents;
// the solver clears what must start at zero (counters every solve; status
// words once, when they are first carved, then by epoch).  Calls that
// overlapped on one GPU would race on that shared array: calls are blocking,
// and calls for one device are serialised by a lock, so within a process
// they never overlap.
#include <mutex>
#include <new>

#include "kmst_solver.hpp"

namespace kmst {
namespace {

// largest accepted solve and the scratch it needs: 40 B per record (two
// record buffers of 16 B and the 8-byte positional keys) and 64 B per vertex,
// plus sort, scan and look-back state (about 73 MiB at these limits)
constexpr u64 NOALLOC_MAX_RECORDS = u64{1} << 20;
constexpr u64 NOALLOC_MAX_VERTICES = u64{1} << 19;
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t{80} << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

}  // namespace
}  // namespace kmst

static __device__ __align__(256) unsigned char scratch[kmst::NOALLOC_SCRATCH_BYTES];

namespace kmst {
namespace {

struct Slot {
  std::mutex mu;
  bool ready = false;
  Arena arena;
  kmst_context* ctx = nullptr;
  alignas(kmst_context) unsigned char ctx_storage[sizeof(kmst_context)];
  u32 host_words[HOST_COUNTER_BYTES / 4];
};
Slot g_slots[NOALLOC_MAX_DEVICES];

// The device's context over its copy of `scratch` (caller holds s.mu).
kmst_context* set_up(Slot& s, int device, hipStream_t st) {
  if (s.ready) return s.ctx;
  KMST_HIP(hipSetDevice(device));
  void* ptr = nullptr;
  KMST_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  s.arena = Arena{static_cast<char*>(ptr), NOALLOC_SCRATCH_BYTES, 0, false};
  // storage of static duration, never destroyed (like the array itself)
  s.ctx = new (s.ctx_storage) kmst_context(device, s.host_words, &s.arena);
  s.ctx->stream = st;
  const kmst_status r =
      kmst_context_reserve(s.ctx, NOALLOC_MAX_RECORDS, NOALLOC_MAX_VERTICES);
  if (r != KMST_OK) throw Error(r, "no-allocation set-up: " + g_last_error);
  s.arena.frozen = true;
  KMST_HIP(hipStreamSynchronize(st));
  s.ready = true;
  return s.ctx;
}

}  // namespace
}  // namespace kmst

extern "C" kmst_status kmst_solve_device_noalloc(const kmst_dev_edges* in, const kmst_config* cfg,
                                                 kmst_dev_mst* out, uint64_t* out_n,
                                                 uint64_t* total_weight, void* stream) {
  using namespace kmst;
  kmst_context* ctx = nullptr;
  Slot* slot = nullptr;
  const kmst_status pre = guarded_call([&] {
    if (!in || !cfg || !out || !out_n || !total_weight) throw Error(KMST_EINVAL, "null argument");
    if (in->m_directed > NOALLOC_MAX_RECORDS || in->n > NOALLOC_MAX_VERTICES)
      throw Error(KMST_EINVAL,
                  "solve beyond the no-allocation limits (2^20 records, 2^19 vertices): " +
                      std::to_string(in->m_directed) + " records, n = " + std::to_string(in->n));
    if (cfg->device < 0 || cfg->device >= NOALLOC_MAX_DEVICES)
      throw Error(KMST_EINVAL, "device ordinal out of range");
    slot = &g_slots[cfg->device];
  });
  if (pre != KMST_OK) return pre;
  std::lock_guard<std::mutex> lk(slot->mu);
  const hipStream_t st = static_cast<hipStream_t>(stream);
  const kmst_status su = guarded_call([&] { ctx = set_up(*slot, cfg->device, st); });
  if (su != KMST_OK) return su;
  ctx->stream = st;
  return kmst_solve_device(ctx, in, cfg, out, out_n, total_weight);
}
