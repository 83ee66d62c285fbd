// kmst_codec.hip -- the difference codec of compression level 1
// (DifferenceEncodedGraph::encode_edges / get_WEdges / decode_edge,
// include/datastructures/compression/difference_encoded_graph.hpp:92-212, with the
// 7-bit varints of encode_value / decode_value,
// include/datastructures/variable_length_encoding.hpp:6-160).
//
// Format: records are cut into chunks of chunk_records.  Chunk k keeps its
// first record raw in first[3k .. 3k + 2] = (src, dst, weight); every later
// record of the chunk is three varints (src - previous src mod 2^32, dst,
// weight), little-endian groups of 7 bits, the high bit set on every byte
// but the last.  offset[k] is the byte where chunk k's varints start and
// offset[chunks] the stream's length.  The host encodes in parallel over
// chunks; the device decodes one chunk per thread, so chunks are short (32
// records) to give every wave work.
#include <algorithm>
#include <thread>
#include <vector>

#include "kmst_kernels.hpp"

namespace kmst {

namespace {

inline u32 varint_bytes(u32 v) { return v < (1u << 7) ? 1 : v < (1u << 14) ? 2 : v < (1u << 21) ? 3 : v < (1u << 28) ? 4 : 5; }

inline u8* put_varint(u8* p, u32 v) {
  while (v >= 0x80) {
    *p++ = static_cast<u8>((v & 0x7F) | 0x80);
    v >>= 7;
  }
  *p++ = static_cast<u8>(v);
  return p;
}

template <typename F>
void parallel_chunks(u64 chunks, F&& f) {
  const u64 threads = std::max<u64>(1, std::min<u64>(std::thread::hardware_concurrency(), 64));
  const u64 per = (chunks + threads - 1) / threads;
  std::vector<std::thread> pool;
  for (u64 t = 0; t < threads && t * per < chunks; ++t)
    pool.emplace_back([&, t] {
      for (u64 k = t * per; k < std::min(chunks, (t + 1) * per); ++k) f(k);
    });
  for (auto& th : pool) th.join();
}

}  // namespace

u64 encoded_chunks(u64 m, u32 chunk_records) { return (m + chunk_records - 1) / chunk_records; }

void encode_host(const u32* src, const u32* dst, const u32* w, u64 m, u32 chunk_records,
                 std::vector<u8>& bytes, std::vector<u64>& offset, std::vector<u32>& first) {
  const u64 chunks = encoded_chunks(m, chunk_records);
  offset.assign(chunks + 1, 0);
  first.assign(3 * chunks, 0);
  parallel_chunks(chunks, [&](u64 k) {
    const u64 lo = k * chunk_records, hi = std::min<u64>(m, lo + chunk_records);
    first[3 * k] = src[lo];
    first[3 * k + 1] = dst[lo];
    first[3 * k + 2] = w[lo];
    u64 b = 0;
    for (u64 i = lo + 1; i < hi; ++i)
      b += varint_bytes(src[i] - src[i - 1]) + varint_bytes(dst[i]) + varint_bytes(w[i]);
    offset[k + 1] = b;
  });
  for (u64 k = 0; k < chunks; ++k) offset[k + 1] += offset[k];
  bytes.resize(offset[chunks]);
  parallel_chunks(chunks, [&](u64 k) {
    const u64 lo = k * chunk_records, hi = std::min<u64>(m, lo + chunk_records);
    u8* p = bytes.data() + offset[k];
    for (u64 i = lo + 1; i < hi; ++i) {
      p = put_varint(p, src[i] - src[i - 1]);
      p = put_varint(p, dst[i]);
      p = put_varint(p, w[i]);
    }
  });
}

// One value; a stream that ends early or a value wider than 32 bits raises
// ERR_INVALID and yields 0.
__device__ __forceinline__ u32 get_varint(const u8* __restrict__ bytes, u64& p, u64 end, u32& bad) {
  u32 v = 0;
  for (u32 shift = 0; shift < 35; shift += 7) {
    if (p >= end) {
      bad = 1;
      return 0;
    }
    const u32 b = bytes[p++];
    v |= (b & 0x7Fu) << shift;
    if (!(b & 0x80u)) return v;
  }
  bad = 1;
  return 0;
}

__global__ __launch_bounds__(BLOCK) void decode_kernel(const u8* __restrict__ bytes,
                                                       const u64* __restrict__ offset,
                                                       const u32* __restrict__ first, u64 m,
                                                       u32 chunk_records, u64 chunks,
                                                       u32* __restrict__ src,
                                                       u32* __restrict__ dst, u32* __restrict__ w,
                                                       u32* err) {
  const u64 k = u64(blockIdx.x) * BLOCK + threadIdx.x;
  if (k >= chunks) return;
  const u64 lo = k * chunk_records, hi = min(m, lo + chunk_records);
  u32 s = first[3 * k];
  src[lo] = s;
  dst[lo] = first[3 * k + 1];
  w[lo] = first[3 * k + 2];
  u64 p = offset[k];
  const u64 end = offset[k + 1];
  u32 bad = 0;
  for (u64 i = lo + 1; i < hi; ++i) {
    s += get_varint(bytes, p, end, bad);
    src[i] = s;
    dst[i] = get_varint(bytes, p, end, bad);
    w[i] = get_varint(bytes, p, end, bad);
  }
  if (bad || p != end) atomicOr(err, ERR_INVALID);
}

void decode_device(hipStream_t st, const u8* bytes, const u64* offset, const u32* first, u64 m,
                   u32 chunk_records, u32* src, u32* dst, u32* w, u32* err) {
  if (m == 0) return;
  const u64 chunks = encoded_chunks(m, chunk_records);
  KTScope t(st, K_OTHER, 12.0 * m);
  decode_kernel<<<ceil_div_u32(chunks, BLOCK), BLOCK, 0, st>>>(bytes, offset, first, m,
                                                              chunk_records, chunks, src, dst, w,
                                                              err);
  KMST_HIP(hipGetLastError());
}

}  // namespace kmst
