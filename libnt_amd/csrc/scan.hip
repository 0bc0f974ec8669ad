// scan.hip -- device-wide exclusive prefix sums (reduce-then-scan), used by
// the lattice product for product offsets, segment ids and output offsets.
// Chunk = 256 threads x 16 items; chunk sums are scanned recursively.
#include "lnt_internal.h"

namespace lnt {

namespace {

constexpr int kBlock = 256;
constexpr int kItems = 16;
constexpr uint64_t kChunk = (uint64_t)kBlock * kItems;

template <typename T>
__global__ __launch_bounds__(kBlock) void chunk_sum_kernel(const T *__restrict__ in, uint64_t n,
                                                           uint64_t *__restrict__ sums) {
  __shared__ uint64_t s[kBlock / kWave];
  const uint64_t base = (uint64_t)blockIdx.x * kChunk;
  uint64_t acc = 0;
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    uint64_t p = base + (uint64_t)i * kBlock + threadIdx.x;
    uint64_t x = (uint64_t)in[p < n ? p : n - 1];  // unconditional load
    acc += p < n ? x : 0;
  }
  uint64_t total;
  block_exclusive_sum64_lds(acc, s, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of the BT * kItems elements from `base`: wave w owns the
// kWave * kItems consecutive elements from base + w * kWave * kItems, as
// kItems rows of kWave (lane l of row i holds element i * kWave + l), so
// every load and store is one coalesced row; each row is scanned across the
// wave with shuffles, carrying the running sum from row to row, and the
// waves' totals are scanned in LDS.  Returns the block's total; out[p] =
// `offset` + the sum of the elements before p.
template <int BT, typename T>
__device__ inline uint64_t block_rows_scan(const T *__restrict__ in, uint64_t n, uint64_t base, uint64_t offset,
                                           uint64_t *__restrict__ out, uint64_t *__restrict__ s) {
  const uint32_t lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
  const uint64_t wbase = base + (uint64_t)wave * kWave * kItems;
  uint64_t v[kItems];
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t p = wbase + (uint64_t)i * kWave + lane;
    const uint64_t x = (uint64_t)in[p < n ? p : n - 1];  // unconditional load
    v[i] = p < n ? x : 0;
  }
  uint64_t run = 0;  // sum of the wave's earlier rows
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t incl = wave_inclusive_sum64(v[i]);
    const uint64_t row_total = __shfl(incl, kWave - 1, kWave);
    v[i] = run + incl - v[i];
    run += row_total;
  }
  uint64_t total;
  const uint64_t before = block_exclusive_sum_waves<BT / kWave>(lane == 0 ? run : 0, s, &total);
  const uint64_t wave_off = offset + __shfl(before, 0, kWave);
#pragma unroll
  for (int i = 0; i < kItems; ++i) {
    const uint64_t p = wbase + (uint64_t)i * kWave + lane;
    if (p < n) out[p] = wave_off + v[i];
  }
  return total;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void chunk_scan_kernel(const T *__restrict__ in, uint64_t n,
                                                            const uint64_t *__restrict__ offsets,
                                                            uint64_t *__restrict__ out,
                                                            uint64_t *__restrict__ total_out) {
  __shared__ uint64_t s[kBlock / kWave];
  const uint64_t offset = offsets ? offsets[blockIdx.x] : 0;
  const uint64_t total = block_rows_scan<kBlock, T>(in, n, (uint64_t)blockIdx.x * kChunk, offset, out, s);
  if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *total_out = offset + total;
}

// Scans of up to kOneBlockMax elements in one launch: one 1024-thread
// workgroup walks the array in rounds of 1024 x kItems elements, carrying the
// running total (three launches become one for the lattice product's
// per-column tables).
constexpr int kOneBlock = 1024;
constexpr uint64_t kOneBlockMax = (uint64_t)kOneBlock * kItems * 4;

template <typename T>
__global__ __launch_bounds__(kOneBlock) void one_block_scan_kernel(const T *__restrict__ in, uint64_t n,
                                                                   uint64_t *__restrict__ out,
                                                                   uint64_t *__restrict__ total_out) {
  // Each thread owns kItems consecutive elements of a round (one workgroup:
  // the shuffle-based row scan of chunk_scan_kernel measured slower here).
  __shared__ uint64_t s[kOneBlock / kWave];
  uint64_t carry = 0;
  for (uint64_t base0 = 0; base0 < n; base0 += (uint64_t)kOneBlock * kItems) {
    const uint64_t base = base0 + (uint64_t)threadIdx.x * kItems;
    uint64_t v[kItems];
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const uint64_t p = base + i;
      const uint64_t x = (uint64_t)in[p < n ? p : n - 1];  // unconditional load
      v[i] = p < n ? x : 0;
      acc += v[i];
    }
    uint64_t total;
    uint64_t run = block_exclusive_sum_waves<kOneBlock / kWave>(acc, s, &total) + carry;
#pragma unroll
    for (int i = 0; i < kItems; ++i) {
      const uint64_t p = base + i;
      if (p < n) out[p] = run;
      run += v[i];
    }
    carry += total;
  }
  if (total_out && threadIdx.x == 0) *total_out = carry;
}

}  // namespace

// out[i] = sum_{j<i} in[j] for i < n; *total_dev (optional, device) = sum of all.
// `in` and `out` may alias only when T is uint64_t.
template <typename T>
int exclusive_scan(lnt_context ctx, const T *in, uint64_t *out, uint64_t n, uint64_t *total_dev,
                   DeviceBuffer &tmp, size_t tmp_offset) {
  if (n == 0) {
    if (total_dev) LNT_TRY(launch_fill_u32(ctx, (uint32_t *)total_dev, 0, 2));
    return LNT_OK;
  }
  uint64_t chunks = (n + kChunk - 1) / kChunk;
  if (chunks == 1) {
    chunk_scan_kernel<T><<<1, kBlock, 0, ctx->stream>>>(in, n, nullptr, out, total_dev);
    LNT_HIP(hipGetLastError());
    return LNT_OK;
  }
  if (n <= kOneBlockMax) {
    one_block_scan_kernel<T><<<1, kOneBlock, 0, ctx->stream>>>(in, n, out, total_dev);
    LNT_HIP(hipGetLastError());
    return LNT_OK;
  }
  // chunk sums at tmp[tmp_offset ...], then their scan in place (recursively).
  size_t need = (tmp_offset + chunks + 1) * sizeof(uint64_t);
  if (tmp.bytes < need) return fail(LNT_ERR_DEVICE, "scan workspace too small");
  uint64_t *sums = (uint64_t *)tmp.ptr + tmp_offset;
  chunk_sum_kernel<T><<<(uint32_t)chunks, kBlock, 0, ctx->stream>>>(in, n, sums);
  LNT_HIP(hipGetLastError());
  LNT_TRY(exclusive_scan<uint64_t>(ctx, sums, sums, chunks, nullptr, tmp, tmp_offset + chunks + 1));
  chunk_scan_kernel<T><<<(uint32_t)chunks, kBlock, 0, ctx->stream>>>(in, n, sums, out, total_dev);
  LNT_HIP(hipGetLastError());
  return LNT_OK;
}

// Words of scan workspace needed for n elements.
size_t scan_workspace_words(uint64_t n) {
  size_t words = 0;
  while (n > kChunk) {
    n = (n + kChunk - 1) / kChunk;
    words += n + 1;
  }
  return words + 2;
}

template int exclusive_scan<uint32_t>(lnt_context, const uint32_t *, uint64_t *, uint64_t, uint64_t *,
                                      DeviceBuffer &, size_t);
template int exclusive_scan<uint64_t>(lnt_context, const uint64_t *, uint64_t *, uint64_t, uint64_t *,
                                      DeviceBuffer &, size_t);

}  // namespace lnt
