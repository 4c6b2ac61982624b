// k_xfer.hip -- import / export / set kernels.
//
// The reference repacks words with xmpTransform (src/xmp.cu:461-489) into a
// compact buffer and transposes lazily to its strided layout with
// xmpC2S/xmpS2C (src/xmp.cu:817-843).  Here the repack and the transpose are
// one kernel: compact host-format words go straight to the interleaved
// device layout and back.  The common case (32-bit words, least significant
// first, native endian, no nails, no indices) uses an LDS-tiled transpose so
// that both the compact reads and the interleaved writes are coalesced.
#include "kernels.hpp"

namespace xmp {

namespace {

constexpr int kTileI = 64;  // instances per tile
constexpr int kTileJ = 32;  // limbs per tile

// byte k (little-endian significance) of a compact integer row held as
// `words` words of `size` bytes with the given order/endian/nails.
__device__ __forceinline__ uint32_t import_byte(const uint8_t *row, uint32_t k, uint32_t words, int32_t order,
                                                int32_t endian, uint32_t nails, uint32_t size) {
  uint32_t w = k / size;
  uint32_t kb = k - w * size;
  if (w >= words) return 0;
  uint32_t pos = (order == -1) ? w : words - 1 - w;
  uint32_t mem = (endian == -1) ? kb : size - 1 - kb;
  uint32_t v = row[(size_t)pos * size + mem];
  // mask word(-1) >> nails applied to the word value (reference xmp.cu:481-483)
  uint64_t wmask = (size == 8 ? ~0ull : ((1ull << (8 * size)) - 1)) >> nails;
  return v & (uint32_t)((wmask >> (8 * kb)) & 0xFF);
}

__global__ void import_fast_kernel(uint32_t *__restrict__ dst, uint32_t stride, uint32_t nlimbs,
                                   const uint32_t *__restrict__ src, uint32_t count, uint32_t words) {
  __shared__ uint32_t tile[kTileI][kTileJ + 1];
  uint32_t i0 = blockIdx.x * kTileI, j0 = blockIdx.y * kTileJ;
  for (uint32_t r = threadIdx.x; r < kTileI * kTileJ; r += blockDim.x) {
    uint32_t ii = r / kTileJ, jj = r % kTileJ;
    uint32_t i = i0 + ii, j = j0 + jj;
    tile[ii][jj] = (i < count && j < words) ? src[(size_t)i * words + j] : 0u;
  }
  __syncthreads();
  for (uint32_t r = threadIdx.x; r < kTileI * kTileJ; r += blockDim.x) {
    uint32_t jj = r / kTileI, ii = r % kTileI;
    uint32_t i = i0 + ii, j = j0 + jj;
    if (i < count && j < nlimbs) dst[(size_t)j * stride + i] = tile[ii][jj];
  }
}

__global__ void import_general_kernel(uint32_t *__restrict__ dst, uint32_t stride, uint32_t nlimbs,
                                      const uint8_t *__restrict__ src, uint32_t count, uint32_t words,
                                      int32_t order, int32_t endian, uint32_t nails, uint32_t size,
                                      const uint32_t *__restrict__ in_idx, uint32_t in_idx_count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t j = blockIdx.y;
  if (i >= count || j >= nlimbs) return;
  uint32_t row = in_idx ? in_idx[i % in_idx_count] : i;
  const uint8_t *r = src + (size_t)row * words * size;
  uint32_t v = 0;
#pragma unroll
  for (uint32_t b = 0; b < 4; b++) v |= import_byte(r, 4 * j + b, words, order, endian, nails, size) << (8 * b);
  dst[(size_t)j * stride + i] = v;
}

__global__ void export_fast_kernel(uint32_t *__restrict__ dst, const uint32_t *__restrict__ src, uint32_t stride,
                                   uint32_t nlimbs, uint32_t count) {
  __shared__ uint32_t tile[kTileI][kTileJ + 1];
  uint32_t i0 = blockIdx.x * kTileI, j0 = blockIdx.y * kTileJ;
  for (uint32_t r = threadIdx.x; r < kTileI * kTileJ; r += blockDim.x) {
    uint32_t jj = r / kTileI, ii = r % kTileI;
    uint32_t i = i0 + ii, j = j0 + jj;
    tile[ii][jj] = (i < count && j < nlimbs) ? src[(size_t)j * stride + i] : 0u;
  }
  __syncthreads();
  for (uint32_t r = threadIdx.x; r < kTileI * kTileJ; r += blockDim.x) {
    uint32_t ii = r / kTileJ, jj = r % kTileJ;
    uint32_t i = i0 + ii, j = j0 + jj;
    if (i < count && j < nlimbs) dst[(size_t)i * nlimbs + j] = tile[ii][jj];
  }
}

// one thread per output word (size bytes) of the compact output
__global__ void export_general_kernel(uint8_t *__restrict__ dst, const uint32_t *__restrict__ src, uint32_t stride,
                                      uint32_t nlimbs, uint32_t count, uint32_t words, int32_t order,
                                      int32_t endian, uint32_t nails, uint32_t size,
                                      const uint32_t *__restrict__ out_idx, uint32_t out_idx_count) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t j = blockIdx.y;
  if (i >= count || j >= words) return;
  uint32_t orow = out_idx ? out_idx[i % out_idx_count] : i;
  uint32_t pos = (order == -1) ? j : words - 1 - j;
  uint64_t wmask = (size == 8 ? ~0ull : ((1ull << (8 * size)) - 1)) >> nails;
  uint8_t *o = dst + ((size_t)orow * words + j) * size;
  for (uint32_t p = 0; p < size; p++) {
    uint32_t sig = (endian == -1) ? p : size - 1 - p;   // significance of memory byte p after swap
    uint32_t k = pos * size + sig;                       // byte of the integer
    uint32_t limb = k >> 2;
    uint32_t v = limb < nlimbs ? src[(size_t)limb * stride + i] : 0u;
    uint32_t byte = (v >> (8 * (k & 3))) & 0xFF;
    o[p] = (uint8_t)(byte & (uint32_t)((wmask >> (8 * p)) & 0xFF));
  }
}

__global__ void set_kernel(uint32_t *__restrict__ out, uint32_t out_stride, uint32_t out_limbs,
                           const uint32_t *__restrict__ in, uint32_t in_stride, uint32_t in_limbs, uint32_t in_count,
                           uint32_t count, const uint32_t *__restrict__ out_idx, uint32_t out_idx_count,
                           const uint32_t *__restrict__ in_idx, uint32_t in_idx_count) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t j = blockIdx.y;
  if (k >= count || j >= out_limbs) return;
  uint32_t i = in_idx ? in_idx[k % in_idx_count] % in_count : k;
  uint32_t o = out_idx ? out_idx[k % out_idx_count] : k;
  out[(size_t)j * out_stride + o] = j < in_limbs ? in[(size_t)j * in_stride + i] : 0u;
}

}  // namespace

hipError_t launch_import(uint32_t *dst, uint32_t dst_stride, uint32_t nlimbs, const uint8_t *src, uint32_t count,
                         uint32_t words, int32_t order, int32_t endian, uint32_t nails, uint32_t size,
                         const uint32_t *in_idx, uint32_t in_idx_count, hipStream_t stream) {
  if (size == 4 && order == -1 && endian == -1 && nails == 0 && in_idx == nullptr) {
    dim3 grid(div_round_up(count, kTileI), div_round_up(nlimbs, kTileJ));
    hipLaunchKernelGGL(import_fast_kernel, grid, dim3(256), 0, stream, dst, dst_stride, nlimbs,
                       reinterpret_cast<const uint32_t *>(src), count, words);
  } else {
    dim3 grid(div_round_up(count, 256), nlimbs);
    hipLaunchKernelGGL(import_general_kernel, grid, dim3(256), 0, stream, dst, dst_stride, nlimbs, src, count,
                       words, order, endian, nails, size, in_idx, in_idx_count);
  }
  return hipGetLastError();
}

hipError_t launch_export(uint8_t *dst, const uint32_t *src, uint32_t src_stride, uint32_t nlimbs, uint32_t count,
                         uint32_t words, int32_t order, int32_t endian, uint32_t nails, uint32_t size,
                         const uint32_t *out_idx, uint32_t out_idx_count, hipStream_t stream) {
  if (size == 4 && order == -1 && endian == -1 && nails == 0 && out_idx == nullptr) {
    dim3 grid(div_round_up(count, kTileI), div_round_up(nlimbs, kTileJ));
    hipLaunchKernelGGL(export_fast_kernel, grid, dim3(256), 0, stream, reinterpret_cast<uint32_t *>(dst), src,
                       src_stride, nlimbs, count);
  } else {
    dim3 grid(div_round_up(count, 256), words);
    hipLaunchKernelGGL(export_general_kernel, grid, dim3(256), 0, stream, dst, src, src_stride, nlimbs, count,
                       words, order, endian, nails, size, out_idx, out_idx_count);
  }
  return hipGetLastError();
}

hipError_t launch_set(uint32_t *out, uint32_t out_stride, uint32_t out_limbs, const uint32_t *in, uint32_t in_stride,
                      uint32_t in_limbs, uint32_t in_count, uint32_t count, const uint32_t *out_idx,
                      uint32_t out_idx_count, const uint32_t *in_idx, uint32_t in_idx_count, hipStream_t stream) {
  dim3 grid(div_round_up(count, 256), out_limbs);
  hipLaunchKernelGGL(set_kernel, grid, dim3(256), 0, stream, out, out_stride, out_limbs, in, in_stride, in_limbs,
                     in_count, count, out_idx, out_idx_count, in_idx, in_idx_count);
  return hipGetLastError();
}

hipError_t launch_copy_words(uint32_t *dst, const uint32_t *src, size_t words, hipStream_t stream) {
  return hipMemcpyAsync(dst, src, words * sizeof(uint32_t), hipMemcpyDeviceToDevice, stream);
}

}  // namespace xmp
