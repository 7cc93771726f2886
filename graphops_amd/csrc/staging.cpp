// staging.cpp -- large copies between pageable host memory and the device
// through a ring of pinned chunks: host threads fill (or drain) one chunk
// while the DMA engine moves another.  A copy straight from pageable memory
// goes through the runtime's own staging, one chunk at a time on one thread.
#include <omp.h>

#include <algorithm>
#include <cstring>

#include "internal.h"

namespace gops {

namespace {

constexpr size_t kChunk = size_t(32) << 20;
constexpr size_t kSmall = size_t(4) << 20;  // below this, a plain copy

// memcpy of len bytes on the OpenMP threads, 1 MiB pieces
void par_copy(void *dst, const void *src, size_t len) {
  constexpr size_t kPiece = size_t(1) << 20;
  const int64_t pieces = (int64_t)((len + kPiece - 1) / kPiece);
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < pieces; p++) {
    const size_t o = (size_t)p * kPiece;
    std::memcpy((char *)dst + o, (const char *)src + o, std::min(kPiece, len - o));
  }
}

hipError_t ensure_chunks(gops_context *ctx) {
  if (ctx->pin[0]) return hipSuccess;
  for (int i = 0; i < kPinChunks; i++) {
    hipError_t e = hipHostMalloc(&ctx->pin[i], kChunk, hipHostMallocDefault);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->pin_done[i], hipEventDisableTiming);
    if (e != hipSuccess) {
      free_pinned_chunks(ctx);
      return e;
    }
  }
  return hipSuccess;
}

}  // namespace

hipError_t staged_upload(gops_context *ctx, void *dst, const void *src, size_t bytes,
                         hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  if (bytes < kSmall) return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s);
  hipError_t e = ensure_chunks(ctx);
  if (e != hipSuccess) return e;
  int i = 0;
  for (size_t o = 0; o < bytes; o += kChunk, i = (i + 1) % kPinChunks) {
    const size_t len = std::min(kChunk, bytes - o);
    // the chunk's previous transfer has drained (an event never recorded
    // counts as complete)
    if ((e = hipEventSynchronize(ctx->pin_done[i])) != hipSuccess) return e;
    par_copy(ctx->pin[i], (const char *)src + o, len);
    if ((e = hipMemcpyAsync((char *)dst + o, ctx->pin[i], len, hipMemcpyHostToDevice, s)) !=
        hipSuccess)
      return e;
    if ((e = hipEventRecord(ctx->pin_done[i], s)) != hipSuccess) return e;
  }
  // the source has been read in full; transfers may still be in flight from
  // the pinned chunks, which the next staged copy waits for
  return hipSuccess;
}

hipError_t staged_download(gops_context *ctx, void *dst, const void *src, size_t bytes,
                           hipStream_t s) {
  if (bytes == 0) return hipSuccess;
  if (bytes < kSmall) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
  }
  hipError_t e = ensure_chunks(ctx);
  if (e != hipSuccess) return e;
  const size_t chunks = (bytes + kChunk - 1) / kChunk;
  auto issue = [&](size_t c) -> hipError_t {
    const int i = (int)(c % kPinChunks);
    const size_t o = c * kChunk, len = std::min(kChunk, bytes - o);
    hipError_t r = hipEventSynchronize(ctx->pin_done[i]);  // an earlier upload's use
    if (r == hipSuccess)
      r = hipMemcpyAsync(ctx->pin[i], (const char *)src + o, len, hipMemcpyDeviceToHost, s);
    if (r == hipSuccess) r = hipEventRecord(ctx->pin_done[i], s);
    return r;
  };
  // keep kPinChunks - 1 transfers ahead of the host copies
  size_t next = 0;
  for (; next < chunks && next < (size_t)kPinChunks - 1; next++)
    if ((e = issue(next)) != hipSuccess) return e;
  for (size_t c = 0; c < chunks; c++) {
    if (next < chunks) {
      if ((e = issue(next)) != hipSuccess) return e;
      next++;
    }
    const int i = (int)(c % kPinChunks);
    if ((e = hipEventSynchronize(ctx->pin_done[i])) != hipSuccess) return e;
    const size_t o = c * kChunk;
    par_copy((char *)dst + o, ctx->pin[i], std::min(kChunk, bytes - o));
  }
  return hipSuccess;
}

void free_pinned_chunks(gops_context *ctx) {
  for (int i = 0; i < kPinChunks; i++) {
    if (ctx->pin_done[i]) {
      hipEventSynchronize(ctx->pin_done[i]);
      hipEventDestroy(ctx->pin_done[i]);
      ctx->pin_done[i] = nullptr;
    }
    if (ctx->pin[i]) {
      hipHostFree(ctx->pin[i]);
      ctx->pin[i] = nullptr;
    }
  }
}

}  // namespace gops
