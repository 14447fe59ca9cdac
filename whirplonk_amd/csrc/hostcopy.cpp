// hostcopy.cpp — copies between pageable host memory and the device for the
// host-pointer entry points (whir_gpu_*, whir_memcpy_*).
//
// A copy to or from pageable memory runs at 10-30 GB/s through the runtime's
// own staging on this machine, against 57 GB/s for DMA to and from pinned
// memory.  Large copies therefore go through the context's two pinned
// staging chunks: the DMA of one chunk runs while host threads move the other
// chunk between pinned and pageable memory, so the copy runs at about the
// slower of the DMA and the host copy rate (several threads: the destination
// pages are first touched in parallel).  Pinned host memory and small copies
// take one plain DMA.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <mutex>

#include "ctx.h"

namespace whir_gpu {
namespace {

constexpr size_t STAGE_CHUNK = size_t{16} << 20;  // bytes per staging chunk
constexpr size_t STAGED_MIN = size_t{4} << 20;    // smaller copies take one plain DMA
constexpr size_t PART = size_t{1} << 20;          // bytes per host-thread part

// dst[0..bytes) = src[0..bytes) split over the copy threads
void parallel_memcpy(void* dst, const void* src, size_t bytes) {
    const size_t parts = (bytes + PART - 1) / PART;
    if (parts <= 1) {
        std::memcpy(dst, src, bytes);
        return;
    }
    CopyPool::get().run(parts, [&](size_t i) {
        const size_t off = i * PART;
        std::memcpy((uint8_t*)dst + off, (const uint8_t*)src + off, std::min(PART, bytes - off));
    });
}

bool is_pinned(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();  // unregistered pageable memory may report an error: clear it
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

}  // namespace

int ensure_stage(whir_ctx* ctx) {
    if (ctx->stage[0]) return 0;
    std::lock_guard<std::mutex> lk(alloc_mutex());
    for (int k = 0; k < 2; ++k) {
        HIPCHK(hipHostMalloc(&ctx->stage[k], STAGE_CHUNK, hipHostMallocDefault));
        HIPCHK(hipEventCreateWithFlags(&ctx->stage_ev[k], hipEventDisableTiming));
        HIPCHK(hipEventRecord(ctx->stage_ev[k], ctx->stream));
    }
    return 0;
}

int copy_h2d(whir_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return 0;
    if (bytes < STAGED_MIN || is_pinned(src)) {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
        return 0;
    }
    if (int rc = ensure_stage(ctx)) return rc;
    for (size_t off = 0, k = 0; off < bytes; off += STAGE_CHUNK, k ^= 1) {
        const size_t len = std::min(STAGE_CHUNK, bytes - off);
        // the DMA out of this chunk two steps ago has finished
        HIPCHK(hipEventSynchronize(ctx->stage_ev[k]));
        parallel_memcpy(ctx->stage[k], (const uint8_t*)src + off, len);
        HIPCHK(hipMemcpyAsync((uint8_t*)dst + off, ctx->stage[k], len, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipEventRecord(ctx->stage_ev[k], ctx->stream));
    }
    return 0;
}

int copy_d2h_each(whir_ctx* ctx, const void* src, size_t bytes, size_t unit, const D2hConsume& consume) {
    if (!bytes) return 0;
    if (!unit || bytes % unit) return bad_arg("copy size is not a multiple of its unit");
    if (int rc = ensure_stage(ctx)) return rc;
    const size_t chunk = STAGE_CHUNK / unit * unit, part = std::max<size_t>(1, PART / unit) * unit;
    const size_t nchunks = (bytes + chunk - 1) / chunk;
    auto enqueue = [&](size_t c) -> int {
        const size_t off = c * chunk, len = std::min(chunk, bytes - off);
        HIPCHK(hipMemcpyAsync(ctx->stage[c & 1], (const uint8_t*)src + off, len, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipEventRecord(ctx->stage_ev[c & 1], ctx->stream));
        return 0;
    };
    // chunk c + 1 is in flight while the host moves chunk c; chunk c + 2 is
    // enqueued only after chunk c has left its buffer
    if (int rc = enqueue(0)) return rc;
    for (size_t c = 0; c < nchunks; ++c) {
        if (c + 1 < nchunks)
            if (int rc = enqueue(c + 1)) return rc;
        HIPCHK(hipEventSynchronize(ctx->stage_ev[c & 1]));
        const size_t off = c * chunk, len = std::min(chunk, bytes - off);
        const uint8_t* p = (const uint8_t*)ctx->stage[c & 1];
        const size_t parts = (len + part - 1) / part;
        if (parts <= 1) {
            consume(off, p, len);
        } else {
            CopyPool::get().run(parts, [&](size_t i) {
                const size_t o = i * part;
                consume(off + o, p + o, std::min(part, len - o));
            });
        }
    }
    return 0;
}

int copy_d2h(whir_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return 0;
    if (bytes < STAGED_MIN || is_pinned(dst)) {
        HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return 0;
    }
    return copy_d2h_each(ctx, src, bytes, 1,
                         [dst](size_t off, const uint8_t* p, size_t len) { std::memcpy((uint8_t*)dst + off, p, len); });
}

}  // namespace whir_gpu
