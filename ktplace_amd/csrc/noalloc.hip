// HPWL that allocates nothing (kt_hpwl_noalloc): for callers that cannot call
// hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
// forbid allocation).  It returns the same double as kt_hpwl, bit for bit:
// getHpwl's per-net bounding-box half-perimeters (simpl_placer.cc:256-275)
// summed in 4096-net chunks, each chunk sequentially, then the chunk sums in
// order (getDeterministicSum, simpl_parallel.h:35-51).
//
// Where kt_hpwl keeps per-net values and chunk sums in a context buffer that
// it grows on demand, this entry point keeps a chunk's per-net values in LDS
// and the chunk sums in `scratch`, a file-scope device array of fixed size:
// at most kNoAllocMaxChunks chunks, that is 67,104,768 nets.  Pin positions
// are read from the context's positions and fixed coordinates directly (no
// vertex-position buffer).  Each GPU holds its own copy of `scratch` until the
// process ends; its address on a device is looked up once, with that device
// current, when the first context on the device is created.  All calls on one
// GPU share that GPU's copy: calls that overlap on one GPU (two host threads,
// or two streams) race on it.  Every call writes every word it reads back, so
// nothing needs clearing between calls.
#include <mutex>

#include "kt_internal.h"

namespace kt {

namespace {

constexpr uint32_t kNoAllocChunk = 4096;
constexpr uint32_t kNoAllocMaxChunks = 16383;  // chunk sums + total: 16384 doubles, 128 KiB
constexpr int kNoAllocMaxDevices = 64;
constexpr int kNoAllocBlock = 256;

}  // namespace

// The chunk sums ([0, chunks)) and the total ([chunks]) of the last call on
// this GPU.  Host code never uses this name as an address: it uses the
// per-device pointer that hipGetSymbolAddress returned.
static __device__ __align__(256) double scratch[kNoAllocMaxChunks + 1];

namespace {

std::mutex g_noAllocMutex;
double *g_noAllocScratch[kNoAllocMaxDevices] = {};

// One block per 4096-net chunk: the nets' half-perimeters into LDS, then one
// thread adds them in net order.
__global__ void __launch_bounds__(kNoAllocBlock)
hpwlChunkNoAlloc(const uint32_t *__restrict__ netStart, const uint32_t *__restrict__ pinCell,
                 const double *__restrict__ offX, const double *__restrict__ offY,
                 const uint32_t *__restrict__ varOfVertex, const double *__restrict__ x,
                 const double *__restrict__ y, const double *__restrict__ fixedX,
                 const double *__restrict__ fixedY, uint32_t N, double *__restrict__ part) {
    __shared__ double perNet[kNoAllocChunk];
    const uint32_t first = blockIdx.x * kNoAllocChunk;
    const uint32_t count = min(kNoAllocChunk, N - first);
    for (uint32_t t = threadIdx.x; t < count; t += kNoAllocBlock) {
        const uint32_t j = first + t;
        const uint32_t b = netStart[j], e = netStart[j + 1];
        double h = 0.0;
        if (e - b >= 2) {
            double lx = INFINITY, ly = INFINITY, ux = -INFINITY, uy = -INFINITY;
            for (uint32_t q = b; q < e; ++q) {
                const uint32_t c = pinCell[q];
                const uint32_t v = varOfVertex[c];
                const double px = (v != kNoVar ? x[v] : fixedX[c]) + offX[q];
                const double py = (v != kNoVar ? y[v] : fixedY[c]) + offY[q];
                lx = px < lx ? px : lx;
                ly = py < ly ? py : ly;
                ux = ux < px ? px : ux;
                uy = uy < py ? py : uy;
            }
            h = (lx > ux || ly > uy) ? 0.0 : (ux - lx) + (uy - ly);
        }
        perNet[t] = h;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (uint32_t t = 0; t < count; ++t) s += perNet[t];
        part[blockIdx.x] = s;
    }
}

__global__ void hpwlTotalNoAlloc(double *__restrict__ part, uint32_t chunks) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double t = 0.0;
    for (uint32_t k = 0; k < chunks; ++k) t += part[k];
    part[chunks] = t;
}

}  // namespace

// With `device` current: look up this device's copy of `scratch` once.
int noAllocInit(int device) {
    KT_CHECK(device >= 0 && device < kNoAllocMaxDevices, KT_EINVAL, "no-alloc scratch: device index too large");
    std::lock_guard<std::mutex> lock(g_noAllocMutex);
    if (g_noAllocScratch[device] != nullptr) return KT_OK;
    void *p = nullptr;
    KT_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    KT_CHECK(p != nullptr, KT_EHIP, "no-alloc scratch: hipGetSymbolAddress gave no address");
    g_noAllocScratch[device] = static_cast<double *>(p);
    return KT_OK;
}

}  // namespace kt

extern "C" int kt_hpwl_noalloc(kt_ctx *ctx, double *hpwl) {
    using namespace kt;
    if (ctx == nullptr) {
        setError("null context");
        return KT_EINVAL;
    }
    KT_CHECK(hpwl != nullptr, KT_EINVAL, "kt_hpwl_noalloc: null output");
    KT_HIP(hipSetDevice(ctx->device));
    KT_CHECK(ctx->haveNetlist, KT_ESTATE, "kt_hpwl_noalloc: no netlist");
    KT_CHECK(ctx->posCount == ctx->n, KT_ESTATE, "kt_hpwl_noalloc: positions must have one entry per movable cell");
    const uint32_t N = ctx->N;
    const uint32_t chunks = (N + kNoAllocChunk - 1) / kNoAllocChunk;
    KT_CHECK(chunks <= kNoAllocMaxChunks, KT_EINVAL,
             "kt_hpwl_noalloc: more than 67,104,768 nets (16383 chunks of 4096) do not fit its fixed scratch");
    double *part = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_noAllocMutex);
        part = g_noAllocScratch[ctx->device];
    }
    KT_CHECK(part != nullptr, KT_ESTATE, "kt_hpwl_noalloc: scratch of this device not set up");
    if (chunks) {
        hpwlChunkNoAlloc<<<dim3(chunks), kNoAllocBlock, 0, ctx->stream>>>(
            ctx->netStart.p, ctx->pinCell.p, ctx->offX.p, ctx->offY.p, ctx->varOfVertex.p, ctx->posX.p, ctx->posY.p,
            ctx->fixedX.p, ctx->fixedY.p, N, part);
        KT_HIP(hipGetLastError());
    }
    hpwlTotalNoAlloc<<<1, 64, 0, ctx->stream>>>(part, chunks);
    KT_HIP(hipGetLastError());
    double h = 0.0;
    KT_HIP(hipMemcpyAsync(&h, part + chunks, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KT_HIP(hipStreamSynchronize(ctx->stream));
    *hpwl = h;
    return KT_OK;
}
