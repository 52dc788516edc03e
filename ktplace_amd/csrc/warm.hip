// Learned warm start of the x/y solve (no counterpart in the reference; the
// reference starts CG from the previous iterate, simpl_lss.cc:214-222).
//
// At the current positions x and the freshly built systems A x = b:
//   1. global statistics of the placement (centroid, spread, mean diagonal);
//   2. r = b - A x and the per-cell features (KT_MLP_FEATURES, below);
//   3. the MLP predicts each cell's displacement in units of the spread:
//      delta = s * MLP(features);
//   4. exact line search along delta on the quadratic the solve minimises,
//      f(x) = x'Ax/2 - b'x:  gamma = (delta . r) / (delta . A delta);
//   5. x += gamma delta when gamma > 0, else the previous iterate is kept.
// Step 4 makes the previous-iterate start the fallback: f never increases, so
// the A-norm error of the start never grows, whatever the network predicts.
// All sums are in a fixed order, so the start is deterministic.
//
// Features of cell i (scale s_x, s_y = standard deviation of the movable
// cells' positions; c_x, c_y their mean):
//   0,1   (x - c_x) / s_x, (y - c_y) / s_y
//   2,3   position relative to the die, 2 (x - die centre) / die width (height)
//   4,5   s_x / die width, s_y / die height
//   6     log2(1 + pins of the cell) / 4
//   7,8   log(diag_i / mean diag) / 4 per axis, clamped to +-2
//   9     occupancy / free area of the cell's density bin, / 4, at most 2
//   10,11 Jacobi step r_i / diag_i / s per axis, clamped to +-4
//   12,13 anchor pull (anchor - x) / s per axis (0 without anchors), clamped
//   14    min(alpha, 4) / 4 of the last B2B build
//   15,16 share of the diagonal not coupled to movable cells per axis,
//         (diag_i - sum_j |a_ij|) / diag_i: pads, fixed cells and anchors
#include <algorithm>
#include <cmath>

#include "kt_internal.h"

namespace kt {

int lalOccupancy(kt_ctx *ctx, std::vector<double> *occHost, std::vector<uint32_t> *countHost = nullptr);
int mlpInferLaunch(hipStream_t stream, const uint16_t *w, const float *b, const uint16_t *feat, uint64_t count,
                   float *out);
int mlpPackLaunch(hipStream_t stream, const float *params, uint16_t *w, float *b);
uint32_t mlpWeightHalfs();
uint32_t mlpBiasFloats();

namespace {

constexpr int kB = 256;
constexpr uint32_t kChunk = 4096;

struct AxisSys {
    const uint32_t *rowPtr, *col;
    const double *val, *diag, *rhs;
};

struct WarmScalars {
    double cx, cy, sx, sy, mdx, mdy;
    double gamma[2], num[2], den[2];
};

__device__ __forceinline__ uint16_t toBf16(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

template <int K>
__device__ __forceinline__ void blockSums(double (&v)[K], double (*s)[kB]) {
#pragma unroll
    for (int k = 0; k < K; ++k) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int off = kB / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off)
#pragma unroll
            for (int k = 0; k < K; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + off];
        __syncthreads();
    }
}

// per 4096-row chunk: sum x, x^2, y, y^2, diag_x, diag_y
__global__ void __launch_bounds__(kB) warmStats(const double *__restrict__ x, const double *__restrict__ y,
                                                const double *__restrict__ dx, const double *__restrict__ dy,
                                                uint32_t n, double *__restrict__ part) {
    __shared__ double s[6][kB];
    double v[6] = {0, 0, 0, 0, 0, 0};
    const uint64_t lo = static_cast<uint64_t>(blockIdx.x) * kChunk;
    const uint64_t hi = std::min<uint64_t>(lo + kChunk, n);
    for (uint64_t i = lo + threadIdx.x; i < hi; i += kB) {
        v[0] += x[i];
        v[1] += x[i] * x[i];
        v[2] += y[i];
        v[3] += y[i] * y[i];
        v[4] += dx[i];
        v[5] += dy[i];
    }
    blockSums<6>(v, s);
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = s[threadIdx.x][0];
}

// One wave: lane l sums chunks l, l + 64, ... in order, then lane 0 adds the
// 64 lane sums in lane order (a fixed order, so the result is reproducible).
__global__ void __launch_bounds__(64) warmStatsFinish(const double *__restrict__ part, uint32_t chunks, uint32_t n,
                                                      double dieW, double dieH, WarmScalars *__restrict__ sc) {
    __shared__ double lanes[6][64];
    double t[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t k = threadIdx.x; k < chunks; k += 64)
#pragma unroll
        for (int j = 0; j < 6; ++j) t[j] += part[k * 6 + j];
#pragma unroll
    for (int j = 0; j < 6; ++j) lanes[j][threadIdx.x] = t[j];
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int j = 0; j < 6; ++j) {
        t[j] = 0.0;
        for (int l = 0; l < 64; ++l) t[j] += lanes[j][l];
    }
    const double inv = 1.0 / static_cast<double>(n);
    sc->cx = t[0] * inv;
    sc->cy = t[2] * inv;
    const double vx = t[1] * inv - sc->cx * sc->cx, vy = t[3] * inv - sc->cy * sc->cy;
    sc->sx = fmax(sqrt(fmax(vx, 0.0)), 1e-6 * dieW + 1e-300);
    sc->sy = fmax(sqrt(fmax(vy, 0.0)), 1e-6 * dieH + 1e-300);
    sc->mdx = t[4] * inv;
    sc->mdy = t[5] * inv;
}

constexpr int kG = 4;               // lanes per CSR row in the row passes
constexpr int kPer = 4;             // entries per lane loaded together
constexpr int kRowsPerBlock = kB / kG;

// One pass over the rows of one axis (blockIdx.y), kG lanes per row:
//   residual mode: r_i = b_i - (A v)_i and the share of diag_i not coupled to
//                  movable cells, (diag_i - sum_j |a_ij|) / diag_i;
//   curvature mode: per block, sum_i d_i r_i and sum_i d_i (A d)_i with d = v.
template <bool kCurvature>
__global__ void __launch_bounds__(kB) warmRows(AxisSys sx, AxisSys sy, const double *__restrict__ vx,
                                               const double *__restrict__ vy, double *__restrict__ rx,
                                               double *__restrict__ ry, float *__restrict__ fsx,
                                               float *__restrict__ fsy, uint32_t n, double *__restrict__ part) {
    const int axis = blockIdx.y;
    const AxisSys s = axis ? sy : sx;
    const double *__restrict__ v = axis ? vy : vx;
    double *__restrict__ r = axis ? ry : rx;
    const int lane = threadIdx.x % kG;
    const uint32_t row = blockIdx.x * kRowsPerBlock + threadIdx.x / kG;
    double t = 0.0, off = 0.0;
    // the row's own data is loaded first, so it arrives with the products
    const uint32_t rq = row < n ? row : 0u;  // n > 0
    const double dRow = s.diag[rq], vRow = v[rq], bRow = kCurvature ? r[rq] : s.rhs[rq];
    if (row < n) {
        // a lane's entries k, k + kG, ..., k + (kPer - 1) kG are loaded
        // together (a row of up to kPer kG entries in one round), and summed in
        // that order
        const uint32_t k1 = s.rowPtr[row + 1];
        for (uint32_t k = s.rowPtr[row] + lane; k < k1; k += kPer * kG) {
            double a[kPer], g[kPer];
            uint32_t c[kPer];
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const uint32_t kk = k + u * kG < k1 ? k + u * kG : k;  // unconditional loads
                a[u] = s.val[kk];
                c[u] = s.col[kk];
            }
#pragma unroll
            for (int u = 0; u < kPer; ++u) g[u] = v[c[u]];
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                if (k + u * kG < k1) {
                    t += a[u] * g[u];
                    off += fabs(a[u]);
                }
            }
        }
    }
#pragma unroll
    for (int o = kG / 2; o > 0; o >>= 1) {
        t += __shfl_xor(t, o);
        off += __shfl_xor(off, o);
    }
    if constexpr (!kCurvature) {
        if (row < n && lane == 0) {
            const double d = dRow;
            r[row] = bRow - (d * vRow + t);
            const float share = d > 0.0 ? static_cast<float>(fmin(fmax((d - off) / d, 0.0), 1.0)) : 0.f;
            (axis ? fsy : fsx)[row] = share;
        }
    } else {
        __shared__ double sh[2][kB];
        double num = 0.0, den = 0.0;
        if (row < n && lane == 0) {
            const double di = vRow;
            num = di * bRow;
            den = di * (dRow * di + t);
        }
        sh[0][threadIdx.x] = num;
        sh[1][threadIdx.x] = den;
        __syncthreads();
        for (int o = kB / 2; o > 0; o >>= 1) {
            if (threadIdx.x < o) {
                sh[0][threadIdx.x] += sh[0][threadIdx.x + o];
                sh[1][threadIdx.x] += sh[1][threadIdx.x + o];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            part[(static_cast<uint64_t>(axis) * gridDim.x + blockIdx.x) * 2 + 0] = sh[0][0];
            part[(static_cast<uint64_t>(axis) * gridDim.x + blockIdx.x) * 2 + 1] = sh[1][0];
        }
    }
}

__device__ __forceinline__ float clampf(double v, double lo, double hi) {
    return static_cast<float>(v < lo ? lo : (v > hi ? hi : v));
}

struct FeatArgs {
    AxisSys sys[2];
    const double *x, *y, *ax, *ay;
    const uint32_t *pinStart, *binOf;
    const double *occ, *avail;
    const WarmScalars *sc;
    const double *r[2];      // b - A x per axis
    const float *fs[2];      // fixed share of the diagonal per axis
    uint16_t *feat;   // [n][32] BF16
    float *feat32;    // [n][KT_MLP_FEATURES] or null
    double dieCx, dieCy, dieW, dieH, alpha;
    uint32_t n;
    int anchored, haveGrid;
};

__global__ void __launch_bounds__(kB) warmFeatures(FeatArgs a) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= a.n) return;
    const WarmScalars sc = *a.sc;
    const double rx = a.r[0][i], ry = a.r[1][i];
    const double x = a.x[i], y = a.y[i];
    const double dgx = a.sys[0].diag[i], dgy = a.sys[1].diag[i];
    float f[KT_MLP_FEATURES];
    f[0] = static_cast<float>((x - sc.cx) / sc.sx);
    f[1] = static_cast<float>((y - sc.cy) / sc.sy);
    f[2] = static_cast<float>(2.0 * (x - a.dieCx) / a.dieW);
    f[3] = static_cast<float>(2.0 * (y - a.dieCy) / a.dieH);
    f[4] = static_cast<float>(sc.sx / a.dieW);
    f[5] = static_cast<float>(sc.sy / a.dieH);
    f[6] = static_cast<float>(log2(1.0 + static_cast<double>(a.pinStart[i + 1] - a.pinStart[i])) * 0.25);
    f[7] = clampf(log(dgx / sc.mdx) * 0.25, -2.0, 2.0);
    f[8] = clampf(log(dgy / sc.mdy) * 0.25, -2.0, 2.0);
    float ov = 0.f;
    if (a.haveGrid) {
        const uint32_t b = a.binOf[i];
        const double av = a.avail[b], oc = a.occ[b];
        ov = av > 0.0 ? clampf(oc / av * 0.25, 0.0, 2.0) : (oc > 0.0 ? 2.f : 0.f);
    }
    f[9] = ov;
    f[10] = clampf(rx / dgx / sc.sx, -4.0, 4.0);
    f[11] = clampf(ry / dgy / sc.sy, -4.0, 4.0);
    f[12] = a.anchored ? clampf((a.ax[i] - x) / sc.sx, -4.0, 4.0) : 0.f;
    f[13] = a.anchored ? clampf((a.ay[i] - y) / sc.sy, -4.0, 4.0) : 0.f;
    f[14] = static_cast<float>(fmin(a.alpha, 4.0) * 0.25);
    f[15] = a.fs[0][i];
    f[16] = a.fs[1][i];
    union {
        uint16_t h[KT_MLP_INPUTS];
        uint4 q[KT_MLP_INPUTS / 8];
    } row;
#pragma unroll
    for (int k = 0; k < KT_MLP_INPUTS; ++k) row.h[k] = k < KT_MLP_FEATURES ? toBf16(f[k]) : 0;
    uint4 *dst = reinterpret_cast<uint4 *>(a.feat + static_cast<uint64_t>(i) * KT_MLP_INPUTS);
#pragma unroll
    for (int k = 0; k < KT_MLP_INPUTS / 8; ++k) dst[k] = row.q[k];
    if (a.feat32)
#pragma unroll
        for (int k = 0; k < KT_MLP_FEATURES; ++k) a.feat32[static_cast<uint64_t>(i) * KT_MLP_FEATURES + k] = f[k];
}

// delta = s * prediction; zero for a cell with no net path to a fixed vertex
// when the systems have no anchors (keepFloating == 0): moving such a cell's
// component is in the null space of A, so no solve would ever undo it and the
// solution would leave the reference's (the previous iterate's) there.
__global__ void __launch_bounds__(kB) warmDelta(const float2 *__restrict__ pred, const WarmScalars *__restrict__ sc,
                                                const uint8_t *__restrict__ floating, int keepFloating, uint32_t n,
                                                double *__restrict__ dx, double *__restrict__ dy) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const float2 p = pred[i];
    const bool off = !keepFloating && floating[i];
    dx[i] = off ? 0.0 : sc->sx * static_cast<double>(p.x);
    dy[i] = off ? 0.0 : sc->sy * static_cast<double>(p.y);
}

// gamma per axis from the per-block partials (fixed order: thread t sums
// blocks t, t + kB, ..., then a fixed tree)
__global__ void __launch_bounds__(kB) warmGamma(const double *__restrict__ part, uint32_t blocks,
                                                WarmScalars *__restrict__ sc) {
    __shared__ double sh[4][kB];
    double v[4] = {0, 0, 0, 0};
    for (uint32_t b = threadIdx.x; b < blocks; b += kB)
        for (int a = 0; a < 2; ++a) {
            v[2 * a] += part[(static_cast<uint64_t>(a) * blocks + b) * 2];
            v[2 * a + 1] += part[(static_cast<uint64_t>(a) * blocks + b) * 2 + 1];
        }
    blockSums<4>(v, sh);
    if (threadIdx.x == 0) {
        for (int a = 0; a < 2; ++a) {
            const double num = sh[2 * a][0], den = sh[2 * a + 1][0];
            const double g = den > 0.0 ? num / den : 0.0;
            sc->num[a] = num;
            sc->den[a] = den;
            sc->gamma[a] = (g > 0.0 && isfinite(g)) ? g : 0.0;
        }
    }
}

__global__ void __launch_bounds__(kB) warmApply(const WarmScalars *__restrict__ sc, const double *__restrict__ dx,
                                                const double *__restrict__ dy, uint32_t n, double *__restrict__ x,
                                                double *__restrict__ y) {
    const uint32_t i = blockIdx.x * kB + threadIdx.x;
    if (i >= n) return;
    const double gx = sc->gamma[0], gy = sc->gamma[1];
    if (gx > 0.0) x[i] += gx * dx[i];
    if (gy > 0.0) y[i] += gy * dy[i];
}

inline dim3 gridFor(uint64_t n) { return dim3(static_cast<uint32_t>(std::max<uint64_t>((n + kB - 1) / kB, 1))); }

AxisSys axisSys(kt_ctx *ctx, int a) {
    const Sys &s = ctx->sys[a];
    return AxisSys{s.rowPtr.p, s.col.p, s.val.p, s.diag.p, s.rhs.p};
}

int warmCheck(kt_ctx *ctx) {
    KT_CHECK(ctx->haveNetlist && ctx->haveDesign, KT_ESTATE, "warm start: netlist and rows needed");
    KT_CHECK(ctx->n > 0 && ctx->posCount == ctx->n, KT_ESTATE, "warm start: positions must cover the movable cells");
    for (int a = 0; a < 2; ++a)
        KT_CHECK(ctx->sys[a].rows == ctx->n && ctx->sys[a].rowPtr.p, KT_ESTATE,
                 "warm start: build the systems at the current positions first");
    return KT_OK;
}

// Steps 1-2: statistics, residual and features (feat32: optional FP32 copy).
int warmFeaturesRun(kt_ctx *ctx, float *feat32) {
    KT_TRY(warmCheck(ctx));
    const uint32_t n = ctx->n;
    const uint32_t chunks = (n + kChunk - 1) / kChunk;
    const bool haveGrid = ctx->grid.size() > 0 && ctx->gridAvail.cap >= ctx->grid.size();
    if (haveGrid) KT_TRY(lalOccupancy(ctx, nullptr, nullptr));
    KT_TRY(ctx->warmPart.reserve(static_cast<uint64_t>(chunks) * 6));
    KT_TRY(ctx->warmScalars.reserve(sizeof(WarmScalars) / sizeof(double)));
    KT_TRY(ctx->warmFeat.reserve(static_cast<uint64_t>(n) * KT_MLP_INPUTS));
    for (int a = 0; a < 2; ++a) KT_TRY(ctx->r[a].reserve(n));
    const lal::Box &die = ctx->lalDie;
    const double dieW = std::max(die.x1 - die.x0, 1e-12), dieH = std::max(die.y1 - die.y0, 1e-12);
    auto *sc = reinterpret_cast<WarmScalars *>(ctx->warmScalars.p);
    warmStats<<<chunks, kB, 0, ctx->stream>>>(ctx->posX.p, ctx->posY.p, ctx->sys[0].diag.p, ctx->sys[1].diag.p, n,
                                             ctx->warmPart.p);
    KT_HIP(hipGetLastError());
    warmStatsFinish<<<1, 64, 0, ctx->stream>>>(ctx->warmPart.p, chunks, n, dieW, dieH, sc);
    KT_HIP(hipGetLastError());
    KT_TRY(ctx->warmShare.reserve(2ull * n));
    const dim3 rowGrid((n + kRowsPerBlock - 1) / kRowsPerBlock, 2);
    warmRows<false><<<rowGrid, kB, 0, ctx->stream>>>(axisSys(ctx, 0), axisSys(ctx, 1), ctx->posX.p, ctx->posY.p,
                                                     ctx->r[0].p, ctx->r[1].p, ctx->warmShare.p,
                                                     ctx->warmShare.p + n, n, nullptr);
    KT_HIP(hipGetLastError());
    FeatArgs fa{};
    fa.sys[0] = axisSys(ctx, 0);
    fa.sys[1] = axisSys(ctx, 1);
    fa.x = ctx->posX.p;
    fa.y = ctx->posY.p;
    fa.anchored = ctx->haveAnchors && ctx->lastAlpha > 0.0;
    fa.ax = fa.anchored ? ctx->anchorX.p : nullptr;
    fa.ay = fa.anchored ? ctx->anchorY.p : nullptr;
    fa.pinStart = ctx->pinStart.p;
    fa.haveGrid = haveGrid;
    fa.binOf = haveGrid ? ctx->binOf.p : nullptr;
    fa.occ = haveGrid ? ctx->occ.p : nullptr;
    fa.avail = haveGrid ? ctx->gridAvail.p : nullptr;
    fa.sc = sc;
    fa.r[0] = ctx->r[0].p;
    fa.r[1] = ctx->r[1].p;
    fa.fs[0] = ctx->warmShare.p;
    fa.fs[1] = ctx->warmShare.p + n;
    fa.feat = ctx->warmFeat.p;
    fa.feat32 = feat32;
    fa.dieCx = 0.5 * (die.x0 + die.x1);
    fa.dieCy = 0.5 * (die.y0 + die.y1);
    fa.dieW = dieW;
    fa.dieH = dieH;
    fa.alpha = fa.anchored ? ctx->lastAlpha : 0.0;
    fa.n = n;
    warmFeatures<<<gridFor(n), kB, 0, ctx->stream>>>(fa);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

// Step 3: delta into p[0], p[1].
int warmPredictRun(kt_ctx *ctx) {
    KT_CHECK(ctx->haveWarm, KT_ESTATE, "warm start: no network parameters set");
    const uint32_t n = ctx->n;
    KT_TRY(ctx->warmOut.reserve(static_cast<uint64_t>(n) * 2));
    for (int a = 0; a < 2; ++a) KT_TRY(ctx->p[a].reserve(n));
    hipEvent_t t0{};
    if (ctx->profile) KT_TRY(timedBegin(ctx, KT_KERNEL_MLP, &t0));
    KT_TRY(mlpInferLaunch(ctx->stream, ctx->warmW.p, ctx->warmB.p, ctx->warmFeat.p, n, ctx->warmOut.p));
    if (ctx->profile) KT_TRY(timedEnd(ctx, KT_KERNEL_MLP, t0));
    warmDelta<<<gridFor(n), kB, 0, ctx->stream>>>(reinterpret_cast<const float2 *>(ctx->warmOut.p),
                                                  reinterpret_cast<const WarmScalars *>(ctx->warmScalars.p),
                                                  ctx->floating.p, ctx->lastAlpha > 0.0 ? 1 : 0, n, ctx->p[0].p,
                                                  ctx->p[1].p);
    KT_HIP(hipGetLastError());
    return KT_OK;
}

}  // namespace

int warmSetParams(kt_ctx *ctx, const float *params, uint64_t count) {
    if (!params) {
        ctx->haveWarm = false;
        return KT_OK;
    }
    KT_CHECK(count == KT_MLP_PARAMS, KT_EINVAL, "kt_set_warm_start: need KT_MLP_PARAMS floats");
    for (uint64_t i = 0; i < count; ++i)
        KT_CHECK(std::isfinite(params[i]), KT_EINVAL, "kt_set_warm_start: non-finite parameter");
    DevBuf<float> master;
    KT_TRY(master.reserve(count));
    KT_TRY(ctx->warmW.reserve(mlpWeightHalfs()));
    KT_TRY(ctx->warmB.reserve(mlpBiasFloats()));
    hipError_t e = hipMemsetAsync(ctx->warmW.p, 0, ctx->warmW.bytes(), ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(master.p, params, count * sizeof(float), hipMemcpyHostToDevice, ctx->stream);
    int rc = e == hipSuccess ? mlpPackLaunch(ctx->stream, master.p, ctx->warmW.p, ctx->warmB.p) : KT_OK;
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    master.release();
    if (e != hipSuccess) return hipFail(e, "kt_set_warm_start", __FILE__, __LINE__);
    KT_TRY(rc);
    ctx->haveWarm = true;
    return KT_OK;
}

int warmFeaturesHost(kt_ctx *ctx, float *features, double *scale) {
    DevBuf<float> f32;
    if (features) KT_TRY(f32.reserve(static_cast<uint64_t>(ctx->n) * KT_MLP_FEATURES));
    int rc = warmFeaturesRun(ctx, features ? f32.p : nullptr);
    hipError_t e = hipSuccess;
    WarmScalars sc{};
    if (rc == KT_OK) {
        if (features)
            e = hipMemcpyAsync(features, f32.p, static_cast<uint64_t>(ctx->n) * KT_MLP_FEATURES * sizeof(float),
                               hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(&sc, ctx->warmScalars.p, sizeof(sc), hipMemcpyDeviceToHost, ctx->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    }
    f32.release();
    if (e != hipSuccess) return hipFail(e, "kt_warm_features", __FILE__, __LINE__);
    KT_TRY(rc);
    if (scale) {
        scale[0] = sc.sx;
        scale[1] = sc.sy;
    }
    return KT_OK;
}

int warmPredictHost(kt_ctx *ctx, double *dx, double *dy) {
    KT_TRY(warmFeaturesRun(ctx, nullptr));
    KT_TRY(warmPredictRun(ctx));
    const uint64_t n = ctx->n;
    KT_HIP(hipMemcpyAsync(dx, ctx->p[0].p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KT_HIP(hipMemcpyAsync(dy, ctx->p[1].p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    KT_HIP(hipStreamSynchronize(ctx->stream));
    return KT_OK;
}

// The whole warm start on the context's stream; gamma per axis is copied to
// ctx->pinned[62..63] (valid after the next stream synchronisation).
int warmStart(kt_ctx *ctx) {
    hipEvent_t t0{};
    if (ctx->profile) KT_TRY(timedBegin(ctx, KT_KERNEL_WARM, &t0));
    KT_TRY(warmFeaturesRun(ctx, nullptr));
    KT_TRY(warmPredictRun(ctx));
    const uint32_t n = ctx->n;
    auto *sc = reinterpret_cast<WarmScalars *>(ctx->warmScalars.p);
    const uint32_t rowBlocks = (n + kRowsPerBlock - 1) / kRowsPerBlock;
    KT_TRY(ctx->warmPart.reserve(4ull * rowBlocks));
    warmRows<true><<<dim3(rowBlocks, 2), kB, 0, ctx->stream>>>(axisSys(ctx, 0), axisSys(ctx, 1), ctx->p[0].p,
                                                               ctx->p[1].p, ctx->r[0].p, ctx->r[1].p, nullptr,
                                                               nullptr, n, ctx->warmPart.p);
    KT_HIP(hipGetLastError());
    warmGamma<<<1, kB, 0, ctx->stream>>>(ctx->warmPart.p, rowBlocks, sc);
    KT_HIP(hipGetLastError());
    warmApply<<<gridFor(n), kB, 0, ctx->stream>>>(sc, ctx->p[0].p, ctx->p[1].p, n, ctx->posX.p, ctx->posY.p);
    KT_HIP(hipGetLastError());
    KT_HIP(hipMemcpyAsync(ctx->pinned + 62, &sc->gamma[0], 2 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    if (ctx->profile) KT_TRY(timedEnd(ctx, KT_KERNEL_WARM, t0));
    return KT_OK;
}

}  // namespace kt
