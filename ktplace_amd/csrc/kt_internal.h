// Internal definitions shared by the HIP translation units of libktplace_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "ktplace_amd.h"
#include "lal_geom.h"

namespace kt {

// ------------------------------------------------------------------ errors
void setError(const std::string &msg);
int hipFail(hipError_t e, const char *what, const char *file, int line);

#define KT_HIP(call)                                                        \
    do {                                                                    \
        hipError_t kt_e_ = (call);                                          \
        if (kt_e_ != hipSuccess) return ::kt::hipFail(kt_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define KT_CHECK(cond, code, msg)           \
    do {                                    \
        if (!(cond)) {                      \
            ::kt::setError(msg);            \
            return (code);                  \
        }                                   \
    } while (0)

#define KT_TRY(expr)                        \
    do {                                    \
        int kt_rc_ = (expr);                \
        if (kt_rc_ != KT_OK) return kt_rc_; \
    } while (0)

// ------------------------------------------------------------------ buffers
// Grow-only device buffer.  Reallocation drops the old contents.  Owns its
// allocation (freed on destruction, on the device current at that time).
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;

    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }

    int reserve(size_t count) {
        if (count <= cap) return KT_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        const size_t bytes = (count == 0 ? 1 : count) * sizeof(T);
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), bytes);
        if (e != hipSuccess) return hipFail(e, "hipMalloc", __FILE__, __LINE__);
        cap = count;
        return KT_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
    size_t bytes() const { return cap * sizeof(T); }
};

constexpr uint32_t kNoVar = KT_NO_VAR;
// Nets with more pins than this get a wave of their own in the B2B build's
// per-net extremes (the others one thread each).
constexpr uint32_t kWideNetPins = 32;
// Spare entries after a system's column and value arrays: the entry-block
// SpMV's wide loads read whole aligned groups of four entries.
constexpr uint64_t kQuadPad = 4;

// CG state of one axis at the start of an iteration.  Double buffered by
// iteration parity (AxisCtl::s[k & 1]).
struct CgState {
    double rho;     // r.z
    double resid;   // ||r|| / ||b|| (recurrence; FP64 after a mixed-precision refresh)
    uint32_t steps; // iterations this axis has stepped
    int32_t live;   // steps this iteration: active && resid > tol && steps < maxIter
    int32_t active; // the reference's Axis::active
    int32_t pad;    // fused FP32 iteration: tiles of the r.r partials rho is summed from (0: rho)
};

// Per-axis control block in device memory.
struct AxisCtl {
    CgState s[2];
    double bNorm;         // ||b||
    double tol;
    double trueResid;     // ||b - A x|| / ||b|| from the last FP64 pass
    double refResid;      // residual at the last mixed-precision refresh
    double refRz;         // r.z just after that refresh (or the start): on the scaled FP32
                          // system the iterations judge resid = refResid sqrt(r.z / refRz)
    double refreshFactor; // refresh once the FP32 residual falls by this factor
    uint32_t maxIter;
    int32_t brk[2];       // p.Ap <= 0 in the iteration of that parity
    double alpha[2];      // rho / p.Ap of the iteration of that parity (cgUpdate -> cgPUpdate)
    int32_t frozen;       // excluded from a restart (converged or broken down)
    int32_t needUpdate;   // mixed precision: FP64 refresh requested
    uint32_t refreshes;
};

// One axis' CSR as the CG kernels see it.
struct SysView {
    const uint32_t *rowPtr;
    const uint32_t *col;
    const double *val;
    const double *diag;
    const double *rhs;
    const float *val32;
    const float *diag32;
    uint32_t rows;
    uint32_t blocks;  // blocks this axis occupies in a two-axis launch
    // entry-block SpMV layout: block b takes rows [x, y) of flatInfo[b] =
    // (first row, end row, first entry, end entry)
    const uint4 *flatInfo;
    uint32_t sblocks;  // SpMV blocks (and p.Ap partials) of this axis
    // entry-block SpMV with 3-byte columns: col = colLo | colHi << 16
    const uint16_t *colLo;
    const uint8_t *colHi;
    // entry-block SpMV with wide loads and 3-byte columns: the columns of
    // entries 4q .. 4q + 3 packed little-endian into words 3q .. 3q + 2
    const uint32_t *colPk;
    const double *isd;  // mixed precision, scaled: 1 / sqrt(d) per row (null otherwise)
};

struct Sys {
    uint64_t rows = 0;
    uint64_t nnzCap = 0;  // entries allocated
    int64_t nnzHost = -1; // nnz when known on the host (-1: read rowPtr[rows])
    DevBuf<uint32_t> rowPtr, col;
    DevBuf<double> val, diag, rhs;
    DevBuf<float> val32, diag32;
    bool have32 = false;
    bool scaled32 = false;  // val32 holds val_ij (s_i s_j), s = 1 / sqrt(d), rather than val_ij
    DevBuf<double> isd;     // s = 1 / sqrt(d) per row (the scaled FP32 system)
    DevBuf<uint16_t> colLo;  // 3-byte columns: low 16 bits
    DevBuf<uint8_t> colHi;   // 3-byte columns: high 8 bits
    bool have24 = false;
    DevBuf<uint32_t> colPk;  // 3-byte columns packed four to three words (wide loads)
    bool havePk = false;
    DevBuf<uint32_t> rowStart;  // entry-block SpMV: (first row, end row, first entry, end entry) per block
};

// Arguments of every CG kernel (both axes).
struct CgArgs {
    SysView sys[2];
    double *x64[2];     // FP64 iterate (the context's positions)
    void *r[2];         // residual, working precision
    void *p[2];         // search direction
    void *ap[2];        // A p
    void *e[2];         // FP32 correction to x64 (mixed precision), else null
    double *partA;      // [axis][stride]    p.Ap per tile
    double *partB;      // [axis][stride][2] r.r, r.z per tile
    double *partS;      // [axis][stride][3] start partials
    double *partU;      // fused iteration: [parity][axis][uStride] r'.r' per update tile
    uint32_t uStride;
    AxisCtl *ctl;       // [2]
    double *history;    // [axis][histCap] residual after each step
    uint32_t histCap;
    uint32_t stride;    // partial slots per axis
    uint32_t tile;      // rows per block
    uint32_t xcdSplit;  // 1: blocks on XCDs 0-3 take axis x, 4-7 axis y
    uint32_t halo;      // SpMV: entries of p staged in LDS on each side of a tile (0: none)
    uint32_t layout;    // SpMV: 0 lane groups per row, 1 entry blocks (KT_SPMV_ENTRY_BLOCKS)
    uint32_t flatEntries;  // entry blocks: entries per block (4096, 8192, 16384)
    uint32_t flatUnroll;   // entry blocks: entries per thread and round (4, 8)
    uint32_t flatRowLanes; // entry blocks: lanes per row in the row sums (1, 4)
    uint32_t flatThreads;  // entry blocks: threads per block (256, 512, 1024)
    uint32_t col24;        // entry blocks: 1 = columns read as 16 + 8 bits
    uint32_t flatResidual; // 1: FP64 residuals (refresh, true residual) by entry blocks of 256 threads
    uint32_t wide;         // entry blocks: 1 = four consecutive entries per thread by wide loads
    uint32_t unitDiag;     // mixed precision: the FP32 system is D^-1/2 A D^-1/2 (unit diagonal);
                           // r, p, Ap and the correction e live in those coordinates
    uint32_t fused;        // mixed precision, scaled: two kernels per iteration (the SpMV also
                           // sums r.Ap and Ap.Ap; cgFUpdate does both vector updates and sums
                           // r'.r' per tile for the next iteration)
};

// Launch-shape knobs (set by the autotuner or the caller; 0 = default).
struct Tune {
    int lanes = 0;            // lanes per row in the SpMV (1..16)
    int rows = 0;             // rows per lane group in flight in the SpMV (2, 4, 8)
    uint32_t tile = 0;        // rows per block (multiple of 1024)
    bool useGraphs = true;
    int xcdSplit = -1;        // -1 default (on), 0 off, 1 on
    double refreshFactor = 0; // mixed precision refresh factor (default 0.05)
    uint32_t halo = 0;        // SpMV LDS window: entries on each side of a tile (0: no staging)
    int layout = 0;           // SpMV layout (KT_SPMV_ROW_GROUPS / KT_SPMV_ENTRY_BLOCKS)
    uint32_t flatEntries = 0; // entry blocks: entries per block (0: 8192)
    int flatUnroll = 0;       // entry blocks: entries per thread and round (0: 4)
    int flatRowLanes = 0;     // entry blocks: lanes per row in the row sums (0, 1: entry order; 4)
    int flatThreads = 0;      // entry blocks: threads per block (0: 1024; 256, 512)
    int colBytes = 0;         // entry blocks: bytes per column index (0, 4; 3 with 256 threads)
    int wideLoads = 0;        // entry blocks: four consecutive entries per thread by wide loads
    int mixedScaling = 1;     // mixed precision: FP32 solve on the symmetrically scaled system
    int fusedUpdate = 1;      // mixed precision, scaled: one vector kernel per iteration where the shape allows
};

// A captured hipGraph of kPairs iteration pairs for one argument set.
struct GraphCache {
    static constexpr int kPairs = 8;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    CgArgs args{};
    int lanes = 0;
    int rows = 0;
};

// Per-block statistics of one LAL level (cells in the block's sorted order).
struct LalSegStat {
    uint32_t count;
    uint32_t m;          // cells in the low half
    uint64_t cellArea;   // fixed point, 2^-16 units
    uint64_t lowArea;
};

struct LalOptions {
    size_t maxLevel = 20;          // simpl_options.h:24
    size_t minCellsToSplit = 4;    // :25
    double stripeAreaFraction = 0.1;  // :26
};

struct KernelTimer {
    uint64_t launches = 0;
    double ms = 0.0;
    std::vector<hipEvent_t> pending;  // start/stop pairs not yet read back
};

}  // namespace kt

struct kt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;

    // netlist (device)
    bool haveNetlist = false;
    uint32_t n = 0, V = 0, N = 0;
    uint64_t P = 0, Pm = 0, nnzBound = 0;
    kt::DevBuf<uint32_t> netStart, pinCell, varOfVertex, movVertex, pinStart, refNet, refPin, netMov;
    kt::DevBuf<uint32_t> wideNets;  // nets of more than kWideNetPins pins (one wave each in netExtremes)
    uint32_t numWideNets = 0;
    kt::DevBuf<double> offX, offY, weight, fixedX, fixedY, width, height, area;
    // per movable cell: 1 if no net path joins it to a fixed vertex (its
    // component's translation is in the null space of the systems without
    // anchors, so no start may move it)
    kt::DevBuf<uint8_t> floating;
    uint32_t floatingCells = 0;

    // positions (rows of the current systems) and anchors
    uint64_t posCount = 0;
    kt::DevBuf<double> posX, posY, saveX, saveY, anchorX, anchorY;
    bool haveAnchors = false;
    kt::DevBuf<double2> vxy;        // vertex-indexed positions (px, py), one 16-byte record per vertex
    kt::DevBuf<uint32_t> ext;       // [N][32]: per-net record of the B2B build (pins, weight, lo/hi per axis)
    kt::DevBuf<double> pinAt[2];    // per pin and axis: (vertex position + offset, rhs term of an edge to it)
    kt::DevBuf<uint32_t> pinVar;    // per pin: (its vertex's variable or kNoVar, movable pins before it in its net), uploaded
    kt::DevBuf<uint64_t> b2bEdgeOff, b2bEntOff;  // B2B: per reference, first edge / CSR slot (x | y << 32)
    kt::DevBuf<double> b2bEdges[2];              // B2B: (w, rhs term) per edge in row-walk order
    kt::DevBuf<uint32_t> b2bLong;                // B2B: first reference and first row of each edge block, per axis
    kt::DevBuf<uint32_t> b2bRowEdge;             // B2B: first edge of each row's run, per axis
    kt::DevBuf<uint32_t> scanTmp;   // block sums of the u32 scans
    kt::DevBuf<uint64_t> scanTmp64; // block sums of the u64 scans
    kt::DevBuf<uint8_t> scanStorage; // temporary storage of the single-pass device scans

    kt::Sys sys[2];

    // CG state
    kt::DevBuf<double> r[2], p[2], ap[2], xs[2];
    kt::DevBuf<float> r32[2], p32[2], ap32[2], x32[2];
    kt::DevBuf<double> partials;    // [2][maxBlocks][3]
    kt::DevBuf<kt::AxisCtl> ctl;    // [2]
    kt::AxisCtl *ctlHost = nullptr; // pinned mirror
    double *pinned = nullptr;       // 64 pinned doubles for small host <-> device scalars
    kt::DevBuf<double> history;     // [2][maxIter]
    uint64_t historyCap = 0;

    // look-ahead legalization: geometry (host), grid, per-cell scratch (device)
    bool haveDesign = false;
    std::vector<kt::lal::RowSpec> lalRows;
    kt::lal::Box lalDie;
    std::vector<kt::lal::Box> obstacles;
    kt::lal::Grid grid;
    kt::lal::FreeSpace freeSpace;
    double densityLimit = 1.0;
    kt::LalOptions lalOpts;
    uint64_t lalBatches = 0;
    uint64_t lalFusedBatches = 0;                      // of those, run in one launch each
    uint32_t lalFusedCells = 0;                        // batches of at most this many cells run fused (0: none)
    kt::DevBuf<double> gridAvail, occ, lalScratch, lalSpan, lalCap, lalCut;
    kt::DevBuf<double> hpwlScratch;  // per-net half-perimeters and chunk sums of getHpwl
    // device-resident bookkeeping of the outer loop (HPWL, overflow, Convergence)
    kt::DevBuf<double> loopDev;
    double *loopHost = nullptr;      // pinned mirror of loopDev
    kt::DevBuf<uint64_t> occFix, lalKeys, lalKeysTmp, lalKey2, lalPrefix, lalStripeLo, lalStripeHi;
    kt::DevBuf<uint32_t> binOf, binRegion, lalSeg, lalFlag, lalIds, lalVals, lalValsTmp, lalHalf, lalStripeOf,
        lalBegin, lalEnd, lalStripeStart, lalStripeEnd, lalRemap, sortCounts, sortRaw;
    kt::DevBuf<uint32_t> lalList, lalRects;  // batch cells, region rectangles
    kt::DevBuf<uint32_t> lalRegionOf, lalSaveBin, lalBinStart;  // in-order rule: per-cell restore state
    kt::DevBuf<double> lalSaveX, lalSaveY;
    uint64_t lalReruns = 0;                            // region runs repeated for the in-order rule
    kt::DevBuf<unsigned long long> lalMoved;           // moved-cell records (2 words each)
    kt::DevBuf<uint32_t> binCells;                      // cells per bin (with occFix)
    kt::lal::AreaTables lalTables;                     // free-area table of the grid, occupancy rows
    uint8_t *lalPinned = nullptr;                      // pinned staging of per-level tables
    size_t lalPinnedCap = 0;
    uint8_t *lalLand = nullptr;                        // pinned landing of a batch's small transfers
    size_t lalLandCap = 0;
    kt::DevBuf<double> lalUp;                          // device copy of the staged tables
    kt::DevBuf<kt::LalSegStat> lalStats;
    kt::DevBuf<uint8_t> lalSplit;
    kt::DevBuf<uint8_t> lalFusedScratch, lalFusedRecord;  // fused one-launch batches: tables, packed record
    kt::DevBuf<unsigned long long> lalFusedOut;          // its status and moved cells

    // learned warm start: BF16 weights / FP32 biases, features, prediction
    bool haveWarm = false;
    double lastAlpha = 0.0;              // alpha of the last B2B build
    kt::DevBuf<uint16_t> warmW, warmFeat;
    kt::DevBuf<float> warmB, warmOut, warmShare;
    kt::DevBuf<double> warmPart, warmScalars;

    kt::DevBuf<uint32_t> starScan, starNet;  // star model: scan of star flags, star -> net

    kt::DevBuf<double> exact;            // reference-order solver vectors
    std::vector<double> exactHistory;    // its residuals per iteration

    kt::Tune tune;
    kt::GraphCache graphs[2][3];  // FP64, FP32 iteration graphs; both axes, x only, y only

    // timing
    bool profile = false;
    kt::KernelTimer timers[KT_KERNEL_COUNT];
};

namespace kt {
int ensureSystem(kt_ctx *ctx, int axis, uint64_t rows, uint64_t nnzCap);
int scanRowPtr(kt_ctx *ctx, uint32_t *rowPtr, uint64_t rows);
int systemNnz(kt_ctx *ctx, int axis, uint64_t *nnz);
int timedBegin(kt_ctx *ctx, int kernel, hipEvent_t *start);
int timedEnd(kt_ctx *ctx, int kernel, hipEvent_t start);
int timedKernelEvents(kt_ctx *ctx, int kernel, hipEvent_t *start, hipEvent_t *stop);
int collectTimers(kt_ctx *ctx);
int warmSetParams(kt_ctx *ctx, const float *params, uint64_t count);
int warmFeaturesHost(kt_ctx *ctx, float *features, double *scale);
int warmPredictHost(kt_ctx *ctx, double *dx, double *dy);
int warmStart(kt_ctx *ctx);
int buildStar(kt_ctx *ctx, uint64_t *size);
int noAllocInit(int device);
}  // namespace kt
