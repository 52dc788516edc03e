// Large inclusive prefix sums (the B2B assembly's reference counts, LAL
// prefixes) by rocPRIM's single-pass device scan.  Integer addition is
// associative, so the values do not depend on how the library splits the
// work.  Its temporary storage is kept in the context between calls.
#include <rocprim/device/device_scan.hpp>

#include "kt_internal.h"

namespace kt {
namespace {

template <typename T>
int deviceScan(kt_ctx *ctx, T *data, uint64_t m) {
    if (m == 0) return KT_OK;
    size_t bytes = 0;
    KT_HIP(rocprim::inclusive_scan(nullptr, bytes, data, data, static_cast<size_t>(m), rocprim::plus<T>(),
                                   ctx->stream));
    KT_TRY(ctx->scanStorage.reserve(std::max<size_t>(bytes, 1)));
    KT_HIP(rocprim::inclusive_scan(ctx->scanStorage.p, bytes, data, data, static_cast<size_t>(m), rocprim::plus<T>(),
                                   ctx->stream));
    KT_HIP(hipGetLastError());
    return KT_OK;
}

}  // namespace

int deviceScan32(kt_ctx *ctx, uint32_t *data, uint64_t m) { return deviceScan<uint32_t>(ctx, data, m); }
int deviceScan64(kt_ctx *ctx, uint64_t *data, uint64_t m) { return deviceScan<uint64_t>(ctx, data, m); }

}  // namespace kt
