// Hybrid scheduling (placeholder until the tile-binning kernels land).
#include "sgs_internal.h"

namespace sgs {

size_t hybrid_st_scratch_bytes(uint32_t, uint32_t, const Tuning&) { return 256; }
size_t hybrid_gps_scratch_bytes(const Tuning&) { return 256; }

hipError_t run_hybrid_st(const SceneArgs&, const float*, const FbCamera&, uint64_t*, uint32_t, const uint8_t*, uint32_t,
                         const Tuning&, void*, uint64_t*, hipStream_t)
{
    return hipErrorNotSupported;
}

hipError_t run_hybrid_gps(const SceneArgs&, const float*, const FbCamera&, uint64_t*, uint32_t, uint32_t, const Tuning&,
                          void*, uint64_t*, hipStream_t)
{
    return hipErrorNotSupported;
}

}  // namespace sgs
