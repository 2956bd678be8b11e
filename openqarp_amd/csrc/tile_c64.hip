// tile_c64.hip — tile-pass kernels for complex64 states (tile arrays of
// 2^11 .. 2^13 amplitudes: 16 / 32 / 64 KiB per state).
#include "tile_kernel.h"

namespace qhip {

void launch_tile_pass(const PassDesc& P, float2* psi, float2* phi, double2* ip_part, const DevOp* d_ops,
                      const float2* d_coefs, int64_t n_tiles, hipStream_t stream) {
    tile::launch<float2, kTileArrayQubitsC64>(P, psi, phi, ip_part, d_ops, d_coefs, n_tiles, stream);
}

void launch_tile_batch_c64(bool two, const BatchItem* items, const int64_t* first, int n_items,
                          int64_t total, int tmax, int kmax, hipStream_t stream) {
    tile::launch_batch<float2, kTileArrayQubitsC64>(two, items, first, n_items, total, tmax, kmax, stream);
}

}  // namespace qhip
