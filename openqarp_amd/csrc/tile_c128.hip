// tile_c128.hip — tile-pass kernels for complex128 states (tile arrays of
// 2^10 .. 2^12 amplitudes: 16 / 32 / 64 KiB per state).
#include "tile_kernel.h"

namespace qhip {

void launch_tile_pass(const PassDesc& P, double2* psi, double2* phi, double2* ip_part, const DevOp* d_ops,
                      const double2* d_coefs, int64_t n_tiles, hipStream_t stream) {
    tile::launch<double2, kTileArrayQubitsC128>(P, psi, phi, ip_part, d_ops, d_coefs, n_tiles, stream);
}

void launch_tile_batch_c128(bool two, const BatchItem* items, const int64_t* first, int n_items,
                          int64_t total, int tmax, int kmax, hipStream_t stream) {
    tile::launch_batch<double2, kTileArrayQubitsC128>(two, items, first, n_items, total, tmax, kmax, stream);
}

}  // namespace qhip
