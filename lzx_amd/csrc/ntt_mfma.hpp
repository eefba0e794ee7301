// The int8-MFMA batched NTT (degree 64, 2^20 < q < 2^31 - 2^16): table layout and
// entry points used by the NTT dispatcher.
#pragma once

#include <stdint.h>

#include <vector>

#include "ntt.hpp"

namespace lzx {

// Per direction: the limb matrices (2 groups x 4 limbs x 2 row blocks x 2 K
// blocks x 64 lanes x 16 bytes), then MF_AUX bytes of u32 constants:
//   [0] q  [1] q^-1 mod 2^32  [2] the L = 2 accumulator bias  [3] the L = 0 bias
//   [MF_B0 + 32 group + row] the L = 0 accumulator bias (the same for every row)
//   [MF_K + 32 group + row]  the row's offset K in [q, 2q - 2^16)
constexpr uint32_t MF_MAT = 32768, MF_AUX = 1024, MF_DIR = MF_MAT + MF_AUX;
constexpr uint32_t MF_B0 = 16, MF_K = 80;
constexpr uint32_t MF_TP = 32, MF_WAVES = 8; // polynomials per tile, waves per block
constexpr uint32_t MF_Q_LIMIT = (1u << 31) - (1u << 16), MF_Q_MIN = 1u << 20;

struct MfSet {
    const uint8_t *p[MAX_PLANS_PER_LAUNCH];
};

bool ntt_mfma_supported(const lzx_ntt_plan *p);
std::vector<uint8_t> ntt_mfma_tables(const lzx_ntt_plan *p);
void ntt_mfma_launch(const lzx_ntt_plan *const *plans, size_t nplans, uint32_t *data, uint32_t polys, NttMode mode,
                     hipStream_t stream);

} // namespace lzx
