// Ajtai key object shared by the NTT-domain and int8-MFMA commit paths.
#pragma once

#include <stdint.h>

#include "ntt.hpp"

struct lzx_ajtai_key {
    lzx_ntt_plan *plan; // owned
    size_t k, m, n;
    uint32_t norm_bound;
    uint32_t *A_dev;    // coefficient domain [k*m][n]
    uint32_t *Ahat_dev; // NTT domain [k*m][n]
    int8_t *limbs_dev;  // int8 path: signed limb planes of the centred A (layout in ajtai_i8.hip)
    size_t limbs_bytes;
    int8_t *nm_afrag;   // NTT-domain MFMA product: limb fragments of the centred A^ (layout in ajtai_nttmm.hip)
    uint32_t *nm_corr;  // and its per-(row, coefficient) byte-offset constant
    int8_t *nm_psi;     // n = 64: the forward-transform matrix as byte-limb fragments (digit witnesses)
    int device;
};

namespace lzx {

// NTT-domain commit of W witnesses already in canonical u32 form.
lzx_status ajtai_commit_ntt(const lzx_ajtai_key *key, const uint32_t *s_dev, const int8_t *digits_dev, size_t W,
                            uint32_t *out_dev, hipStream_t stream);

// int8 MFMA coefficient-domain commit of small-digit witnesses.
lzx_status ajtai_commit_int8(const lzx_ajtai_key *key, const int8_t *digits_dev, size_t W, uint32_t *out_dev,
                             hipStream_t stream);
bool ajtai_int8_supported(const lzx_ajtai_key *key);
lzx_status ajtai_build_limbs(lzx_ajtai_key *key);

// NTT-domain commit with the multiply-accumulate on int8 MFMA (q < 2^32 - 2^25,
// k <= 16, 16 <= n <= 64); ajtai_commit_ntt takes it when the key holds its
// fragments, unless LZX_AJTAI_MAC=valu.
bool ajtai_nttmm_supported(const lzx_ajtai_key *key);
bool ajtai_nttmm_enabled(const lzx_ajtai_key *key);
lzx_status ajtai_build_nttmm(lzx_ajtai_key *key);
lzx_status ajtai_commit_nttmm(const lzx_ajtai_key *key, const uint32_t *s_dev, const int8_t *digits_dev, size_t W,
                              uint32_t *out_dev, hipStream_t stream);

} // namespace lzx
