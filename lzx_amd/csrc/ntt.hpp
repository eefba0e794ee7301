// Negacyclic NTT plans and launch helpers shared by the NTT, ring and Ajtai
// modules.
#pragma once

#include <stdint.h>

#include <vector>

#include "modarith.hpp"
#include "runtime.hpp"

namespace lzx {

// Layout of one plan's constant table in HBM (u32 words).  The forward and
// inverse kernels use twiddles with psi folded in (Cooley-Tukey from natural
// order, Gentleman-Sande back), and permute to the reference's natural-order
// evaluations on store.  The ref* tables restate NttTables' pow_psi and
// pow_omega (ntt.rs:60-79) for the incomplete transforms, which must follow
// the reference's own level structure.
//   [0] q  [1] qinv  [2] n  [3] log_n  [4] bigq  [5] c  [6] k (q = c 2^k + 1)
//   [7] arithmetic class  [12] 1 if q < 2^30 (Harvey's lazy range fits 32 bits)
//   [8] n^{-1} (w, w')   [10] z_1^{-1} n^{-1} (w, w')
//   fwd zeta  z_k = psi^brv(k), k in [1, n):   FWD_W + k, FWD_WP + k
//   inv zeta  z_k^{-1}:                         INV_W + k, INV_WP + k
//   pow_psi[i], i < n:                          PSI_W + i, PSI_WP + i
//   pow_omega[e], e < max(n/2, 1):              OMG_W + e, OMG_WP + e
// (w, w') is the Shoup pair for q < 2^31 and (w*2^32 mod q, 0) otherwise.
struct TabLayout {
    static constexpr uint32_t HDR = 16;
    static __host__ __device__ constexpr uint32_t fwd_w(uint32_t n) { return HDR; }
    static __host__ __device__ constexpr uint32_t fwd_wp(uint32_t n) { return HDR + n; }
    static __host__ __device__ constexpr uint32_t inv_w(uint32_t n) { return HDR + 2 * n; }
    static __host__ __device__ constexpr uint32_t inv_wp(uint32_t n) { return HDR + 3 * n; }
    static __host__ __device__ constexpr uint32_t psi_w(uint32_t n) { return HDR + 4 * n; }
    static __host__ __device__ constexpr uint32_t psi_wp(uint32_t n) { return HDR + 5 * n; }
    static __host__ __device__ constexpr uint32_t omg_w(uint32_t n) { return HDR + 6 * n; }
    static __host__ __device__ constexpr uint32_t omg_wp(uint32_t n) { return HDR + 6 * n + (n > 1 ? n / 2 : 1); }
    // FP64 constants (8-byte aligned; offsets in doubles from dbl(n)):
    // [q, 1/q, n^-1, fl(n^-1/q), last-level w, fl(w/q), 0, 0], then a
    // (w, fl(w/q)) pair per forward twiddle k < n, then one per inverse twiddle
    static __host__ __device__ constexpr uint32_t dbl(uint32_t n) { return (HDR + 7 * n + 2 + 1) & ~1u; }
    static __host__ __device__ constexpr uint32_t dbl_fwd(uint32_t) { return 8; }
    static __host__ __device__ constexpr uint32_t dbl_inv(uint32_t n) { return 8 + 2 * n; }
    static __host__ __device__ constexpr uint32_t words(uint32_t n) { return dbl(n) + 2 * (8 + 4 * n); }
};

constexpr uint32_t NTT_MAX_LOG_N = 13; // LDS tile limit of the generic kernel
constexpr int MAX_PLANS_PER_LAUNCH = 8;

struct TabSet {
    const uint32_t *p[MAX_PLANS_PER_LAUNCH];
};

} // namespace lzx

struct lzx_ntt_plan {
    uint32_t q, two_adicity, generator, log_n, n;
    uint32_t psi, psi_inv, n_inv;
    bool bigq;
    int arith;         // lzx::Arith class of q
    uint32_t prime_c;  // q = prime_c * 2^prime_k + 1
    uint32_t prime_k;
    bool stage_dma;       // persistent DMA-staged register kernel (else LDS-staged blocks)
    int cus;              // compute units of the plan's device (the persistent kernel's grid scale)
    lzx::Mod32 mod;
    std::vector<uint32_t> host_tab; // same words as dev_tab
    uint32_t *dev_tab;
    uint8_t *mfma_dev;    // int8-MFMA limb matrices (n = 64, 2^20 < q < 2^31 - 2^16), else null
    uint32_t kernel_pref; // LZX_NTT_PATH_AUTO, or the class an MFMA-capable launch must run
    int device;
};

namespace lzx {

enum NttMode { NTT_FWD = 0, NTT_INV = 1, NTT_PARTIAL = 2, NTT_COMPLETE = 3 };

// Launch one transform mode over nplans plans sharing n and bigq; data is
// laid out [plan][poly][n] with `polys` polynomials per plan.
// Forward NTT of `polys` small-digit polynomials (i8, any value, 16-byte
// aligned) into u32 NTT-domain output; 16 <= n <= 64, else LZX_UNSUPPORTED.
lzx_status ntt_forward_digits(const lzx_ntt_plan *plan, const int8_t *digits, uint32_t *out, size_t polys,
                              hipStream_t stream);
lzx_status ntt_launch(const lzx_ntt_plan *const *plans, size_t nplans, uint32_t *data, size_t polys, NttMode mode,
                      uint32_t levels, hipStream_t stream);
// The kernel class a batched forward/inverse of these plans runs (or
// LZX_NTT_PATH_AUTO while an MFMA-capable shape has not been measured).
lzx_status ntt_batch_path(const lzx_ntt_plan *const *plans, size_t nplans, size_t polys, NttMode mode,
                          uint32_t &path);

} // namespace lzx
