// Uneven commitment batch: many independent commitments t = A s, each on its
// own ring Z_q[X]/(X^n + 1) (n in {64, 128, 256}, q < 2^50) and with its own
// length m, processed by ONE persistent launch.
//
// Each job has the semantics of AjtaiPublicKey::commit
// (crates/lattice-commitment/src/ajtai.rs:118-153): t_i = sum_j A_ij s_j with
// the NTT-domain key cached (ajtai.rs:76-85 stores a_ntt), the witness
// transformed, multiplied pointwise, accumulated and transformed back.  The
// reference commits one witness per call with one modulus class (Modulus32);
// here jobs of every class share one launch.
//
// Work decomposition (device work queue, one atomic head counter):
//   phase A item (job, chunk): `chunk` witness elements -> LDS, forward NTT,
//       multiply-accumulate against the key rows (all k rows), one partial
//       [k][n] of canonical residues to scratch; then done[job] += 1;
//   phase B item (job, row):   waits for done[job] == items(job), sums the
//       row's partials mod q, inverse NTT, writes t_row.
// All phase A items precede all phase B items in the queue, so a phase B item
// only ever waits for items that running blocks already hold: the wait always
// ends (and is bounded anyway: past the bound an error flag is raised).
// Items are ordered largest job first.  The launch (with the counter reset,
// a plain fill kernel) is captured into a hipGraph and replayed.
//
// Arithmetic by modulus class ("kind"):
//   kind 0, q < 2^31:  u32 residues, Shoup twiddles, products summed in u64
//                      with an overflow count;
//   kind 1, q < 2^32:  u32 residues, Montgomery twiddles, same sums;
//   kind 2, q < 2^50:  u64 residues, 64-bit Shoup twiddles; products split
//                      into 25-bit limbs, a*b = a1b1 2^50 + (a1b0 + a0b1) 2^25
//                      + a0b0, three u64 sums exact for 4096 terms, folded
//                      mod q with Shoup constants at the end.
// Inside a job the transforms use the merged-psi Cooley-Tukey forward
// (bit-reversed evaluation order) and Gentleman-Sande inverse, so the key's
// NTT copy is in that order; only t leaves the library, and t is the exact
// ring product.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "modarith.hpp"
#include "runtime.hpp"

namespace lzx {

enum : uint32_t { WK_S32 = 0, WK_M32 = 1, WK_W64 = 2 };

constexpr uint32_t CB_THREADS = 256;
constexpr uint32_t CB_DATA_BYTES = 32768; // LDS for one sub-batch of witness elements
constexpr uint32_t CB_MIN_LOG_N = 6, CB_MAX_LOG_N = 8;
constexpr uint32_t CB_CLASSES = 9;
constexpr uint32_t CB_MAX_CHUNK = 4096;   // 25-bit limb sums exact up to 4096 products
constexpr uint32_t CB_MAX_ITEMS = 8192;   // per job: phase B sums stay below 2^63
constexpr uint32_t CB_ROWS = 16;          // key rows per pass of a work item
constexpr uint32_t CB_INFLIGHT = 8;       // 16-byte key loads a thread issues before consuming them
constexpr uint32_t CB_SPIN_LIMIT = 1u << 26;

struct KeyDev {
    const void *ahat;   // [k][m][n] NTT domain (bit-reversed evaluation order)
    const uint64_t *tw; // [4n]: forward w, w'; inverse w, w'
    uint64_t q, m;
    uint32_t k, log_n, kind, qinv;
    uint64_t barrett;                   // floor((2^64-1)/q), 32-bit kinds
    uint64_t r64;                       // 2^64 mod q, 32-bit kinds
    uint64_t one_p, c25, c25p, c50, c50p; // Shoup pairs of 1, 2^25, 2^50 (kind 2)
    uint64_t ninv, ninvp;               // n^{-1} as a twiddle pair
};

struct JobDev {
    uint32_t key, nitems, chunk, pad;
    uint64_t m;
    const int8_t *digits;
    void *out;
    uint64_t part; // byte offset of [nitems][k][n] partials
};

// ---- global-memory access ----------------------------------------------------
// Key, witness, partial and output pointers reach the kernels through
// descriptors in memory, so the compiler cannot infer their address space and
// would emit flat accesses, which also count against the LDS wait counter and
// so serialize with the LDS traffic of the same loop.  Every global access
// goes through an explicit address_space(1) pointer instead.
#define CB_GLOBAL __attribute__((address_space(1)))
typedef unsigned cb_v4u __attribute__((ext_vector_type(4)));

template <typename T> __device__ __forceinline__ CB_GLOBAL T *gbl(T *p) { return (CB_GLOBAL T *)p; }
template <typename T> __device__ __forceinline__ const CB_GLOBAL T *gbl(const T *p)
{
    return (const CB_GLOBAL T *)p;
}

// 16-byte load of streamed-once data (the key): nontemporal, kept out of the caches' reuse set.
__device__ __forceinline__ uint4 load16_stream(const void *p)
{
    const cb_v4u x = __builtin_nontemporal_load(gbl(reinterpret_cast<const cb_v4u *>(p)));
    return make_uint4(x[0], x[1], x[2], x[3]);
}

__device__ __forceinline__ uint4 load16(const void *p)
{
    const cb_v4u x = *gbl(reinterpret_cast<const cb_v4u *>(p));
    return make_uint4(x[0], x[1], x[2], x[3]);
}

// ---- arithmetic per kind ----------------------------------------------------

LZX_HD uint64_t shoup64(uint64_t a, uint64_t w, uint64_t wp, uint64_t q)
{
    const uint64_t qh = mulhi64(a, wp);
    const uint64_t r = a * w - qh * q; // in [0, 2q) for any a < 2^64, q < 2^63
    return r >= q ? r - q : r;
}

template <int KIND> struct WA;

template <> struct WA<WK_S32> {
    typedef uint32_t T;
    struct Acc {
        uint64_t lo;
        uint32_t c;
    };
    static __device__ __forceinline__ T tw(T a, uint64_t w, uint64_t wp, const KeyDev &K)
    {
        return shoup_mul(a, (uint32_t)w, (uint32_t)wp, (uint32_t)K.q);
    }
    static __device__ __forceinline__ T add(T a, T b, const KeyDev &K) { return add_small(a, b, (uint32_t)K.q); }
    static __device__ __forceinline__ T sub(T a, T b, const KeyDev &K) { return sub_small(a, b, (uint32_t)K.q); }
    static __device__ __forceinline__ void zero(Acc &s) { s.lo = 0, s.c = 0; }
    static __device__ __forceinline__ void mac(Acc &s, T a, T b)
    {
        const uint64_t p = (uint64_t)a * b;
        s.lo += p;
        s.c += s.lo < p;
    }
    static __device__ __forceinline__ T finish(const Acc &s, const KeyDev &K)
    {
        const uint32_t q = (uint32_t)K.q;
        return add_big(reduce_u64(s.lo, q, K.barrett), reduce_u64((uint64_t)s.c * K.r64, q, K.barrett), q);
    }
    static __device__ __forceinline__ T reduce_any(T x, const KeyDev &K)
    {
        return reduce_u64(x, (uint32_t)K.q, K.barrett);
    }
    static __device__ __forceinline__ T reduce_sum(uint64_t x, const KeyDev &K)
    {
        return reduce_u64(x, (uint32_t)K.q, K.barrett);
    }
};

template <> struct WA<WK_M32> : WA<WK_S32> {
    static __device__ __forceinline__ T tw(T a, uint64_t w, uint64_t wp, const KeyDev &K)
    {
        return mont_mul(a, (uint32_t)w, (uint32_t)K.q, K.qinv);
    }
    static __device__ __forceinline__ T add(T a, T b, const KeyDev &K) { return add_big(a, b, (uint32_t)K.q); }
    static __device__ __forceinline__ T sub(T a, T b, const KeyDev &K) { return sub_big(a, b, (uint32_t)K.q); }
};

template <> struct WA<WK_W64> {
    typedef uint64_t T;
    struct Acc {
        uint64_t s0, s1, s2;
    };
    static constexpr uint32_t M25 = (1u << 25) - 1;
    static __device__ __forceinline__ T tw(T a, uint64_t w, uint64_t wp, const KeyDev &K)
    {
        return shoup64(a, w, wp, K.q);
    }
    static __device__ __forceinline__ T add(T a, T b, const KeyDev &K)
    {
        const uint64_t s = a + b;
        return s >= K.q ? s - K.q : s;
    }
    static __device__ __forceinline__ T sub(T a, T b, const KeyDev &K) { return a >= b ? a - b : a + (K.q - b); }
    static __device__ __forceinline__ void zero(Acc &s) { s.s0 = 0, s.s1 = 0, s.s2 = 0; }
    static __device__ __forceinline__ void mac(Acc &s, T a, T b)
    {
        const uint32_t a0 = (uint32_t)a & M25, a1 = (uint32_t)(a >> 25);
        const uint32_t b0 = (uint32_t)b & M25, b1 = (uint32_t)(b >> 25);
        s.s0 += (uint64_t)a0 * b0;
        s.s1 += (uint64_t)a0 * b1;
        s.s1 += (uint64_t)a1 * b0;
        s.s2 += (uint64_t)a1 * b1;
    }
    static __device__ __forceinline__ T finish(const Acc &s, const KeyDev &K)
    {
        const T x0 = shoup64(s.s0, 1, K.one_p, K.q);
        const T x1 = shoup64(s.s1, K.c25, K.c25p, K.q);
        const T x2 = shoup64(s.s2, K.c50, K.c50p, K.q);
        return add(add(x0, x1, K), x2, K);
    }
    static __device__ __forceinline__ T reduce_any(T x, const KeyDev &K) { return shoup64(x, 1, K.one_p, K.q); }
    static __device__ __forceinline__ T reduce_sum(uint64_t x, const KeyDev &K) { return shoup64(x, 1, K.one_p, K.q); }
};

// ---- LDS transforms ------------------------------------------------------------

template <int LOG_N>
__device__ __forceinline__ void load_twiddles(const KeyDev &K, uint64_t *tws)
{
    constexpr uint32_t N = 1u << LOG_N;
    for (uint32_t i = threadIdx.x; i < 4 * N; i += CB_THREADS) tws[i] = gbl(K.tw)[i];
}

// Merged-psi Cooley-Tukey, natural coefficients -> bit-reversed evaluations,
// over nb polynomials of xs.  Ends with a barrier.
template <int KIND, int LOG_N>
__device__ void forward_ntt(const KeyDev &K, typename WA<KIND>::T *xs, uint32_t nb, const uint64_t *tws)
{
    using A = WA<KIND>;
    using T = typename A::T;
    constexpr uint32_t N = 1u << LOG_N, H = N / 2;
#pragma unroll 1
    for (int lg = LOG_N - 1; lg >= 0; lg--) {
        const uint32_t len = 1u << lg, grp = 1u << (LOG_N - 1 - lg);
        for (uint32_t b = threadIdx.x; b < nb * H; b += CB_THREADS) {
            const uint32_t e = b >> (LOG_N - 1), r = b & (H - 1);
            const uint32_t g = r >> lg, o = r & (len - 1);
            const uint32_t i = e * N + (g << (lg + 1)) + o, j = i + len;
            const T u = xs[i];
            const T v = A::tw(xs[j], tws[grp + g], tws[N + grp + g], K);
            xs[i] = A::add(u, v, K);
            xs[j] = A::sub(u, v, K);
        }
        __syncthreads();
    }
}

// Gentleman-Sande inverse of one polynomial (bit-reversed evaluations ->
// natural coefficients), scaled by n^{-1}.  Ends with a barrier.
template <int KIND, int LOG_N>
__device__ void inverse_ntt(const KeyDev &K, typename WA<KIND>::T *x, const uint64_t *tws)
{
    using A = WA<KIND>;
    using T = typename A::T;
    constexpr uint32_t N = 1u << LOG_N, H = N / 2;
#pragma unroll 1
    for (int lg = 0; lg < LOG_N; lg++) {
        const uint32_t len = 1u << lg, grp = H >> lg;
        for (uint32_t b = threadIdx.x; b < H; b += CB_THREADS) {
            const uint32_t g = b >> lg, o = b & (len - 1);
            const uint32_t i = (g << (lg + 1)) + o, j = i + len;
            const T u = x[i], v = x[j];
            x[i] = A::add(u, v, K);
            x[j] = A::tw(A::sub(u, v, K), tws[2 * N + grp + g], tws[3 * N + grp + g], K);
        }
        __syncthreads();
    }
    for (uint32_t c = threadIdx.x; c < N; c += CB_THREADS) x[c] = A::tw(x[c], K.ninv, K.ninvp, K);
    __syncthreads();
}

// nb witness elements (nb * N int8 digits, 16-byte aligned) -> residues in xs.
template <int KIND, int LOG_N>
__device__ __forceinline__ void load_witness(const KeyDev &K, const int8_t *digits, uint32_t nb,
                                             typename WA<KIND>::T *xs)
{
    using T = typename WA<KIND>::T;
    constexpr uint32_t N = 1u << LOG_N;
    const uint32_t vecs = nb * N / 16;
    for (uint32_t v = threadIdx.x; v < vecs; v += CB_THREADS) {
        const uint4 d = load16(digits + 16 * (size_t)v);
        const uint32_t w[4] = {d.x, d.y, d.z, d.w};
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int s = (int8_t)(w[i >> 2] >> (8 * (i & 3)));
            xs[16 * v + i] = s < 0 ? (T)(K.q - (uint64_t)(-s)) : (T)s;
        }
    }
}

template <typename T> struct Vec16;
template <> struct Vec16<uint32_t> {
    static __device__ __forceinline__ void unpack(const uint4 &x, uint32_t (&v)[4])
    {
        v[0] = x.x, v[1] = x.y, v[2] = x.z, v[3] = x.w;
    }
};
template <> struct Vec16<uint64_t> {
    static __device__ __forceinline__ void unpack(const uint4 &x, uint64_t (&v)[2])
    {
        v[0] = x.x | ((uint64_t)x.y << 32), v[1] = x.z | ((uint64_t)x.w << 32);
    }
};

// ---- work items -------------------------------------------------------------------

// Phase A: ring elements [item * chunk, min(m, (item + 1) * chunk)) of one job.
template <int KIND, int LOG_N>
__device__ void mac_item(const KeyDev &K, const JobDev &J, uint32_t item, unsigned char *smem, uint64_t *tws,
                         unsigned char *part_base)
{
    using A = WA<KIND>;
    using T = typename A::T;
    using Acc = typename A::Acc;
    constexpr uint32_t N = 1u << LOG_N;
    constexpr uint32_t VEC = 16 / sizeof(T);        // coefficients per thread (one 16-byte load)
    constexpr uint32_t TPR = N / VEC;               // threads per key row
    constexpr uint32_t RS = CB_THREADS / TPR;       // row stride between a thread's rows
    constexpr uint32_t RP = RS < CB_ROWS ? CB_ROWS / RS : 1; // rows per thread per pass
    constexpr uint32_t U = CB_INFLIGHT / RP;        // ring elements whose key loads are issued together
    constexpr uint32_t SB = CB_DATA_BYTES / (N * sizeof(T)); // witness elements per LDS sub-batch
    T *xs = reinterpret_cast<T *>(smem);
    const uint32_t tid = threadIdx.x;
    const uint32_t cv = (tid % TPR) * VEC, r0 = tid / TPR;
    const uint64_t jb = (uint64_t)item * J.chunk, je = min(J.m, jb + J.chunk);
    const T *ahat = reinterpret_cast<const T *>(K.ahat);
    T *part = reinterpret_cast<T *>(part_base + J.part) + (size_t)item * K.k * N;

    load_twiddles<LOG_N>(K, tws);
    for (uint32_t rbase = 0; rbase < K.k; rbase += RP * RS) {
        Acc acc[RP][VEC];
#pragma unroll
        for (uint32_t r = 0; r < RP; r++)
#pragma unroll
            for (uint32_t v = 0; v < VEC; v++) A::zero(acc[r][v]);
        for (uint64_t j0 = jb; j0 < je; j0 += SB) {
            const uint32_t nb = (uint32_t)min<uint64_t>(SB, je - j0);
            __syncthreads(); // the previous sub-batch (and the twiddle fill) is done with LDS
            load_witness<KIND, LOG_N>(K, J.digits + j0 * N, nb, xs);
            __syncthreads();
            forward_ntt<KIND, LOG_N>(K, xs, nb, tws);
            const T *arow = ahat + ((size_t)(rbase + r0) * K.m + j0) * N + cv;
            const size_t rstep = (size_t)RS * K.m * N;
            // U ring elements x RP rows of key loads in flight, then their MACs
            for (uint32_t jj = 0; jj < nb; jj += U) {
                uint4 raw[U][RP];
#pragma unroll
                for (uint32_t u = 0; u < U; u++)
#pragma unroll
                    for (uint32_t r = 0; r < RP; r++)
                        raw[u][r] = (jj + u < nb && rbase + r0 + RS * r < K.k)
                                        ? load16_stream(arow + r * rstep + (size_t)(jj + u) * N)
                                        : make_uint4(0, 0, 0, 0);
#pragma unroll
                for (uint32_t u = 0; u < U; u++) {
                    if (jj + u < nb) {
                        T b[VEC];
#pragma unroll
                        for (uint32_t v = 0; v < VEC; v++) b[v] = xs[(jj + u) * N + cv + v];
#pragma unroll
                        for (uint32_t r = 0; r < RP; r++) {
                            T a[VEC];
                            Vec16<T>::unpack(raw[u][r], a);
#pragma unroll
                            for (uint32_t v = 0; v < VEC; v++) A::mac(acc[r][v], a[v], b[v]);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < RP; r++) {
            const uint32_t row = rbase + r0 + RS * r;
            if (row < K.k) {
#pragma unroll
                for (uint32_t v = 0; v < VEC; v++) gbl(part)[(size_t)row * N + cv + v] = A::finish(acc[r][v], K);
            }
        }
    }
}

// Phase B: row `row` of one job, after all of the job's phase A items.
template <int KIND, int LOG_N>
__device__ void finish_item(const KeyDev &K, const JobDev &J, uint32_t row, unsigned char *smem, uint64_t *tws,
                            const unsigned char *part_base)
{
    using A = WA<KIND>;
    using T = typename A::T;
    constexpr uint32_t N = 1u << LOG_N, G = CB_THREADS / N;
    uint64_t *red = reinterpret_cast<uint64_t *>(smem + CB_DATA_BYTES / 2);
    T *x = reinterpret_cast<T *>(smem);
    const uint32_t c = threadIdx.x % N, g = threadIdx.x / N;
    const T *part = reinterpret_cast<const T *>(part_base + J.part);
    load_twiddles<LOG_N>(K, tws);
    uint64_t s = 0; // < items * q <= 2^13 * 2^50
    for (uint32_t it = g; it < J.nitems; it += G) s += gbl(part)[((size_t)it * K.k + row) * N + c];
    red[threadIdx.x] = s;
    __syncthreads();
    if (g == 0) {
        uint64_t t = 0;
#pragma unroll
        for (uint32_t i = 0; i < G; i++) t += red[i * N + c];
        x[c] = A::reduce_sum(t, K);
    }
    __syncthreads();
    inverse_ntt<KIND, LOG_N>(K, x, tws);
    T *out = reinterpret_cast<T *>(J.out) + (size_t)row * N;
    for (uint32_t i = threadIdx.x; i < N; i += CB_THREADS) gbl(out)[i] = x[i];
}

template <int KIND, int LOG_N>
__global__ __launch_bounds__(CB_THREADS) void wide_key_ntt_kernel(KeyDev K, const void *A_in, size_t npolys)
{
    using A = WA<KIND>;
    using T = typename A::T;
    constexpr uint32_t N = 1u << LOG_N;
    constexpr uint32_t SB = CB_DATA_BYTES / (N * sizeof(T));
    __shared__ __attribute__((aligned(16))) unsigned char smem[CB_DATA_BYTES];
    __shared__ uint64_t tws[4 * N];
    T *xs = reinterpret_cast<T *>(smem);
    const T *src = reinterpret_cast<const T *>(A_in);
    T *dst = const_cast<T *>(reinterpret_cast<const T *>(K.ahat));
    load_twiddles<LOG_N>(K, tws);
    for (size_t base = (size_t)blockIdx.x * SB; base < npolys; base += (size_t)gridDim.x * SB) {
        const uint32_t nb = (uint32_t)min<size_t>(SB, npolys - base);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nb * N; i += CB_THREADS) xs[i] = A::reduce_any(gbl(src)[base * N + i], K);
        __syncthreads();
        forward_ntt<KIND, LOG_N>(K, xs, nb, tws);
        for (uint32_t i = threadIdx.x; i < nb * N; i += CB_THREADS) gbl(dst)[base * N + i] = xs[i];
    }
}

// state: [0] queue head, [1] error flag, [2 + job] finished phase A items
__global__ void batch_reset_kernel(uint32_t *state, uint32_t words)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < words; i += gridDim.x * blockDim.x) state[i] = 0;
}

#define CB_DISPATCH(FN, ...)                                                                                         \
    switch (K.kind * 3 + K.log_n - CB_MIN_LOG_N) {                                                                   \
    case 0: FN<WK_S32, 6>(__VA_ARGS__); break;                                                                        \
    case 1: FN<WK_S32, 7>(__VA_ARGS__); break;                                                                        \
    case 2: FN<WK_S32, 8>(__VA_ARGS__); break;                                                                        \
    case 3: FN<WK_M32, 6>(__VA_ARGS__); break;                                                                        \
    case 4: FN<WK_M32, 7>(__VA_ARGS__); break;                                                                        \
    case 5: FN<WK_M32, 8>(__VA_ARGS__); break;                                                                        \
    case 6: FN<WK_W64, 6>(__VA_ARGS__); break;                                                                        \
    case 7: FN<WK_W64, 7>(__VA_ARGS__); break;                                                                        \
    default: FN<WK_W64, 8>(__VA_ARGS__); break;                                                                       \
    }

__global__ __launch_bounds__(CB_THREADS) __attribute__((amdgpu_waves_per_eu(2, 2))) void commit_batch_kernel(const KeyDev *__restrict__ keys,
                                                                 const JobDev *__restrict__ jobs,
                                                                 const uint2 *__restrict__ items, uint32_t n_a,
                                                                 uint32_t n_total, uint32_t *state,
                                                                 unsigned char *part_base, uint32_t spin_limit)
{
    __shared__ __attribute__((aligned(16))) unsigned char smem[CB_DATA_BYTES];
    __shared__ uint64_t tws[4 << CB_MAX_LOG_N];
    __shared__ uint32_t s_item, s_fail;
    uint32_t *head = state, *err = state + 1, *done = state + 2;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(head, 1u);
        __syncthreads();
        const uint32_t it = s_item;
        if (it >= n_total) break;
        const uint2 w = items[it];
        const JobDev J = jobs[w.x];
        const KeyDev K = keys[J.key];
        if (it < n_a) {
            CB_DISPATCH(mac_item, K, J, w.y, smem, tws, part_base);
            // Publish the partial: each wave waits until its own stores have
            // completed (vmcnt(0) also counts stores on gfx9), then one agent-scope
            // release fence (L2 write-back) covers the whole block before the
            // count moves.  A fence per wave costs a write-back each.
            __builtin_amdgcn_s_waitcnt(0x0F70);
            __syncthreads();
            if (threadIdx.x == 0) {
                __threadfence();
                atomicAdd(&done[w.x], 1u);
            }
        } else {
            if (threadIdx.x == 0) {
                uint32_t spins = 0;
                while (__hip_atomic_load(&done[w.x], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < J.nitems) {
                    if (++spins > spin_limit) break;
                    __builtin_amdgcn_s_sleep(8);
                }
                s_fail = spins > spin_limit;
                if (s_fail) atomicOr(err, 1u);
            }
            __syncthreads();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            if (!s_fail) CB_DISPATCH(finish_item, K, J, w.y, smem, tws, part_base);
        }
        __syncthreads(); // LDS and s_item are free for the next item
    }
}

// ---- host side -------------------------------------------------------------------

namespace {

uint64_t mulmod_h(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((unsigned __int128)a * b % q); }

uint64_t powmod_h(uint64_t b, uint64_t e, uint64_t q)
{
    uint64_t r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = mulmod_h(r, b, q);
        b = mulmod_h(b, b, q);
        e >>= 1;
    }
    return r;
}

uint32_t brv(uint32_t x, uint32_t bits)
{
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

uint64_t shoup64_companion(uint64_t w, uint64_t q) { return (uint64_t)(((unsigned __int128)w << 64) / q); }

// (w, w') of one twiddle as the kind's kernels consume it.
void tw_pair(uint32_t kind, uint64_t w, uint64_t q, uint64_t &a, uint64_t &b)
{
    if (kind == WK_S32) {
        a = w, b = shoup_companion((uint32_t)w, (uint32_t)q);
    } else if (kind == WK_M32) {
        a = to_mont((uint32_t)w, (uint32_t)q), b = 0;
    } else {
        a = w, b = shoup64_companion(w, q);
    }
}

uint32_t class_of(uint32_t kind, uint32_t log_n) { return kind * 3 + log_n - CB_MIN_LOG_N; }

std::string cache_key(uint32_t kind, uint32_t log_n, size_t k)
{
    char buf[96];
    snprintf(buf, sizeof buf, "commit_batch kind=%u log_n=%u k=%zu", kind, log_n, k);
    return buf;
}

bool cache_get(const std::string &key, uint32_t &val)
{
    return autotune_get(key, val) && val >= 1 && val <= CB_MAX_CHUNK;
}

void cache_put(const std::string &key, uint32_t val) { autotune_put(key, val); }

uint32_t env_chunk()
{
    const char *e = getenv("LZX_BATCH_CHUNK");
    if (!e || !*e) return 0;
    const unsigned long v = strtoul(e, nullptr, 10);
    return (v >= 1 && v <= CB_MAX_CHUNK) ? (uint32_t)v : 0;
}

// Untuned default: 512 ring elements per work item was the best common
// choice for every class on the 512-job mixed batch (16 key rows).
uint32_t default_chunk(size_t k, uint32_t n, uint32_t elem) { return 512; }

} // namespace
} // namespace lzx

using namespace lzx;

struct lzx_wide_key {
    uint64_t q, generator;
    uint32_t log_n, n, kind, elem;
    size_t k, m;
    void *ahat;
    uint64_t *tw;
    KeyDev dev;
    int device;
};

struct lzx_commit_batch {
    std::vector<lzx_commit_job> jobs;
    std::vector<const lzx_wide_key *> keys;
    std::vector<uint32_t> job_key;
    uint32_t chunk[CB_CLASSES] = {};
    bool present[CB_CLASSES] = {};
    KeyDev *d_keys = nullptr;
    JobDev *d_jobs = nullptr;
    uint2 *d_items = nullptr;
    uint32_t *d_state = nullptr;
    unsigned char *d_part = nullptr;
    size_t part_bytes = 0, n_a = 0, n_total = 0;
    uint32_t grid = 1;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
};

namespace {

void free_plan(lzx_commit_batch *b)
{
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr, b->graph = nullptr;
    (void)hipFree(b->d_keys);
    (void)hipFree(b->d_jobs);
    (void)hipFree(b->d_items);
    (void)hipFree(b->d_state);
    (void)hipFree(b->d_part);
    b->d_keys = nullptr, b->d_jobs = nullptr, b->d_items = nullptr, b->d_state = nullptr, b->d_part = nullptr;
}

uint32_t job_class(const lzx_wide_key *k) { return class_of(k->kind, k->log_n); }

// Builds the device work queue and descriptors for the batch's current chunks.
lzx_status plan_batch(lzx_commit_batch *b)
{
    free_plan(b);
    const size_t nj = b->jobs.size();
    std::vector<JobDev> jd(nj);
    std::vector<size_t> order(nj);
    size_t off = 0;
    for (size_t i = 0; i < nj; i++) {
        const lzx_commit_job &J = b->jobs[i];
        const lzx_wide_key *K = J.key;
        uint32_t chunk = b->chunk[job_class(K)];
        // keep per-job item counts within the exact phase B sum
        while ((J.m + chunk - 1) / chunk > CB_MAX_ITEMS && chunk < CB_MAX_CHUNK) chunk *= 2;
        if ((J.m + chunk - 1) / chunk > CB_MAX_ITEMS) return LZX_UNSUPPORTED;
        jd[i].key = b->job_key[i];
        jd[i].chunk = chunk;
        jd[i].nitems = (uint32_t)((J.m + chunk - 1) / chunk);
        jd[i].m = J.m;
        jd[i].digits = J.digits_dev;
        jd[i].out = J.out_dev;
        jd[i].part = off;
        off += ((size_t)jd[i].nitems * K->k * K->n * K->elem + 255) / 256 * 256;
        order[i] = i;
    }
    // largest job (key bytes streamed) first
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        const lzx_wide_key *kx = b->jobs[x].key, *ky = b->jobs[y].key;
        return b->jobs[x].m * kx->k * kx->n * kx->elem > b->jobs[y].m * ky->k * ky->n * ky->elem;
    });
    std::vector<uint2> items;
    for (size_t i : order)
        for (uint32_t c = 0; c < jd[i].nitems; c++) items.push_back(make_uint2((uint32_t)i, c));
    b->n_a = items.size();
    for (size_t i : order)
        for (uint32_t r = 0; r < b->jobs[i].key->k; r++) items.push_back(make_uint2((uint32_t)i, r));
    b->n_total = items.size();
    if (b->n_total >= 0xFFFFFFFFull) return LZX_UNSUPPORTED;
    b->part_bytes = off;

    std::vector<KeyDev> kd(b->keys.size());
    for (size_t i = 0; i < b->keys.size(); i++) kd[i] = b->keys[i]->dev;
    LZX_HIP(hipMalloc(&b->d_keys, std::max<size_t>(1, kd.size()) * sizeof(KeyDev)));
    LZX_HIP(hipMalloc(&b->d_jobs, std::max<size_t>(1, nj) * sizeof(JobDev)));
    LZX_HIP(hipMalloc(&b->d_items, std::max<size_t>(1, items.size()) * sizeof(uint2)));
    LZX_HIP(hipMalloc(&b->d_state, (2 + nj) * sizeof(uint32_t)));
    LZX_HIP(hipMalloc(&b->d_part, std::max<size_t>(256, b->part_bytes)));
    if (!kd.empty()) LZX_HIP(hipMemcpy(b->d_keys, kd.data(), kd.size() * sizeof(KeyDev), hipMemcpyHostToDevice));
    if (nj) LZX_HIP(hipMemcpy(b->d_jobs, jd.data(), nj * sizeof(JobDev), hipMemcpyHostToDevice));
    if (!items.empty())
        LZX_HIP(hipMemcpy(b->d_items, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice));

    int per_cu = 0, cus = 0;
    LZX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, commit_batch_kernel, CB_THREADS, 0));
    LZX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
    b->grid = (uint32_t)std::max<size_t>(1, std::min<size_t>(b->n_total, (size_t)std::max(1, per_cu) * cus));
    return LZX_OK;
}

lzx_status launch_direct(lzx_commit_batch *b, hipStream_t s)
{
    const uint32_t words = (uint32_t)(2 + b->jobs.size());
    hipLaunchKernelGGL(batch_reset_kernel, dim3(std::min<uint32_t>(1024, (words + 255) / 256)), dim3(256), 0, s,
                       b->d_state, words);
    LZX_LAUNCH_CHECK("batch_reset_kernel");
    if (b->n_total == 0) return LZX_OK;
    hipLaunchKernelGGL(commit_batch_kernel, dim3(b->grid), dim3(CB_THREADS), 0, s, b->d_keys, b->d_jobs, b->d_items,
                       (uint32_t)b->n_a, (uint32_t)b->n_total, b->d_state, b->d_part, CB_SPIN_LIMIT);
    LZX_LAUNCH_CHECK("commit_batch_kernel");
    return LZX_OK;
}

bool graph_enabled()
{
    const char *e = getenv("LZX_BATCH_GRAPH");
    return !(e && e[0] == '0');
}

lzx_status run_batch(lzx_commit_batch *b, hipStream_t s)
{
    if (!graph_enabled()) return launch_direct(b, s);
    if (!b->exec) {
        if (!b->cap_stream) LZX_HIP(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
        LZX_HIP(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
        lzx_status st = launch_direct(b, b->cap_stream);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
        if (st != LZX_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e != hipSuccess) return device_fail(e, "hipStreamEndCapture");
        b->graph = g;
        LZX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    LZX_HIP(hipGraphLaunch(b->exec, s));
    return LZX_OK;
}

template <int KIND, int LOG_N> lzx_status launch_key_ntt(const lzx_wide_key *key, const void *A, hipStream_t s)
{
    const size_t npolys = key->k * key->m;
    const uint32_t sb = CB_DATA_BYTES / ((1u << LOG_N) * key->elem);
    const uint32_t blocks = (uint32_t)std::min<size_t>((npolys + sb - 1) / sb, 4096);
    hipLaunchKernelGGL((wide_key_ntt_kernel<KIND, LOG_N>), dim3(blocks), dim3(CB_THREADS), 0, s, key->dev, A,
                       npolys);
    LZX_LAUNCH_CHECK("wide_key_ntt_kernel");
    return LZX_OK;
}

} // namespace

extern "C" lzx_status lzx_wide_key_create(uint64_t q, uint64_t generator, uint32_t log_n, size_t k, size_t m,
                                          const void *A_dev, void *stream, lzx_wide_key **out)
{
    if (!out) return LZX_INVALID_ARGUMENT;
    *out = nullptr;
    if (log_n < CB_MIN_LOG_N || log_n > CB_MAX_LOG_N || k == 0 || k > 4096 || q <= 256 || q >= (1ULL << 50) ||
        !(q & 1) || m > (1ULL << 32) || (m && !A_dev))
        return LZX_INVALID_ARGUMENT;
    const uint32_t n = 1u << log_n;
    if ((q - 1) % (2 * (uint64_t)n)) return LZX_NTT_ADICITY_TOO_SMALL;
    const uint64_t psi = powmod_h(generator, (q - 1) / (2 * n), q);
    if (powmod_h(psi, n, q) != q - 1) return LZX_INVALID_ARGUMENT;
    const uint64_t psi_inv = powmod_h(psi, q - 2, q), n_inv = powmod_h(n, q - 2, q);

    lzx_wide_key *key = new lzx_wide_key();
    key->q = q, key->generator = generator, key->log_n = log_n, key->n = n, key->k = k, key->m = m;
    key->kind = q < (1ULL << 31) ? WK_S32 : q < (1ULL << 32) ? WK_M32 : WK_W64;
    key->elem = key->kind == WK_W64 ? 8 : 4;
    (void)hipGetDevice(&key->device);
    std::vector<uint64_t> tw(4 * n, 0);
    for (uint32_t i = 1; i < n; i++) {
        tw_pair(key->kind, powmod_h(psi, brv(i, log_n), q), q, tw[i], tw[n + i]);
        tw_pair(key->kind, powmod_h(psi_inv, brv(i, log_n), q), q, tw[2 * n + i], tw[3 * n + i]);
    }
    KeyDev &d = key->dev;
    memset(&d, 0, sizeof d);
    d.q = q, d.m = m, d.k = (uint32_t)k, d.log_n = log_n, d.kind = key->kind;
    tw_pair(key->kind, n_inv, q, d.ninv, d.ninvp);
    if (key->kind != WK_W64) {
        d.qinv = host_inv32((uint32_t)q);
        d.barrett = UINT64_MAX / q;
        d.r64 = (uint64_t)(((unsigned __int128)1 << 64) % q);
    } else {
        d.one_p = shoup64_companion(1, q);
        d.c25 = (1ULL << 25) % q, d.c25p = shoup64_companion(d.c25, q);
        d.c50 = (uint64_t)((1ULL << 50) % q), d.c50p = shoup64_companion(d.c50, q);
    }
    hipError_t e = hipMalloc(&key->tw, tw.size() * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemcpy(key->tw, tw.data(), tw.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMalloc(&key->ahat, std::max<size_t>(16, k * m * n * key->elem));
    if (e != hipSuccess) {
        lzx_status st = device_fail(e, "lzx_wide_key_create");
        lzx_wide_key_destroy(key);
        return st;
    }
    d.ahat = key->ahat, d.tw = key->tw;
    lzx_status st = LZX_OK;
    if (m) {
        hipStream_t s = (hipStream_t)stream;
        switch (class_of(key->kind, log_n)) {
        case 0: st = launch_key_ntt<WK_S32, 6>(key, A_dev, s); break;
        case 1: st = launch_key_ntt<WK_S32, 7>(key, A_dev, s); break;
        case 2: st = launch_key_ntt<WK_S32, 8>(key, A_dev, s); break;
        case 3: st = launch_key_ntt<WK_M32, 6>(key, A_dev, s); break;
        case 4: st = launch_key_ntt<WK_M32, 7>(key, A_dev, s); break;
        case 5: st = launch_key_ntt<WK_M32, 8>(key, A_dev, s); break;
        case 6: st = launch_key_ntt<WK_W64, 6>(key, A_dev, s); break;
        case 7: st = launch_key_ntt<WK_W64, 7>(key, A_dev, s); break;
        default: st = launch_key_ntt<WK_W64, 8>(key, A_dev, s); break;
        }
    }
    if (st != LZX_OK) {
        lzx_wide_key_destroy(key);
        return st;
    }
    *out = key;
    return LZX_OK;
}

extern "C" void lzx_wide_key_destroy(lzx_wide_key *key)
{
    if (!key) return;
    (void)hipFree(key->ahat);
    (void)hipFree(key->tw);
    delete key;
}

extern "C" uint32_t lzx_wide_key_elem_bytes(const lzx_wide_key *key) { return key ? key->elem : 0; }

extern "C" lzx_status lzx_commit_batch_create(const lzx_commit_job *jobs, size_t njobs, lzx_commit_batch **out)
{
    if (!out || (njobs && !jobs) || njobs >= (1u << 31)) return LZX_INVALID_ARGUMENT;
    *out = nullptr;
    lzx_commit_batch *b = new lzx_commit_batch();
    (void)hipGetDevice(&b->device);
    std::map<const lzx_wide_key *, uint32_t> kidx;
    for (size_t i = 0; i < njobs; i++) {
        const lzx_commit_job &J = jobs[i];
        if (!J.key || J.m > J.key->m || !J.out_dev || (J.m && !J.digits_dev) ||
            ((uintptr_t)J.digits_dev % 16) || ((uintptr_t)J.out_dev % J.key->elem) || J.key->device != b->device) {
            delete b;
            return J.key && J.m > J.key->m ? LZX_AJTAI_DIMENSION_MISMATCH : LZX_INVALID_ARGUMENT;
        }
        auto it = kidx.find(J.key);
        if (it == kidx.end()) {
            it = kidx.emplace(J.key, (uint32_t)b->keys.size()).first;
            b->keys.push_back(J.key);
        }
        b->jobs.push_back(J);
        b->job_key.push_back(it->second);
        b->present[job_class(J.key)] = true;
    }
    const uint32_t forced = env_chunk();
    for (const lzx_wide_key *K : b->keys) {
        const uint32_t c = job_class(K);
        uint32_t v = 0;
        if (forced)
            v = forced;
        else if (!cache_get(cache_key(K->kind, K->log_n, K->k), v))
            v = default_chunk(K->k, K->n, K->elem);
        b->chunk[c] = std::max(b->chunk[c], v);
    }
    lzx_status st = plan_batch(b);
    if (st != LZX_OK) {
        lzx_commit_batch_destroy(b);
        return st;
    }
    *out = b;
    return LZX_OK;
}

extern "C" lzx_status lzx_commit_batch_run(lzx_commit_batch *b, void *stream)
{
    if (!b) return LZX_INVALID_ARGUMENT;
    return run_batch(b, (hipStream_t)stream);
}

extern "C" lzx_status lzx_commit_batch_check(lzx_commit_batch *b, void *stream)
{
    if (!b) return LZX_INVALID_ARGUMENT;
    LZX_HIP(hipStreamSynchronize((hipStream_t)stream));
    uint32_t err = 0;
    LZX_HIP(hipMemcpy(&err, b->d_state + 1, sizeof err, hipMemcpyDeviceToHost));
    if (err) return device_fail(hipErrorLaunchTimeOut, "commit batch: a phase B item waited past its bound");
    return LZX_OK;
}

namespace {

// Mean of `reps` timed runs of the whole batch after one warm-up run.
lzx_status time_batch(lzx_commit_batch *b, hipStream_t s, hipEvent_t e0, hipEvent_t e1, int reps, float &ms_out)
{
    float total = 0, ms = 0;
    for (int rep = 0; rep <= reps; rep++) {
        LZX_HIP(hipEventRecord(e0, s));
        lzx_status st = launch_direct(b, s);
        if (st != LZX_OK) return st;
        LZX_HIP(hipEventRecord(e1, s));
        LZX_HIP(hipEventSynchronize(e1));
        LZX_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (rep > 0) total += ms;
    }
    uint32_t err = 0;
    LZX_HIP(hipMemcpy(&err, b->d_state + 1, sizeof err, hipMemcpyDeviceToHost));
    if (err) return device_fail(hipErrorLaunchTimeOut, "commit batch autotune: a dependency wait hit its bound");
    ms_out = total / reps;
    return LZX_OK;
}

} // namespace

// One pass of coordinate descent over the classes present: each class's
// candidate chunks are timed on the WHOLE batch with the other classes at their
// current choice (a class alone favours smaller items than the mixed queue).
extern "C" lzx_status lzx_commit_batch_autotune(lzx_commit_batch *b, void *stream)
{
    if (!b) return LZX_INVALID_ARGUMENT;
    hipStream_t s = (hipStream_t)stream;
    static const uint32_t CANDS[] = {64, 128, 256, 512, 1024, 2048};
    hipEvent_t e0, e1;
    LZX_HIP(hipEventCreate(&e0));
    LZX_HIP(hipEventCreate(&e1));
    lzx_status st = LZX_OK;
    for (uint32_t c = 0; c < CB_CLASSES && st == LZX_OK; c++) {
        if (!b->present[c]) continue;
        const uint32_t start = b->chunk[c];
        uint32_t best_chunk = start;
        float best = 0;
        for (uint32_t cand : CANDS) {
            b->chunk[c] = cand;
            st = plan_batch(b);
            if (st == LZX_UNSUPPORTED) { // item limit: this candidate is too small for the class
                st = LZX_OK;
                continue;
            }
            float ms = 0;
            if (st == LZX_OK) st = time_batch(b, s, e0, e1, 3, ms);
            if (st != LZX_OK) break;
            if (best == 0 || ms < best) best = ms, best_chunk = cand;
        }
        b->chunk[c] = st == LZX_OK ? best_chunk : start;
        if (st == LZX_OK)
            for (const lzx_wide_key *K : b->keys)
                if (job_class(K) == c) cache_put(cache_key(K->kind, K->log_n, K->k), best_chunk);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    lzx_status st2 = plan_batch(b);
    return st != LZX_OK ? st : st2;
}

extern "C" lzx_status lzx_commit_batch_info(const lzx_commit_batch *b, size_t *items, size_t *scratch_bytes,
                                            uint32_t chunks[9])
{
    if (!b) return LZX_INVALID_ARGUMENT;
    if (items) *items = b->n_total;
    if (scratch_bytes) *scratch_bytes = b->part_bytes;
    if (chunks)
        for (uint32_t c = 0; c < CB_CLASSES; c++) chunks[c] = b->present[c] ? b->chunk[c] : 0;
    return LZX_OK;
}

extern "C" void lzx_commit_batch_destroy(lzx_commit_batch *b)
{
    if (!b) return;
    free_plan(b);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    delete b;
}
