// Band overlap matrices on the matrix cores: C = alpha * A B^H for
// band-major arrays A (m rows) and B (n rows) of length k -- the subspace
// matrices S = <psi|psi> and H = <psi|H|psi> of the eigensolvers (the
// reference forms them with rk / r2k / gemm, gpaw/utilities/blas.py:39,273,
// 314 -> c/blas.c:48,101,140, called from gpaw/matrix.py and
// gpaw/wavefunctions/arrays.py:262-340).
//
// The shape is tall-skinny: m, n are band counts (a few hundred) and k is a
// grid (10^5 .. 10^7 points), so the work is split along k.  A workgroup
// owns one block pair (BM rows of A x BN rows of B) and one k-chunk; it
// streams 128-byte row segments of both blocks through LDS and multiplies
// them on the MFMA units:
//   float / complex64   v_mfma_f32_32x32x2f32  (FP32 products and sums)
//   double / complex128 v_mfma_f64_16x16x4f64
// Each chunk's partial block is written to a workspace in the element
// precision, and a second kernel adds the chunks in FP64 in a fixed order
// (deterministic; FP32 sums never run over more than one chunk).
//
// Complex data: with A = a + ib and B = a' + ib' interleaved,
//   Re C = sum (a a' + b b')      -- the real product of the interleaved rows
//   Im C = sum (b a' - a b')      -- the same A operand against (-b', a')
// so each B value feeds two MFMAs (one per part) and no conjugate copy of B
// is formed.
//
// Hermitian mode (m == n, the product known to be Hermitian, e.g. S or H):
// only blocks holding some element i >= j are computed and the upper
// triangle is the conjugate of the lower one (what rk / r2k's "lower
// triangle" contract gives the reference's callers).
//
// k-order inside a 128-byte segment: lane group g of an MFMA supplies
// elements [g W, (g + 1) W) of the segment (W = elements per segment / K of
// the instruction), the same permutation for A and B, so each lane reads W
// contiguous elements with 16-byte LDS reads.
#include <algorithm>

#include "common.h"

namespace gh {

namespace {

template <typename T> struct Mfma;

template <> struct Mfma<float> {
    static constexpr int TILE = 32;  // output tile TILE x TILE
    static constexpr int KOP = 2;    // k per instruction
    static constexpr int NACC = 16;  // accumulators per lane
    typedef float acc_t __attribute__((ext_vector_type(16)));
    static __device__ __forceinline__ acc_t op(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, c, 0, 0, 0);
    }
    // output row of accumulator r in lane group g (column = lane % TILE)
    static __device__ __forceinline__ int row(int r, int g) {
        return 8 * (r / 4) + 4 * g + (r % 4);
    }
};

template <> struct Mfma<double> {
    static constexpr int TILE = 16;
    static constexpr int KOP = 4;
    static constexpr int NACC = 4;
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t op(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int r, int g) { return g + 4 * r; }
};

constexpr int kSeg = 128;          // bytes of a row per stage
constexpr int kRowPitch = 144;     // LDS bytes per row (16-byte pad: no bank conflicts)
constexpr int kBM = 128;           // rows of A per block
constexpr int kThreads = 256;      // 4 waves: 2 x 2 over the block

// Block pairs in Hermitian mode: (I, J) with some i >= j, i.e.
// J <= q (I + 1) - 1 for q = BM / BN; row I holds q (I + 1) of them.
__host__ __device__ inline int herm_pair_index(int I, int J, int q) {
    return q * I * (I + 1) / 2 + J;
}

__device__ inline void herm_pair_of(int p, int q, int& I, int& J) {
    int i = 0;
    while (p >= q * (i + 1)) {
        p -= q * (i + 1);
        i++;
    }
    I = i;
    J = p;
}

template <typename T, bool CPLX, int WN>
__global__ void __launch_bounds__(kThreads)
gram_mfma_kernel(const char* __restrict__ A, const char* __restrict__ B,
                 int64_t lda, int64_t ldb,          // row pitch in bytes
                 int m, int n, int64_t nseg,        // segments of kSeg bytes per row
                 int64_t kbytes,                    // valid bytes per row
                 int herm, int npairs, int nJ, int nchunks,
                 T* __restrict__ work) {
    using M = Mfma<T>;
    constexpr int TILE = M::TILE, KOP = M::KOP, NACC = M::NACC;
    constexpr int BN = 2 * WN;
    constexpr int MT = 64 / TILE;            // row tiles per wave
    constexpr int NT = WN / TILE;            // column tiles per wave
    constexpr int EL = kSeg / sizeof(T);     // elements per segment
    constexpr int W = EL / KOP;              // elements per lane group
    constexpr int NC = CPLX ? 2 : 1;
    constexpr int AV = kBM * (kSeg / 16) / kThreads;   // 16-byte loads per thread
    constexpr int BV = BN * (kSeg / 16) / kThreads;
    static_assert(W % 2 == 0 && (W * sizeof(T)) % 16 == 0, "lane run");
    static_assert(AV >= 1 && BV >= 1, "loader");

    __shared__ __align__(16) char sA[kBM * kRowPitch];
    __shared__ __align__(16) char sB[BN * kRowPitch];

    // block id -> (chunk, pair); ids are dealt round-robin to the 8 XCDs,
    // so the pairs of one chunk (which read the same rows) are placed on
    // one XCD
    const int nblk = npairs * nchunks;
    const int lin = xcd_swizzle(blockIdx.x, nblk);
    const int pair = lin % npairs;
    const int chunk = lin / npairs;
    int I, J;
    if (herm) herm_pair_of(pair, kBM / BN, I, J);
    else { I = pair / nJ; J = pair % nJ; }
    const int row0 = I * kBM, col0 = J * BN;
    const bool sameAB = (A == B) && (lda == ldb) && row0 == col0 && BN == kBM;

    const int64_t s_beg = nseg * chunk / nchunks;
    const int64_t s_end = nseg * (chunk + 1) / nchunks;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int lg = lane / TILE, lr = lane % TILE;

    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 ra[AV], rb[BV];

    auto load = [&](int64_t s) {
        const int64_t kb0 = s * kSeg;
#pragma unroll
        for (int v = 0; v < AV; v++) {
            const int idx = tid + v * kThreads;
            const int r = idx / (kSeg / 16), c = idx % (kSeg / 16);
            const int64_t kb = kb0 + c * 16;
            f4 x = {0.f, 0.f, 0.f, 0.f};
            if (row0 + r < m && kb < kbytes)
                x = *(const f4*)(A + (int64_t)(row0 + r) * lda + kb);
            ra[v] = x;
        }
        if (!sameAB) {
#pragma unroll
            for (int v = 0; v < BV; v++) {
                const int idx = tid + v * kThreads;
                const int r = idx / (kSeg / 16), c = idx % (kSeg / 16);
                const int64_t kb = kb0 + c * 16;
                f4 x = {0.f, 0.f, 0.f, 0.f};
                if (col0 + r < n && kb < kbytes)
                    x = *(const f4*)(B + (int64_t)(col0 + r) * ldb + kb);
                rb[v] = x;
            }
        }
    };
    auto stash = [&]() {
#pragma unroll
        for (int v = 0; v < AV; v++) {
            const int idx = tid + v * kThreads;
            const int r = idx / (kSeg / 16), c = idx % (kSeg / 16);
            *(f4*)(sA + r * kRowPitch + c * 16) = ra[v];
        }
        if (!sameAB) {
#pragma unroll
            for (int v = 0; v < BV; v++) {
                const int idx = tid + v * kThreads;
                const int r = idx / (kSeg / 16), c = idx % (kSeg / 16);
                *(f4*)(sB + r * kRowPitch + c * 16) = rb[v];
            }
        }
    };

    typename M::acc_t acc[NC][MT][NT];
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int i = 0; i < MT; i++)
#pragma unroll
            for (int j = 0; j < NT; j++)
#pragma unroll
                for (int r = 0; r < NACC; r++) acc[c][i][j][r] = T(0);

    const char* sBr = sameAB ? sA : sB;
    if (s_beg < s_end) load(s_beg);
    for (int64_t s = s_beg; s < s_end; s++) {
        __syncthreads();          // the previous stage's LDS reads are done
        stash();
        __syncthreads();
        if (s + 1 < s_end) load(s + 1);   // in flight during the products
        T a[MT][W], b[NT][W];
#pragma unroll
        for (int i = 0; i < MT; i++) {
            const char* p = sA + (wr * 64 + i * TILE + lr) * kRowPitch + lg * W * sizeof(T);
#pragma unroll
            for (int q = 0; q < W * (int)sizeof(T) / 16; q++) {
                f4 x = *(const f4*)(p + 16 * q);
                __builtin_memcpy(&a[i][q * 16 / sizeof(T)], &x, 16);
            }
        }
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const char* p = sBr + (wc * WN + j * TILE + lr) * kRowPitch + lg * W * sizeof(T);
#pragma unroll
            for (int q = 0; q < W * (int)sizeof(T) / 16; q++) {
                f4 x = *(const f4*)(p + 16 * q);
                __builtin_memcpy(&b[j][q * 16 / sizeof(T)], &x, 16);
            }
        }
#pragma unroll
        for (int e = 0; e < W; e++) {
#pragma unroll
            for (int i = 0; i < MT; i++)
#pragma unroll
                for (int j = 0; j < NT; j++) {
                    acc[0][i][j] = M::op(a[i][e], b[j][e], acc[0][i][j]);
                    if constexpr (CPLX) {
                        // (re, im) of B at e even / odd -> (-im, re)
                        const T y = (e & 1) ? b[j][e - 1] : -b[j][e + 1];
                        acc[NC - 1][i][j] = M::op(a[i][e], y, acc[NC - 1][i][j]);
                    }
                }
        }
    }

    // partial block -> work[(pair, chunk)][part][BM][BN]
    T* out = work + ((int64_t)pair * nchunks + chunk) * NC * kBM * BN;
#pragma unroll
    for (int c = 0; c < NC; c++)
#pragma unroll
        for (int i = 0; i < MT; i++)
#pragma unroll
            for (int j = 0; j < NT; j++)
#pragma unroll
                for (int r = 0; r < NACC; r++) {
                    const int row = wr * 64 + i * TILE + M::row(r, lg);
                    const int col = wc * WN + j * TILE + lr;
                    out[(int64_t)c * kBM * BN + row * BN + col] = acc[c][i][j][r];
                }
}

// C[i][j] = alpha * sum over chunks (FP64, chunk order) of the partial
// blocks; Hermitian mode mirrors the lower triangle; conj_out conjugates.
template <typename T, bool CPLX, int BN>
__global__ void __launch_bounds__(256)
gram_reduce_kernel(const T* __restrict__ work, int m, int n, int herm, int nJ,
                   int nchunks, double alpha, int conj_out,
                   double* __restrict__ C) {
    constexpr int NC = CPLX ? 2 : 1;
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= (int64_t)m * n) return;
    const int i = (int)(t / n), j = (int)(t % n);
    int si = i, sj = j;
    bool flip = false;
    if (herm && i < j) {
        si = j;
        sj = i;
        flip = true;
    }
    const int I = si / kBM, J = sj / BN;
    const int pair = herm ? herm_pair_index(I, J, kBM / BN) : I * nJ + J;
    const int li = si % kBM, lj = sj % BN;
    const T* p = work + (int64_t)pair * nchunks * NC * kBM * BN + li * BN + lj;
    double re = 0.0, im = 0.0;
    for (int c = 0; c < nchunks; c++) {
        re += (double)p[(int64_t)c * NC * kBM * BN];
        if (CPLX) im += (double)p[(int64_t)c * NC * kBM * BN + kBM * BN];
    }
    if (flip) im = -im;
    if (conj_out) im = -im;
    if (herm && i == j) im = 0.0;   // a Hermitian matrix has a real diagonal
    if (CPLX) {
        C[2 * t] = alpha * re;
        C[2 * t + 1] = alpha * im;
    } else {
        C[t] = alpha * re;
    }
}

struct GramPlan {
    int BN, npairs, nI, nJ, nchunks;
    int64_t nseg, work_bytes;
};

GramPlan gram_plan(int dtype, int64_t m, int64_t n, int64_t kbytes, int herm) {
    GramPlan p;
    p.BN = dtype == GH_C128 ? 64 : 128;
    p.nI = (int)((m + kBM - 1) / kBM);
    p.nJ = (int)((n + p.BN - 1) / p.BN);
    const int q = kBM / p.BN;
    p.npairs = herm ? q * p.nI * (p.nI + 1) / 2 : p.nI * p.nJ;
    p.nseg = (kbytes + kSeg - 1) / kSeg;
    // about 1024 workgroups (4 per CU), a multiple of the XCD count, and
    // at least 16 segments per chunk
    int64_t nc = std::max<int64_t>(1, 1024 / std::max(p.npairs, 1));
    nc = std::min<int64_t>(nc, std::max<int64_t>(1, p.nseg / 16));
    while (nc > 1 && (p.npairs * nc) % kNumXCD) nc--;
    p.nchunks = (int)nc;
    const int es = dtype_size(dtype);   // complex: two parts of half size
    p.work_bytes = (int64_t)p.npairs * p.nchunks * kBM * p.BN * es;
    return p;
}

template <typename T, bool CPLX, int WN>
int launch_gram_mfma(const GramPlan& pl, int m, int n, int64_t kbytes,
                     const void* A, int64_t lda, const void* B, int64_t ldb,
                     double alpha, int herm, int conj_out, double* C, void* work,
                     hipStream_t st) {
    const int nblk = pl.npairs * pl.nchunks;
    if (nblk > 0 && pl.nseg > 0) {
        hipLaunchKernelGGL((gram_mfma_kernel<T, CPLX, WN>), dim3(nblk), dim3(kThreads), 0, st,
                           (const char*)A, (const char*)B, lda, ldb, m, n, pl.nseg, kbytes,
                           herm, pl.npairs, pl.nJ, pl.nchunks, (T*)work);
        GH_CHECK_LAUNCH();
    }
    const int64_t total = (int64_t)m * n;
    const int nchunks = pl.nseg > 0 ? pl.nchunks : 0;
    hipLaunchKernelGGL((gram_reduce_kernel<T, CPLX, 2 * WN>),
                       dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       (const T*)work, m, n, herm, pl.nJ, nchunks, alpha, conj_out, C);
    GH_CHECK_LAUNCH();
    return GH_OK;
}

}  // namespace

}  // namespace gh

using namespace gh;

extern "C" {

int64_t gh_gram_work_bytes(int dtype, int64_t m, int64_t n, int64_t k, int hermitian) {
    if (dtype < GH_F64 || dtype > GH_C64 || m < 0 || n < 0 || k < 0) return -1;
    return gram_plan(dtype, m, n, k * dtype_size(dtype), hermitian).work_bytes;
}

int gh_gram(int dtype, int64_t m, int64_t n, int64_t k, const void* A_dev,
            int64_t lda, const void* B_dev, int64_t ldb, double alpha,
            int hermitian, int conj_out, double* C_dev, void* work_dev,
            int64_t work_bytes, void* stream) {
    GH_REQUIRE(dtype >= GH_F64 && dtype <= GH_C64, GH_ERR_ARG, "gram: bad dtype");
    GH_REQUIRE(m >= 0 && n >= 0 && k >= 0 && m < (1 << 20) && n < (1 << 20), GH_ERR_ARG,
               "gram: bad sizes m=%lld n=%lld k=%lld", (long long)m, (long long)n,
               (long long)k);
    GH_REQUIRE(!hermitian || m == n, GH_ERR_SHAPE, "gram: Hermitian mode needs m == n");
    if (m == 0 || n == 0) return GH_OK;
    GH_REQUIRE(A_dev && B_dev && C_dev, GH_ERR_ARG, "gram: null argument");
    GH_REQUIRE(lda >= k && ldb >= k, GH_ERR_ARG, "gram: row pitch below k");
    const int es = dtype_size(dtype);
    const int64_t kbytes = k * es;
    GH_REQUIRE(kbytes % 16 == 0 && (lda * es) % 16 == 0 && (ldb * es) % 16 == 0,
               GH_ERR_UNSUPPORTED, "gram: rows must be multiples of 16 bytes");
    GH_REQUIRE(((uintptr_t)A_dev & 15) == 0 && ((uintptr_t)B_dev & 15) == 0,
               GH_ERR_UNSUPPORTED, "gram: A and B must be 16-byte aligned");
    const GramPlan pl = gram_plan(dtype, m, n, kbytes, hermitian);
    GH_REQUIRE(work_dev || pl.work_bytes == 0, GH_ERR_ARG, "gram: null workspace");
    GH_REQUIRE(work_bytes >= pl.work_bytes, GH_ERR_ARG,
               "gram: workspace of %lld bytes, %lld needed", (long long)work_bytes,
               (long long)pl.work_bytes);
    hipStream_t st = (hipStream_t)stream;
    const int mi = (int)m, ni = (int)n;
    const int64_t la = lda * es, lb = ldb * es;
    switch (dtype) {
        case GH_F32:
            return launch_gram_mfma<float, false, 64>(pl, mi, ni, kbytes, A_dev, la, B_dev, lb,
                                                      alpha, hermitian, conj_out, C_dev,
                                                      work_dev, st);
        case GH_C64:
            return launch_gram_mfma<float, true, 64>(pl, mi, ni, kbytes, A_dev, la, B_dev, lb,
                                                     alpha, hermitian, conj_out, C_dev,
                                                     work_dev, st);
        case GH_F64:
            return launch_gram_mfma<double, false, 64>(pl, mi, ni, kbytes, A_dev, la, B_dev,
                                                       lb, alpha, hermitian, conj_out, C_dev,
                                                       work_dev, st);
        case GH_C128:
            return launch_gram_mfma<double, true, 32>(pl, mi, ni, kbytes, A_dev, la, B_dev,
                                                      lb, alpha, hermitian, conj_out, C_dev,
                                                      work_dev, st);
    }
    return GH_ERR_ARG;
}

}  // extern "C"
