// Band-wise BLAS-1 kernels and the multigrid V-cycle on gfx950.
//
// multi_axpy follows gpaw/utilities/mblas.py:13-32 (y[x] += a[x] x[x]);
// multi_dotc gives the per-band integrals of rmmdiis.py:124-130,158-164 with
// FP64 accumulation in a fixed reduction order (deterministic).
//
// The V-cycle computes exactly the 14 steps of the reference
// (c/fd_preconditioner.c:199-349, gpaw/preconditioner.py:74-93), regrouped
// so that every scalar update is folded into a stencil or transformer
// epilogue (L = Laplacian of the level, s = step, R0/R1 restrictions,
// I2/I1 interpolations):
//   1. r1 = -R0(R)
//   2. t1 = 4s L1 r1 - r1                      (= q1 of step 3)
//   3. r2 = R1(t1)
//   4. t2 = 32s r2 - 256 s^2 L2 r2             (= d2 after step 7)
//   5. u1 = I2(t2)                             (= q1 of step 8)
//   6. t1 = 4s r1 - u1                         (= d1 after step 9)
//   7. u1 = t1 - 4s (L1 t1 - r1)               (= d1 after step 11)
//   8. d0' = -I1(u1)                           (step 12)
//   9. d0 = s (L0 d0' - R) - d0'               (steps 13-14)
// Nine kernels in place of the reference's twenty array passes.
#include <algorithm>
#include <mutex>

#include "common.h"
#include "internal.h"

namespace gh {

constexpr int kDotChunks = 64;

template <typename V, typename S>
__global__ void multi_axpy_kernel(const double* __restrict__ a,
                                  const V* __restrict__ x, V* __restrict__ y,
                                  int64_t n, int64_t total) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x) {
        int64_t band = i / n;
        y[i] = fma_rs((S)a[band], x[i], y[i]);
    }
}

// z = alpha * x + beta * y  (real scalars)
template <typename V, typename S>
__global__ void axpbyz_kernel(S alpha, const V* __restrict__ x, S beta,
                              const V* __restrict__ y, V* __restrict__ z,
                              int64_t total) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total;
         i += (int64_t)gridDim.x * blockDim.x)
        z[i] = vadd(vscale(x[i], alpha), vscale(y[i], beta));
}

// out[x] = sum_i c[x, i] v_i[x] per band x: the DIIS combination
// psi = alpha_nit psi_nit + sum_{i<nit} alpha_i psi_i of the reference
// (eigensolvers/rmmdiis.py:226-237) and out-of-place axpy (v_0 + lam v_1) in
// one pass.  Terms are accumulated in the order given.  Coefficients are
// FP64 (real data) or FP64 (re, im) pairs (complex data).  `out` may be one
// of the inputs: every element is read before it is written, by the same
// thread.
constexpr int kLincombMax = 8;
template <typename V> struct LincombPtrs { const V* v[kLincombMax]; };

template <typename S>
__device__ __forceinline__ S coef_mul(S v, S cr, S) { return cr * v; }
template <typename S>
__device__ __forceinline__ cplx<S> coef_mul(cplx<S> v, S cr, S ci) {
    return cplx<S>{cr * v.re - ci * v.im, cr * v.im + ci * v.re};
}
template <typename S>
__device__ __forceinline__ S coef_fma(S v, S cr, S, S acc) { return fma(cr, v, acc); }
template <typename S>
__device__ __forceinline__ cplx<S> coef_fma(cplx<S> v, S cr, S ci, cplx<S> acc) {
    return cplx<S>{fma(cr, v.re, fma(-ci, v.im, acc.re)),
                   fma(cr, v.im, fma(ci, v.re, acc.im))};
}

template <typename V, typename S, int K, int EPT>
__global__ void __launch_bounds__(256)
multi_lincomb_kernel(LincombPtrs<V> p, const double* __restrict__ coef, V* out,
                     int64_t n) {
    constexpr bool CPLX = is_complex<V>::value;
    const int64_t band = blockIdx.y;
    S cr[K], ci[K];
#pragma unroll
    for (int i = 0; i < K; i++) {
        if (CPLX) {
            cr[i] = (S)coef[2 * (band * K + i)];
            ci[i] = (S)coef[2 * (band * K + i) + 1];
        } else {
            cr[i] = (S)coef[band * K + i];
            ci[i] = 0;
        }
    }
    const size_t off = (size_t)band * n;
    const int64_t step = (int64_t)gridDim.x * blockDim.x * EPT;
    for (int64_t i0 = (int64_t)blockIdx.x * blockDim.x * EPT + threadIdx.x; i0 < n;
         i0 += step) {
        V x[EPT][K];
#pragma unroll
        for (int e = 0; e < EPT; e++) {
            const int64_t idx = i0 + (int64_t)e * blockDim.x;
#pragma unroll
            for (int i = 0; i < K; i++)
                x[e][i] = idx < n ? p.v[i][off + idx] : vzero<V>();
        }
#pragma unroll
        for (int e = 0; e < EPT; e++) {
            const int64_t idx = i0 + (int64_t)e * blockDim.x;
            V acc = coef_mul(x[e][0], cr[0], ci[0]);
#pragma unroll
            for (int i = 1; i < K; i++) acc = coef_fma(x[e][i], cr[i], ci[i], acc);
            if (idx < n) out[off + idx] = acc;
        }
    }
}

template <typename V, typename S, int K>
void launch_lincomb(int64_t nb, int64_t n, const void* const* vecs,
                    const double* coef, void* out, hipStream_t st) {
    constexpr int EPT = K <= 2 ? 4 : 2;
    LincombPtrs<V> p{};
    for (int i = 0; i < K; i++) p.v[i] = (const V*)vecs[i];
    const int64_t per = 256 * EPT;
    int64_t chunks = (n + per - 1) / per;
    chunks = std::max<int64_t>(1, std::min<int64_t>(chunks, (8192 + nb - 1) / nb));
    hipLaunchKernelGGL((multi_lincomb_kernel<V, S, K, EPT>),
                       dim3((unsigned)chunks, (unsigned)nb), dim3(256), 0, st, p, coef,
                       (V*)out, n);
}

template <typename V, typename S>
int lincomb_k(int k, int64_t nb, int64_t n, const void* const* vecs,
              const double* coef, void* out, hipStream_t st) {
    switch (k) {
        case 1: launch_lincomb<V, S, 1>(nb, n, vecs, coef, out, st); break;
        case 2: launch_lincomb<V, S, 2>(nb, n, vecs, coef, out, st); break;
        case 3: launch_lincomb<V, S, 3>(nb, n, vecs, coef, out, st); break;
        case 4: launch_lincomb<V, S, 4>(nb, n, vecs, coef, out, st); break;
        case 5: launch_lincomb<V, S, 5>(nb, n, vecs, coef, out, st); break;
        case 6: launch_lincomb<V, S, 6>(nb, n, vecs, coef, out, st); break;
        case 7: launch_lincomb<V, S, 7>(nb, n, vecs, coef, out, st); break;
        case 8: launch_lincomb<V, S, 8>(nb, n, vecs, coef, out, st); break;
        default: set_error("multi_lincomb: k = %d out of range 1..%d", k, kLincombMax);
                 return GH_ERR_ARG;
    }
    GH_CHECK_LAUNCH();
    return GH_OK;
}

__device__ __forceinline__ double dotc_term(double a, double b) { return a * b; }
__device__ __forceinline__ double dotc_term(float a, float b) {
    return (double)a * (double)b;
}
template <typename S>
__device__ __forceinline__ double dotc_term(cplx<S> a, cplx<S> b) {
    return (double)a.re * (double)b.re + (double)a.im * (double)b.im;
}

template <typename V>
__global__ void __launch_bounds__(256)
dotc_partial_kernel(const V* __restrict__ a, const V* __restrict__ b,
                    int64_t n, double* __restrict__ partial) {
    __shared__ double red[4];
    const int band = blockIdx.y;
    const int chunk = blockIdx.x;
    const V* ab = a + (size_t)band * n;
    const V* bb = b + (size_t)band * n;
    const int64_t per = (n + kDotChunks - 1) / kDotChunks;
    const int64_t i0 = chunk * per;
    const int64_t i1 = min(n, i0 + per);
    double s = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
        s += dotc_term(ab[i], bb[i]);
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0)
        partial[band * kDotChunks + chunk] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void dotc_final_kernel(const double* __restrict__ partial,
                                  double scale, double* __restrict__ out) {
    const int band = blockIdx.x;
    double s = partial[band * kDotChunks + threadIdx.x];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
    if (threadIdx.x == 0) out[band] = s * scale;
}

// Gram matrices of k vectors per band: G[x][i][j] = scale <v_i[x]|v_j[x]>,
// all k (k + 1) / 2 products in one pass over the data; FP64 accumulation,
// fixed reduction order.  Complex results are (re, im) pairs.
constexpr int kGramMax = 6;

template <typename V> struct GramPtrs { const V* v[kGramMax]; };

__device__ __forceinline__ void gram_term(double a, double b, double& re, double&) {
    re += a * b;
}
__device__ __forceinline__ void gram_term(float a, float b, double& re, double&) {
    re += (double)a * (double)b;
}
template <typename S>
__device__ __forceinline__ void gram_term(cplx<S> a, cplx<S> b, double& re, double& im) {
    re += (double)a.re * (double)b.re + (double)a.im * (double)b.im;
    im += (double)a.re * (double)b.im - (double)a.im * (double)b.re;
}

template <typename V, int K>
__global__ void __launch_bounds__(256)
gram_partial_kernel(GramPtrs<V> vp, int64_t n, double* __restrict__ partial) {
    constexpr int NP = K * (K + 1) / 2;
    __shared__ double red[4][2 * NP];
    const int band = blockIdx.y;
    const int chunk = blockIdx.x;
    const int64_t per = (n + kDotChunks - 1) / kDotChunks;
    const int64_t i0 = chunk * per;
    const int64_t i1 = min(n, i0 + per);
    double re[NP], im[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) re[p] = im[p] = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        V x[K];
#pragma unroll
        for (int a = 0; a < K; a++) x[a] = vp.v[a][(size_t)band * n + i];
        int p = 0;
#pragma unroll
        for (int a = 0; a < K; a++)
#pragma unroll
            for (int b = a; b < K; b++, p++) gram_term(x[a], x[b], re[p], im[p]);
    }
#pragma unroll
    for (int p = 0; p < NP; p++) {
        double r = re[p], m = im[p];
        for (int off = 32; off > 0; off >>= 1) {
            r += __shfl_xor(r, off);
            m += __shfl_xor(m, off);
        }
        if ((threadIdx.x & 63) == 0) {
            red[threadIdx.x >> 6][2 * p] = r;
            red[threadIdx.x >> 6][2 * p + 1] = m;
        }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < 2 * NP; q += blockDim.x)
        partial[((size_t)band * 2 * NP + q) * kDotChunks + chunk] =
            (red[0][q] + red[1][q]) + (red[2][q] + red[3][q]);
}

template <int K>
__global__ void gram_final_kernel(const double* __restrict__ partial, double scale,
                                  int cplx_out, double* __restrict__ out) {
    constexpr int NP = K * (K + 1) / 2;
    const int band = blockIdx.x;
    for (int q = 0; q < 2 * NP; q++) {
        double s = partial[((size_t)band * 2 * NP + q) * kDotChunks + threadIdx.x];
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (threadIdx.x == 0) {
            const int p = q / 2, part = q % 2;
            int a = 0, rem = p;
            while (rem >= K - a) { rem -= K - a; a++; }
            const int b = a + rem;
            s *= scale;
            if (cplx_out) {
                double* o = out + (size_t)band * K * K * 2;
                o[(a * K + b) * 2 + part] = s;
                o[(b * K + a) * 2 + part] = part ? -s : s;   // Hermitian
            } else if (!part) {
                double* o = out + (size_t)band * K * K;
                o[a * K + b] = s;
                o[b * K + a] = s;
            }
        }
    }
}

template <typename V, int K>
int launch_gram(int64_t nbands, int64_t n, const void* const* vecs, double scale,
                int cplx_out, double* out, double* work, hipStream_t st) {
    GramPtrs<V> vp;
    for (int a = 0; a < K; a++) vp.v[a] = (const V*)vecs[a];
    hipLaunchKernelGGL((gram_partial_kernel<V, K>), dim3(kDotChunks, (unsigned)nbands),
                       dim3(256), 0, st, vp, n, work);
    GH_CHECK_LAUNCH();
    hipLaunchKernelGGL((gram_final_kernel<K>), dim3((unsigned)nbands), dim3(kDotChunks), 0,
                       st, (const double*)work, scale, cplx_out, out);
    GH_CHECK_LAUNCH();
    return GH_OK;
}

template <typename V>
int gram_k(int k, int64_t nbands, int64_t n, const void* const* vecs, double scale,
           int cplx_out, double* out, double* work, hipStream_t st) {
    switch (k) {
        case 1: return launch_gram<V, 1>(nbands, n, vecs, scale, cplx_out, out, work, st);
        case 2: return launch_gram<V, 2>(nbands, n, vecs, scale, cplx_out, out, work, st);
        case 3: return launch_gram<V, 3>(nbands, n, vecs, scale, cplx_out, out, work, st);
        case 4: return launch_gram<V, 4>(nbands, n, vecs, scale, cplx_out, out, work, st);
        case 5: return launch_gram<V, 5>(nbands, n, vecs, scale, cplx_out, out, work, st);
        case 6: return launch_gram<V, 6>(nbands, n, vecs, scale, cplx_out, out, work, st);
    }
    set_error("multi_gram: k = %d out of range 1..%d", k, kGramMax);
    return GH_ERR_ARG;
}

// Fixed scratch for gh_multi_gram_noalloc: one array per GPU, alive until
// the process ends.  Host code never uses the symbol's name as an address
// (on the host it is only a stand-in); each device's address comes from
// hipGetSymbolAddress, taken once per device with that device current.
constexpr int64_t kNoAllocScratchBytes = 16ll << 20;   // 16 MiB
__device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double* g_scratch_ptr[kMaxDevices] = {};

int scratch_for_current_device(double** out) {
    int dev = 0;
    GH_CHECK_HIP(hipGetDevice(&dev));
    GH_REQUIRE(dev >= 0 && dev < kMaxDevices, GH_ERR_UNSUPPORTED,
               "multi_gram_noalloc: device %d beyond the %d supported", dev, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        GH_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        GH_REQUIRE(p != nullptr, GH_ERR_HIP, "multi_gram_noalloc: no scratch address");
        g_scratch_ptr[dev] = (double*)p;
    }
    *out = g_scratch_ptr[dev];
    return GH_OK;
}

// Pseudo-density accumulation (c/utilities.c:313-356, called per band by
// wavefunctions/fd.py:95-100): nt[G] += sum_n f_n |psi_n(G)|^2.  One thread
// per grid point sweeps the bands in order (the reference's order of the
// per-band additions), FP64 accumulation for every data type.
__device__ __forceinline__ double abs2(double a) { return a * a; }
__device__ __forceinline__ double abs2(float a) { return (double)a * (double)a; }
template <typename S>
__device__ __forceinline__ double abs2(cplx<S> a) {
    return (double)a.re * (double)a.re + (double)a.im * (double)a.im;
}

template <typename V>
__global__ void __launch_bounds__(256)
add_to_density_kernel(const double* __restrict__ f, const V* __restrict__ psit,
                      double* __restrict__ nt, int64_t nbands, int64_t n) {
    for (int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; g < n;
         g += (int64_t)gridDim.x * blockDim.x) {
        double acc = nt[g];
        for (int64_t b = 0; b < nbands; b++) acc += f[b] * abs2(psit[b * n + g]);
        nt[g] = acc;
    }
}

template <typename V, typename S>
int launch_axpbyz(double alpha, const void* x, double beta, const void* y,
                  void* z, int64_t total, hipStream_t st) {
    int64_t nblk = std::min<int64_t>((total + 255) / 256, 8192);
    hipLaunchKernelGGL((axpbyz_kernel<V, S>), dim3((unsigned)nblk), dim3(256), 0,
                       st, (S)alpha, (const V*)x, (S)beta, (const V*)y, (V*)z,
                       total);
    GH_CHECK_LAUNCH();
    return GH_OK;
}

int axpbyz(int dtype, double alpha, const void* x, double beta, const void* y,
           void* z, int64_t total, hipStream_t st) {
    switch (dtype) {
        case GH_F64: return launch_axpbyz<double, double>(alpha, x, beta, y, z, total, st);
        case GH_C128: return launch_axpbyz<cplx<double>, double>(alpha, x, beta, y, z, total, st);
        case GH_F32: return launch_axpbyz<float, float>(alpha, x, beta, y, z, total, st);
        case GH_C64: return launch_axpbyz<cplx<float>, float>(alpha, x, beta, y, z, total, st);
    }
    set_error("axpbyz: bad dtype");
    return GH_ERR_ARG;
}

}  // namespace gh

using namespace gh;

namespace gh {

// Plain device copy in 16-byte vectors, four in flight per thread,
// non-temporal both ways: the achievable HBM rate on this device, measured
// next to the kernels whose traffic it puts in context.
typedef float v4f __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256)
stream_copy_kernel(const v4f* __restrict__ src, v4f* __restrict__ dst, int64_t n16) {
    const int64_t stride = (int64_t)gridDim.x * 256 * 4;
    for (int64_t base = (int64_t)blockIdx.x * 256 * 4 + threadIdx.x; base < n16;
         base += stride) {
        v4f v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t i = base + u * 256;
            if (i < n16) v[u] = __builtin_nontemporal_load(src + i);
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t i = base + u * 256;
            if (i < n16) __builtin_nontemporal_store(v[u], dst + i);
        }
    }
}

}  // namespace gh

extern "C" {

int gh_stream_copy(const void* src_dev, void* dst_dev, int64_t nbytes, int nblocks,
                   void* stream) {
    GH_REQUIRE(src_dev && dst_dev && src_dev != dst_dev, GH_ERR_ARG, "stream_copy: bad pointers");
    GH_REQUIRE(nbytes >= 0 && nbytes % 16 == 0, GH_ERR_ARG,
               "stream_copy: byte count not a multiple of 16");
    GH_REQUIRE(((uintptr_t)src_dev & 15u) == 0 && ((uintptr_t)dst_dev & 15u) == 0, GH_ERR_ARG,
               "stream_copy: pointers not 16-byte aligned");
    if (nbytes == 0) return GH_OK;
    if (nblocks <= 0) nblocks = 2048;
    hipLaunchKernelGGL(gh::stream_copy_kernel, dim3((unsigned)nblocks), dim3(256), 0,
                       (hipStream_t)stream, (const gh::v4f*)src_dev, (gh::v4f*)dst_dev,
                       nbytes / 16);
    GH_CHECK_LAUNCH();
    return GH_OK;
}

int gh_multi_axpy(int dtype, int64_t nbands, int64_t n, const double* a_dev,
                  const void* x_dev, void* y_dev, void* stream) {
    GH_REQUIRE(n >= 0 && nbands >= 0, GH_ERR_ARG, "multi_axpy: negative size");
    int64_t total = nbands * n;
    if (total == 0) return GH_OK;  // empty arrays may carry null pointers
    GH_REQUIRE(a_dev && x_dev && y_dev, GH_ERR_ARG, "multi_axpy: null argument");
    hipStream_t st = (hipStream_t)stream;
    int64_t nblk = std::min<int64_t>((total + 255) / 256, 8192);
    switch (dtype) {
        case GH_F64: hipLaunchKernelGGL((multi_axpy_kernel<double, double>), dim3(nblk), dim3(256), 0, st, a_dev, (const double*)x_dev, (double*)y_dev, n, total); break;
        case GH_C128: hipLaunchKernelGGL((multi_axpy_kernel<cplx<double>, double>), dim3(nblk), dim3(256), 0, st, a_dev, (const cplx<double>*)x_dev, (cplx<double>*)y_dev, n, total); break;
        case GH_F32: hipLaunchKernelGGL((multi_axpy_kernel<float, float>), dim3(nblk), dim3(256), 0, st, a_dev, (const float*)x_dev, (float*)y_dev, n, total); break;
        case GH_C64: hipLaunchKernelGGL((multi_axpy_kernel<cplx<float>, float>), dim3(nblk), dim3(256), 0, st, a_dev, (const cplx<float>*)x_dev, (cplx<float>*)y_dev, n, total); break;
        default: set_error("multi_axpy: bad dtype"); return GH_ERR_ARG;
    }
    GH_CHECK_LAUNCH();
    return GH_OK;
}

int gh_multi_dotc(int dtype, int64_t nbands, int64_t n, const void* a_dev,
                  const void* b_dev, double scale, double* out_dev,
                  double* work_dev, void* stream) {
    GH_REQUIRE(n >= 0 && nbands >= 0, GH_ERR_ARG, "multi_dotc: negative size");
    if (nbands == 0) return GH_OK;
    GH_REQUIRE(a_dev && b_dev && out_dev && work_dev, GH_ERR_ARG, "multi_dotc: null argument");
    if (nbands > kBandSlice) {
        // one grid row per band: slices, reusing the work buffer in stream order
        const int64_t row = n * dtype_size(dtype);
        GH_REQUIRE(row > 0 || n == 0, GH_ERR_ARG, "multi_dotc: bad dtype");
        for (int64_t x0 = 0; x0 < nbands; x0 += kBandSlice) {
            const int64_t nx = std::min<int64_t>(kBandSlice, nbands - x0);
            const int rc = gh_multi_dotc(dtype, nx, n, (const char*)a_dev + x0 * row,
                                         (const char*)b_dev + x0 * row, scale, out_dev + x0,
                                         work_dev, stream);
            if (rc != GH_OK) return rc;
        }
        return GH_OK;
    }
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(kDotChunks, (unsigned)nbands);
    switch (dtype) {
        case GH_F64: hipLaunchKernelGGL((dotc_partial_kernel<double>), grid, dim3(256), 0, st, (const double*)a_dev, (const double*)b_dev, n, work_dev); break;
        case GH_C128: hipLaunchKernelGGL((dotc_partial_kernel<cplx<double>>), grid, dim3(256), 0, st, (const cplx<double>*)a_dev, (const cplx<double>*)b_dev, n, work_dev); break;
        case GH_F32: hipLaunchKernelGGL((dotc_partial_kernel<float>), grid, dim3(256), 0, st, (const float*)a_dev, (const float*)b_dev, n, work_dev); break;
        case GH_C64: hipLaunchKernelGGL((dotc_partial_kernel<cplx<float>>), grid, dim3(256), 0, st, (const cplx<float>*)a_dev, (const cplx<float>*)b_dev, n, work_dev); break;
        default: set_error("multi_dotc: bad dtype"); return GH_ERR_ARG;
    }
    GH_CHECK_LAUNCH();
    hipLaunchKernelGGL(dotc_final_kernel, dim3((unsigned)nbands), dim3(kDotChunks), 0, st,
                       (const double*)work_dev, scale, out_dev);
    GH_CHECK_LAUNCH();
    return GH_OK;
}

int gh_multi_lincomb(int dtype, int64_t nbands, int64_t n, int k,
                     const void* const* vecs_dev, const double* coef_dev,
                     void* out_dev, void* stream) {
    GH_REQUIRE(n >= 0 && nbands >= 0, GH_ERR_ARG, "multi_lincomb: negative size");
    GH_REQUIRE(k >= 1 && k <= kLincombMax, GH_ERR_ARG,
               "multi_lincomb: k = %d out of range 1..%d", k, kLincombMax);
    if (nbands == 0 || n == 0) return GH_OK;  // empty arrays may carry null pointers
    GH_REQUIRE(vecs_dev && coef_dev && out_dev, GH_ERR_ARG, "multi_lincomb: null argument");
    for (int a = 0; a < k; a++)
        GH_REQUIRE(vecs_dev[a], GH_ERR_ARG, "multi_lincomb: null vector %d", a);
    if (nbands > kBandSlice) {
        // one grid row per band: pass larger band sets in slices
        const int64_t row = n * dtype_size(dtype);
        GH_REQUIRE(row > 0, GH_ERR_ARG, "multi_lincomb: bad dtype");
        const int64_t crow = (int64_t)k * ((dtype == GH_C128 || dtype == GH_C64) ? 2 : 1);
        const void* v[kLincombMax];
        for (int64_t x0 = 0; x0 < nbands; x0 += kBandSlice) {
            const int64_t nx = std::min<int64_t>(kBandSlice, nbands - x0);
            for (int a = 0; a < k; a++) v[a] = (const char*)vecs_dev[a] + x0 * row;
            const int rc = gh_multi_lincomb(dtype, nx, n, k, v, coef_dev + x0 * crow,
                                            (char*)out_dev + x0 * row, stream);
            if (rc != GH_OK) return rc;
        }
        return GH_OK;
    }
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
        case GH_F64: return lincomb_k<double, double>(k, nbands, n, vecs_dev, coef_dev, out_dev, st);
        case GH_C128: return lincomb_k<cplx<double>, double>(k, nbands, n, vecs_dev, coef_dev, out_dev, st);
        case GH_F32: return lincomb_k<float, float>(k, nbands, n, vecs_dev, coef_dev, out_dev, st);
        case GH_C64: return lincomb_k<cplx<float>, float>(k, nbands, n, vecs_dev, coef_dev, out_dev, st);
    }
    set_error("multi_lincomb: bad dtype");
    return GH_ERR_ARG;
}

int64_t gh_multi_gram_work(int k, int64_t nbands) {
    if (k < 1 || k > kGramMax || nbands < 0) return -1;
    return (int64_t)k * (k + 1) * kDotChunks * nbands;
}

int gh_multi_gram(int dtype, int64_t nbands, int64_t n, int k,
                  const void* const* vecs_dev, double scale, double* out_dev,
                  double* work_dev, void* stream) {
    GH_REQUIRE(n >= 0 && nbands >= 0, GH_ERR_ARG, "multi_gram: negative size");
    GH_REQUIRE(k >= 1 && k <= kGramMax, GH_ERR_ARG, "multi_gram: k = %d out of range 1..%d",
               k, kGramMax);
    if (nbands == 0) return GH_OK;  // empty arrays may carry null pointers
    GH_REQUIRE(vecs_dev && out_dev && work_dev, GH_ERR_ARG, "multi_gram: null argument");
    for (int a = 0; a < k; a++)
        GH_REQUIRE(vecs_dev[a], GH_ERR_ARG, "multi_gram: null vector %d", a);
    if (nbands > kBandSlice) {
        // one grid row per band: slices, reusing the work buffer in stream order
        const int64_t row = n * dtype_size(dtype);
        GH_REQUIRE(row > 0 || n == 0, GH_ERR_ARG, "multi_gram: bad dtype");
        const int64_t orow = (int64_t)k * k * ((dtype == GH_C128 || dtype == GH_C64) ? 2 : 1);
        const void* v[kGramMax];
        for (int64_t x0 = 0; x0 < nbands; x0 += kBandSlice) {
            const int64_t nx = std::min<int64_t>(kBandSlice, nbands - x0);
            for (int a = 0; a < k; a++) v[a] = (const char*)vecs_dev[a] + x0 * row;
            const int rc = gh_multi_gram(dtype, nx, n, k, v, scale, out_dev + x0 * orow,
                                         work_dev, stream);
            if (rc != GH_OK) return rc;
        }
        return GH_OK;
    }
    hipStream_t st = (hipStream_t)stream;
    switch (dtype) {
        case GH_F64: return gram_k<double>(k, nbands, n, vecs_dev, scale, 0, out_dev, work_dev, st);
        case GH_C128: return gram_k<cplx<double>>(k, nbands, n, vecs_dev, scale, 1, out_dev, work_dev, st);
        case GH_F32: return gram_k<float>(k, nbands, n, vecs_dev, scale, 0, out_dev, work_dev, st);
        case GH_C64: return gram_k<cplx<float>>(k, nbands, n, vecs_dev, scale, 1, out_dev, work_dev, st);
    }
    set_error("multi_gram: bad dtype");
    return GH_ERR_ARG;
}

int gh_multi_gram_noalloc(int dtype, int64_t nbands, int64_t n, int k,
                          const void* const* vecs_dev, double scale, double* out_dev,
                          void* stream) {
    GH_REQUIRE(k >= 1 && k <= kGramMax && nbands >= 0, GH_ERR_ARG,
               "multi_gram_noalloc: k = %d out of range 1..%d", k, kGramMax);
    const int64_t need = gh_multi_gram_work(k, nbands) * (int64_t)sizeof(double);
    GH_REQUIRE(need <= kNoAllocScratchBytes, GH_ERR_ARG,
               "multi_gram_noalloc: %lld bands x k = %d need %lld bytes of scratch, "
               "more than the fixed %lld",
               (long long)nbands, k, (long long)need, (long long)kNoAllocScratchBytes);
    double* work = nullptr;
    int rc = scratch_for_current_device(&work);
    if (rc != GH_OK) return rc;
    // the partial kernel writes every slot the final kernel reads, so the
    // scratch needs no clearing between calls
    return gh_multi_gram(dtype, nbands, n, k, vecs_dev, scale, out_dev, work, stream);
}

int gh_add_to_density(int dtype, int64_t nbands, int64_t n, const double* f_dev,
                      const void* psit_dev, double* nt_dev, void* stream) {
    GH_REQUIRE(n >= 0 && nbands >= 0, GH_ERR_ARG, "add_to_density: negative size");
    if (n == 0 || nbands == 0) return GH_OK;  // empty arrays may carry null pointers
    GH_REQUIRE(f_dev && psit_dev && nt_dev, GH_ERR_ARG, "add_to_density: null argument");
    hipStream_t st = (hipStream_t)stream;
    const int64_t nblk = std::min<int64_t>((n + 255) / 256, 16384);
    switch (dtype) {
        case GH_F64: hipLaunchKernelGGL((add_to_density_kernel<double>), dim3((unsigned)nblk), dim3(256), 0, st, f_dev, (const double*)psit_dev, nt_dev, nbands, n); break;
        case GH_C128: hipLaunchKernelGGL((add_to_density_kernel<cplx<double>>), dim3((unsigned)nblk), dim3(256), 0, st, f_dev, (const cplx<double>*)psit_dev, nt_dev, nbands, n); break;
        case GH_F32: hipLaunchKernelGGL((add_to_density_kernel<float>), dim3((unsigned)nblk), dim3(256), 0, st, f_dev, (const float*)psit_dev, nt_dev, nbands, n); break;
        case GH_C64: hipLaunchKernelGGL((add_to_density_kernel<cplx<float>>), dim3((unsigned)nblk), dim3(256), 0, st, f_dev, (const cplx<float>*)psit_dev, nt_dev, nbands, n); break;
        default: set_error("add_to_density: bad dtype"); return GH_ERR_ARG;
    }
    GH_CHECK_LAUNCH();
    return GH_OK;
}

int64_t gh_precond_scratch_bytes(const gh_stencil* kin0, const gh_stencil* kin1,
                                 const gh_stencil* kin2, int64_t nbands) {
    if (!kin0 || !kin1 || !kin2) return -1;
    int64_t n0 = stencil_points(kin0), n1 = stencil_points(kin1),
            n2 = stencil_points(kin2);
    return nbands * (n0 + 3 * n1 + 2 * n2) * dtype_size(stencil_dtype(kin0));
}

int gh_precond_apply(const gh_stencil* kin0, const gh_stencil* kin1,
                     const gh_stencil* kin2, const gh_transformer* rest0,
                     const gh_transformer* rest1,
                     const gh_transformer* intp2,
                     const gh_transformer* intp1, double step,
                     const void* R_dev, void* d0_dev, int64_t nbands,
                     const double* phase_cd, void* scratch_dev, void* stream) {
    GH_REQUIRE(kin0 && kin1 && kin2 && rest0 && rest1 && intp2 && intp1, GH_ERR_ARG,
               "precond_apply: null handle");
    GH_REQUIRE(nbands >= 0, GH_ERR_ARG, "precond_apply: nbands < 0");
    // empty arrays may carry null pointers
    GH_REQUIRE(nbands == 0 || (R_dev && d0_dev && scratch_dev), GH_ERR_ARG,
               "precond_apply: null argument");
    GH_REQUIRE(nbands == 0 || R_dev != d0_dev, GH_ERR_ARG,
               "precond_apply: R and d0 must differ");
    const int dt = stencil_dtype(kin0);
    GH_REQUIRE(stencil_dtype(kin1) == dt && stencil_dtype(kin2) == dt, GH_ERR_ARG,
               "precond_apply: dtype mismatch");
    const int64_t n0 = stencil_points(kin0), n1 = stencil_points(kin1),
                  n2 = stencil_points(kin2);
    GH_REQUIRE(transformer_in_points(rest0) == n0 && transformer_out_points(rest0) == n1 &&
                   transformer_in_points(rest1) == n1 && transformer_out_points(rest1) == n2 &&
                   transformer_in_points(intp2) == n2 && transformer_out_points(intp2) == n1 &&
                   transformer_in_points(intp1) == n1 && transformer_out_points(intp1) == n0,
               GH_ERR_SHAPE, "precond_apply: level sizes do not match");
    if (nbands == 0) return GH_OK;
    hipStream_t st = (hipStream_t)stream;
    const int vs = dtype_size(dt);
    char* base = (char*)scratch_dev;
    void* d0p = base;                 base += nbands * n0 * vs;
    void* r1 = base;                  base += nbands * n1 * vs;
    void* t1 = base;                  base += nbands * n1 * vs;
    void* u1 = base;                  base += nbands * n1 * vs;
    void* r2 = base;                  base += nbands * n2 * vs;
    void* t2 = base;
    const double s = step;
    const double* ph = phase_cd;
    int rc;
#define RUN(x) if ((rc = (x)) != GH_OK) return rc
    RUN(transformer_apply_scaled(rest0, R_dev, r1, nbands, ph, -1.0, st));
    RUN(stencil_apply_scaled(kin1, r1, t1, nbands, ph, nullptr, nullptr, 0.0, 4 * s, -1.0, st));
    RUN(transformer_apply_scaled(rest1, t1, r2, nbands, ph, 1.0, st));
    RUN(stencil_apply_scaled(kin2, r2, t2, nbands, ph, nullptr, nullptr, 0.0,
                             -256 * s * s, 32 * s, st));
    RUN(transformer_apply_scaled(intp2, t2, u1, nbands, ph, 1.0, st));
    RUN(axpbyz(dt, 4 * s, r1, -1.0, u1, t1, nbands * n1, st));
    RUN(stencil_apply_scaled(kin1, t1, u1, nbands, ph, nullptr, r1, 4 * s, -4 * s, 1.0, st));
    RUN(transformer_apply_scaled(intp1, u1, d0p, nbands, ph, -1.0, st));
    RUN(stencil_apply_scaled(kin0, d0p, d0_dev, nbands, ph, nullptr, R_dev, -s, s, -1.0, st));
#undef RUN
    return GH_OK;
}

}  // extern "C"
