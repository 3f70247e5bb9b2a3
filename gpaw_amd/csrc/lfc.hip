// Localized functions (PAW projectors) on gfx950: LFC.integrate / LFC.add.
//
// Semantics follow the reference c/lfc.c: integrate (lfc.c:432-505)
//   c[x, M] = dv * sum_G a[x, G] A_M(G) [* phase_W]
// and add (lfc.c:818-885)
//   a[x, G] += sum_M c[x, M] A_M(G) [* conj(phase_W)]
// where the reference sweeps a sorted event list of sphere z-runs per band.
//
// Layout here (built once per geometry by the host):
//   * atom-major entries: for atom a, entries e in [entry_ptr[a],
//     entry_ptr[a+1]) carry the grid index G_e, the periodic-image index w_e
//     and the atom's nm_a function values A[e][0..nm_a) (contiguous);
//   * a point-major CSR over the union of sphere points: for each distinct G
//     (sorted), the list of entries that touch it.
// integrate runs one block per (atom, band tile) and reduces over the atom's
// entries in LDS -- a gathered skinny GEMM, deterministic; add runs one
// thread per distinct point and band tile and sums its entries in a fixed
// order -- no atomics, deterministic, every point written once.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "common.h"

namespace gh {

constexpr int kLfcBT = 4;      // bands per integrate block
constexpr int kLfcMC = 8;      // functions per integrate pass
constexpr int kLfcThreads = 256;

template <typename V, typename S>
__global__ void __launch_bounds__(kLfcThreads)
lfc_integrate_kernel(const V* __restrict__ a, V* __restrict__ c,
                     const int64_t* __restrict__ entry_ptr,
                     const int32_t* __restrict__ M_ptr,
                     const int64_t* __restrict__ A_off,
                     const int32_t* __restrict__ G_e,
                     const int32_t* __restrict__ w_e,
                     const S* __restrict__ A, const S* __restrict__ phase_w,
                     int64_t ngpts, int Mmax, int nbands, S dv) {
    __shared__ V red[kLfcThreads / 64][kLfcBT * kLfcMC];
    const int atom = blockIdx.x;
    const int x0 = blockIdx.y * kLfcBT;
    const int e0 = (int)entry_ptr[atom];
    const int e1 = (int)entry_ptr[atom + 1];
    const int M0 = M_ptr[atom];
    const int nm = M_ptr[atom + 1] - M0;
    const S* Aa = A + A_off[atom];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;

    for (int m0 = 0; m0 < nm; m0 += kLfcMC) {
        const int mc = min(kLfcMC, nm - m0);
        V acc[kLfcBT][kLfcMC];
#pragma unroll
        for (int b = 0; b < kLfcBT; b++)
#pragma unroll
            for (int m = 0; m < kLfcMC; m++) acc[b][m] = vzero<V>();
        for (int e = e0 + threadIdx.x; e < e1; e += kLfcThreads) {
            const int64_t G = G_e[e];
            V av[kLfcBT];
            S pre = 1, pim = 0;
            if (is_complex<V>::value && phase_w) {
                int w = w_e[e];
                if (w >= 0) { pre = phase_w[2 * w]; pim = phase_w[2 * w + 1]; }
            }
#pragma unroll
            for (int b = 0; b < kLfcBT; b++) {
                int x = x0 + b;
                av[b] = x < nbands ? mul_phase(a[(size_t)x * ngpts + G], pre, pim)
                                   : vzero<V>();
            }
            const S* Ae = Aa + (size_t)(e - e0) * nm + m0;
#pragma unroll
            for (int m = 0; m < kLfcMC; m++) {
                if (m < mc) {
                    S Am = Ae[m];
#pragma unroll
                    for (int b = 0; b < kLfcBT; b++) acc[b][m] = fma_rs(Am, av[b], acc[b][m]);
                }
            }
        }
        // wave reduction (64 lanes) then across waves through LDS
#pragma unroll
        for (int b = 0; b < kLfcBT; b++)
#pragma unroll
            for (int m = 0; m < kLfcMC; m++) {
                V v = acc[b][m];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    if constexpr (is_complex<V>::value) {
                        v.re += __shfl_xor(v.re, off);
                        v.im += __shfl_xor(v.im, off);
                    } else {
                        v += __shfl_xor(v, off);
                    }
                }
                if (lane == 0) red[wave][b * kLfcMC + m] = v;
            }
        __syncthreads();
        if (threadIdx.x < kLfcBT * kLfcMC) {
            int b = threadIdx.x / kLfcMC;
            int m = threadIdx.x % kLfcMC;
            V s = red[0][threadIdx.x];
            for (int w = 1; w < kLfcThreads / 64; w++) s = vadd(s, red[w][threadIdx.x]);
            int x = x0 + b;
            if (x < nbands && m < mc) c[(size_t)x * Mmax + M0 + m0 + m] = vscale(s, dv);
        }
        __syncthreads();
    }
}

constexpr int kAddBT = 4;

template <typename V, typename S>
__global__ void __launch_bounds__(256)
lfc_add_kernel(const V* __restrict__ c, V* __restrict__ a,
               const int32_t* __restrict__ uG, const int32_t* __restrict__ uptr,
               const int32_t* __restrict__ uref,
               const int32_t* __restrict__ atom_of_e,
               const int32_t* __restrict__ M_ptr,
               const int64_t* __restrict__ A_off,
               const int64_t* __restrict__ entry_ptr,
               const int32_t* __restrict__ w_e, const S* __restrict__ A,
               const S* __restrict__ phase_w, int nu, int64_t ngpts, int Mmax,
               int nbands) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const int x0 = blockIdx.y * kAddBT;
    V acc[kAddBT];
#pragma unroll
    for (int b = 0; b < kAddBT; b++) acc[b] = vzero<V>();
    for (int r = uptr[u]; r < uptr[u + 1]; r++) {
        const int e = uref[r];
        const int at = atom_of_e[e];
        const int M0 = M_ptr[at];
        const int nm = M_ptr[at + 1] - M0;
        const S* Ae = A + A_off[at] + (size_t)(e - entry_ptr[at]) * nm;
        V part[kAddBT];
#pragma unroll
        for (int b = 0; b < kAddBT; b++) part[b] = vzero<V>();
        for (int m = 0; m < nm; m++) {
            S Am = Ae[m];
#pragma unroll
            for (int b = 0; b < kAddBT; b++) {
                int x = x0 + b;
                if (x < nbands) part[b] = fma_rs(Am, c[(size_t)x * Mmax + M0 + m], part[b]);
            }
        }
        S pre = 1, pim = 0;
        if (is_complex<V>::value && phase_w) {
            int w = w_e[e];
            if (w >= 0) { pre = phase_w[2 * w]; pim = -phase_w[2 * w + 1]; }
        }
#pragma unroll
        for (int b = 0; b < kAddBT; b++) acc[b] = vadd(acc[b], mul_phase(part[b], pre, pim));
    }
    const int64_t G = uG[u];
#pragma unroll
    for (int b = 0; b < kAddBT; b++) {
        int x = x0 + b;
        if (x < nbands) {
            V* p = a + (size_t)x * ngpts + G;
            *p = vadd(*p, acc[b]);
        }
    }
}

// out[x, M0+i] = sum_j P[x, M0+j] dH[j][i]  (- eps[x] sum_j P dS[j][i])
template <typename V, typename S>
__global__ void lfc_dH_kernel(const V* __restrict__ P, const S* __restrict__ dH,
                              const int64_t* __restrict__ dH_off,
                              const S* __restrict__ eps, const S* __restrict__ dS,
                              const int32_t* __restrict__ M_ptr, V* __restrict__ out,
                              int natoms, int Mmax, int nbands) {
    const int x = blockIdx.y;
    const int M = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= nbands || M >= Mmax) return;
    // find the atom of M (M_ptr is sorted)
    int lo = 0, hi = natoms;
    while (hi - lo > 1) {
        int mid = (lo + hi) / 2;
        if (M_ptr[mid] <= M) lo = mid; else hi = mid;
    }
    const int M0 = M_ptr[lo];
    const int nm = M_ptr[lo + 1] - M0;
    const int i = M - M0;
    const S* H = dH + dH_off[lo];
    const V* Px = P + (size_t)x * Mmax + M0;
    V acc = vzero<V>();
    for (int j = 0; j < nm; j++) acc = fma_rs(H[j * nm + i], Px[j], acc);
    if (eps && dS) {
        const S* D = dS + dH_off[lo];
        V s = vzero<V>();
        for (int j = 0; j < nm; j++) s = fma_rs(D[j * nm + i], Px[j], s);
        acc = fma_rs(-eps[x], s, acc);
    }
    out[(size_t)x * Mmax + M] = acc;
}

}  // namespace gh

struct gh_lfc {
    int dtype;
    int natoms;
    int Mmax;
    int nu;
    int64_t ngpts;
    int64_t nentries;
    double dv;
    // device arrays
    int64_t* entry_ptr = nullptr;
    int32_t* M_ptr = nullptr;
    int64_t* A_off = nullptr;
    int64_t* dH_off = nullptr;
    int32_t* G_e = nullptr;
    int32_t* w_e = nullptr;
    int32_t* atom_of_e = nullptr;
    void* A = nullptr;         // double or float
    void* phase_w = nullptr;   // (re, im) pairs, double or float
    int32_t* uG = nullptr;
    int32_t* uptr = nullptr;
    int32_t* uref = nullptr;
};

using namespace gh;

namespace {

template <typename T>
int upload(T** dst, const T* src, size_t n) {
    *dst = nullptr;
    if (n == 0) return GH_OK;
    GH_CHECK_HIP(hipMalloc((void**)dst, n * sizeof(T)));
    GH_CHECK_HIP(hipMemcpy(*dst, src, n * sizeof(T), hipMemcpyHostToDevice));
    return GH_OK;
}

void free_lfc(gh_lfc* h) {
    if (!h) return;
    void* ptrs[] = {h->entry_ptr, h->M_ptr, h->A_off, h->dH_off, h->G_e, h->w_e,
                    h->atom_of_e, h->A, h->phase_w, h->uG, h->uptr, h->uref};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete h;
}

bool single(int dtype) { return dtype == GH_F32 || dtype == GH_C64; }
bool cplxd(int dtype) { return dtype == GH_C128 || dtype == GH_C64; }

}  // namespace

extern "C" {

int gh_lfc_create(gh_lfc** out, int64_t ngpts, int natoms,
                  const int64_t* entry_ptr, const int32_t* M_ptr,
                  const int32_t* G_e, const int32_t* w_e, const double* A,
                  int nw, const double* phase_w, double dv, int dtype) {
    GH_REQUIRE(out && entry_ptr && M_ptr && natoms >= 0, GH_ERR_ARG,
               "lfc_create: null argument");
    GH_REQUIRE(dtype >= GH_F64 && dtype <= GH_C64, GH_ERR_ARG, "lfc_create: bad dtype");
    GH_REQUIRE(ngpts > 0 && ngpts < (1ll << 31), GH_ERR_SHAPE, "lfc_create: bad ngpts");
    const int64_t ne = entry_ptr[natoms];
    GH_REQUIRE(ne == 0 || (G_e && A), GH_ERR_ARG, "lfc_create: missing entries");
    // host-side validation and derived tables
    std::vector<int64_t> A_off(natoms + 1, 0), dH_off(natoms + 1, 0);
    std::vector<int32_t> atom_of_e(ne);
    for (int a = 0; a < natoms; a++) {
        int nm = M_ptr[a + 1] - M_ptr[a];
        GH_REQUIRE(nm >= 0 && entry_ptr[a + 1] >= entry_ptr[a], GH_ERR_ARG,
                   "lfc_create: pointers not monotonic at atom %d", a);
        A_off[a + 1] = A_off[a] + (entry_ptr[a + 1] - entry_ptr[a]) * nm;
        dH_off[a + 1] = dH_off[a] + (int64_t)nm * nm;
        for (int64_t e = entry_ptr[a]; e < entry_ptr[a + 1]; e++) {
            GH_REQUIRE(G_e[e] >= 0 && G_e[e] < ngpts, GH_ERR_ARG,
                       "lfc_create: grid index out of range");
            GH_REQUIRE(!w_e || w_e[e] < nw, GH_ERR_ARG, "lfc_create: image index out of range");
            atom_of_e[e] = a;
        }
    }
    // point-major CSR over distinct G (entries in a fixed, sorted order)
    std::vector<int64_t> order(ne);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return G_e[x] < G_e[y]; });
    std::vector<int32_t> uG, uptr, uref;
    uref.reserve(ne);
    for (int64_t i = 0; i < ne; i++) {
        int64_t e = order[i];
        if (uG.empty() || uG.back() != G_e[e]) {
            uG.push_back(G_e[e]);
            uptr.push_back((int32_t)i);
        }
        uref.push_back((int32_t)e);
    }
    uptr.push_back((int32_t)ne);

    auto* h = new gh_lfc();
    h->dtype = dtype;
    h->natoms = natoms;
    h->Mmax = M_ptr[natoms];
    h->nu = (int)uG.size();
    h->ngpts = ngpts;
    h->nentries = ne;
    h->dv = dv;
    int rc;
#define UP(dst, src, n) if ((rc = upload(dst, src, n)) != GH_OK) { free_lfc(h); return rc; }
    UP(&h->entry_ptr, entry_ptr, (size_t)natoms + 1);
    UP(&h->M_ptr, M_ptr, (size_t)natoms + 1);
    UP(&h->A_off, A_off.data(), A_off.size());
    UP(&h->dH_off, dH_off.data(), dH_off.size());
    UP(&h->G_e, G_e, (size_t)ne);
    UP(&h->atom_of_e, atom_of_e.data(), (size_t)ne);
    UP(&h->uG, uG.data(), uG.size());
    UP(&h->uptr, uptr.data(), uptr.size());
    UP(&h->uref, uref.data(), uref.size());
    std::vector<int32_t> w(ne, -1);
    if (w_e) std::copy(w_e, w_e + ne, w.begin());
    UP(&h->w_e, w.data(), w.size());
    const size_t nA = (size_t)A_off[natoms];
    if (single(dtype)) {
        std::vector<float> Af(A, A + nA);
        float* d = nullptr;
        UP(&d, Af.data(), nA);
        h->A = d;
        if (phase_w && nw > 0) {
            std::vector<float> pf(phase_w, phase_w + 2 * nw);
            float* dp = nullptr;
            UP(&dp, pf.data(), pf.size());
            h->phase_w = dp;
        }
    } else {
        double* d = nullptr;
        UP(&d, A, nA);
        h->A = d;
        if (phase_w && nw > 0) {
            double* dp = nullptr;
            UP(&dp, phase_w, (size_t)2 * nw);
            h->phase_w = dp;
        }
    }
#undef UP
    *out = h;
    return GH_OK;
}

int gh_lfc_destroy(gh_lfc* h) {
    free_lfc(h);
    return GH_OK;
}

int gh_lfc_nfunctions(const gh_lfc* h) { return h ? h->Mmax : 0; }

int gh_lfc_integrate(const gh_lfc* h, const void* a_dev, void* c_dev,
                     int64_t nbands, void* stream) {
    GH_REQUIRE(h && a_dev && c_dev, GH_ERR_ARG, "lfc_integrate: null argument");
    if (nbands == 0 || h->natoms == 0) return GH_OK;
    // every (band, M) is written by the block of its atom, also for atoms
    // without grid points, so no clearing pass is needed
    hipStream_t st = (hipStream_t)stream;
    dim3 grid(h->natoms, (unsigned)((nbands + kLfcBT - 1) / kLfcBT));
    const bool ph = cplxd(h->dtype) && h->phase_w;
#define ARGS(V, S) (const V*)a_dev, (V*)c_dev, h->entry_ptr, h->M_ptr, h->A_off, h->G_e, h->w_e, \
                   (const S*)h->A, ph ? (const S*)h->phase_w : nullptr, h->ngpts, h->Mmax, (int)nbands, (S)h->dv
    switch (h->dtype) {
        case GH_F64: hipLaunchKernelGGL((lfc_integrate_kernel<double, double>), grid, dim3(kLfcThreads), 0, st, ARGS(double, double)); break;
        case GH_C128: hipLaunchKernelGGL((lfc_integrate_kernel<cplx<double>, double>), grid, dim3(kLfcThreads), 0, st, ARGS(cplx<double>, double)); break;
        case GH_F32: hipLaunchKernelGGL((lfc_integrate_kernel<float, float>), grid, dim3(kLfcThreads), 0, st, ARGS(float, float)); break;
        case GH_C64: hipLaunchKernelGGL((lfc_integrate_kernel<cplx<float>, float>), grid, dim3(kLfcThreads), 0, st, ARGS(cplx<float>, float)); break;
    }
#undef ARGS
    GH_CHECK_LAUNCH();
    return GH_OK;
}

int gh_lfc_add(const gh_lfc* h, const void* c_dev, void* a_dev,
               int64_t nbands, void* stream) {
    GH_REQUIRE(h && a_dev && c_dev, GH_ERR_ARG, "lfc_add: null argument");
    if (nbands == 0 || h->nu == 0) return GH_OK;
    hipStream_t st = (hipStream_t)stream;
    dim3 grid((h->nu + 255) / 256, (unsigned)((nbands + kAddBT - 1) / kAddBT));
    const bool ph = cplxd(h->dtype) && h->phase_w;
#define ARGS(V, S) (const V*)c_dev, (V*)a_dev, h->uG, h->uptr, h->uref, h->atom_of_e, h->M_ptr, \
                   h->A_off, h->entry_ptr, h->w_e, (const S*)h->A, ph ? (const S*)h->phase_w : nullptr, \
                   h->nu, h->ngpts, h->Mmax, (int)nbands
    switch (h->dtype) {
        case GH_F64: hipLaunchKernelGGL((lfc_add_kernel<double, double>), grid, dim3(256), 0, st, ARGS(double, double)); break;
        case GH_C128: hipLaunchKernelGGL((lfc_add_kernel<cplx<double>, double>), grid, dim3(256), 0, st, ARGS(cplx<double>, double)); break;
        case GH_F32: hipLaunchKernelGGL((lfc_add_kernel<float, float>), grid, dim3(256), 0, st, ARGS(float, float)); break;
        case GH_C64: hipLaunchKernelGGL((lfc_add_kernel<cplx<float>, float>), grid, dim3(256), 0, st, ARGS(cplx<float>, float)); break;
    }
#undef ARGS
    GH_CHECK_LAUNCH();
    return GH_OK;
}

int gh_lfc_apply_dH(const gh_lfc* h, const void* P_dev, const double* dH_dev,
                    const double* eps_dev, const double* dS_dev,
                    void* out_dev, int64_t nbands, void* stream) {
    GH_REQUIRE(h && P_dev && dH_dev && out_dev, GH_ERR_ARG, "lfc_apply_dH: null argument");
    GH_REQUIRE(!single(h->dtype), GH_ERR_UNSUPPORTED,
               "lfc_apply_dH: single-precision coefficients are converted by the caller");
    GH_REQUIRE(P_dev != out_dev, GH_ERR_ARG, "lfc_apply_dH: P and out must differ");
    if (nbands == 0 || h->Mmax == 0) return GH_OK;
    hipStream_t st = (hipStream_t)stream;
    dim3 grid((h->Mmax + 127) / 128, (unsigned)nbands);
    if (h->dtype == GH_F64)
        hipLaunchKernelGGL((lfc_dH_kernel<double, double>), grid, dim3(128), 0, st,
                           (const double*)P_dev, dH_dev, h->dH_off, eps_dev, dS_dev,
                           h->M_ptr, (double*)out_dev, h->natoms, h->Mmax, (int)nbands);
    else
        hipLaunchKernelGGL((lfc_dH_kernel<cplx<double>, double>), grid, dim3(128), 0, st,
                           (const cplx<double>*)P_dev, dH_dev, h->dH_off, eps_dev, dS_dev,
                           h->M_ptr, (cplx<double>*)out_dev, h->natoms, h->Mmax, (int)nbands);
    GH_CHECK_LAUNCH();
    return GH_OK;
}

}  // extern "C"

// ------------------------------------------------- per-geometry setup --
// Host code: radial spline tables and the values of the localized functions
// on the grid points of a box.  Runs once per geometry.

namespace {

// Real solid harmonics r^l Y_lm times f for l <= 2, m order of the
// reference (c/bmgs/spherical_harmonics.h): l=1 (y, z, x); l=2 (xy, yz,
// 3z^2 - r^2, xz, x^2 - y^2).
void solid_harmonics(int l, double f, double x, double y, double z,
                     double r2, double* p) {
    switch (l) {
        case 0:
            p[0] = f * 0.28209479177387814;
            break;
        case 1:
            p[0] = f * 0.48860251190291992 * y;
            p[1] = f * 0.48860251190291992 * z;
            p[2] = f * 0.48860251190291992 * x;
            break;
        case 2:
            p[0] = f * 1.0925484305920792 * x * y;
            p[1] = f * 1.0925484305920792 * y * z;
            p[2] = f * 0.31539156525252005 * (-r2 + 3 * z * z);
            p[3] = f * 1.0925484305920792 * x * z;
            p[4] = f * 0.54627421529603959 * (-y * y + x * x);
            break;
    }
}

}  // namespace

extern "C" {

// Cubic spline with four coefficients per bin and a zero bin appended:
// reference c/bmgs/spline.c:11-47 (second derivatives from the tridiagonal
// recurrence with f2[0] = -1/2), evaluated as in :50-58.
int gh_spline_coefs(int nbins, double dr, const double* f, double* out) {
    GH_REQUIRE(nbins >= 1 && dr > 0 && f && out, GH_ERR_ARG, "spline_coefs: bad argument");
    const double c = 3.0 / (dr * dr);
    std::vector<double> f2(nbins + 1), u(nbins);
    f2[0] = -0.5;
    u[0] = (f[1] - f[0]) * c;
    for (int b = 1; b < nbins; b++) {
        double p = 0.5 * f2[b - 1] + 2.0;
        f2[b] = -0.5 / p;
        u[b] = ((f[b + 1] - 2.0 * f[b] + f[b - 1]) * c - 0.5 * u[b - 1]) / p;
    }
    f2[nbins] = (f[nbins - 1] * c - 0.5 * u[nbins - 1]) / (0.5 * f2[nbins - 1] + 1.0);
    for (int b = nbins - 1; b >= 0; b--) f2[b] = f2[b] * f2[b + 1] + u[b];
    for (int b = 0; b < nbins; b++) {
        out[4 * b + 0] = f[b];
        out[4 * b + 1] = (f[b + 1] - f[b]) / dr - (f2[b] / 3 + f2[b + 1] / 6) * dr;
        out[4 * b + 2] = 0.5 * f2[b];
        out[4 * b + 3] = (f2[b + 1] - f2[b]) / (6 * dr);
    }
    for (int k = 0; k < 4; k++) out[4 * nbins + k] = 0.0;
    return GH_OK;
}

// Values of spline(r) * r^l Y_lm on the points of box [beg_c, end_c) with
// r < rcut, in C order of the box; G_out gets the flattened index of each
// point relative to the grid corner gdcorner_c (reference lfc.c:887-977).
// Returns the number of points written, or a negative error code.
int64_t gh_spline_to_grid(int l, int nbins, double dr, const double* coefs,
                          const int64_t* beg_c, const int64_t* end_c,
                          const double* pos_v, const double* h_cv,
                          const int64_t* n_c, const int64_t* gdcorner_c,
                          double* A_out, int32_t* G_out) {
    GH_REQUIRE(l >= 0 && l <= 2, GH_ERR_UNSUPPORTED, "spline_to_grid: l=%d > 2", l);
    GH_REQUIRE(coefs && beg_c && end_c && pos_v && h_cv && n_c && gdcorner_c &&
                   A_out && G_out,
               GH_ERR_ARG, "spline_to_grid: null argument");
    const int nm = 2 * l + 1;
    const double rcut = dr * nbins;
    int64_t ng = 0;
    for (int64_t g0 = beg_c[0]; g0 < end_c[0]; g0++)
        for (int64_t g1 = beg_c[1]; g1 < end_c[1]; g1++)
            for (int64_t g2 = beg_c[2]; g2 < end_c[2]; g2++) {
                double x = h_cv[0] * g0 + h_cv[3] * g1 + h_cv[6] * g2 - pos_v[0];
                double y = h_cv[1] * g0 + h_cv[4] * g1 + h_cv[7] * g2 - pos_v[1];
                double z = h_cv[2] * g0 + h_cv[5] * g1 + h_cv[8] * g2 - pos_v[2];
                double r2 = x * x + y * y + z * z;
                double r = std::sqrt(r2);
                if (r >= rcut) continue;
                int b = (int)(r / dr);
                double f = 0.0;
                if (b < nbins) {
                    double u = r - b * dr;
                    const double* s = coefs + 4 * b;
                    f = s[0] + u * (s[1] + u * (s[2] + u * s[3]));
                }
                solid_harmonics(l, f, x, y, z, r2, A_out + ng * nm);
                G_out[ng] = (int32_t)(((g0 - gdcorner_c[0]) * n_c[1] +
                                       (g1 - gdcorner_c[1])) * n_c[2] +
                                      (g2 - gdcorner_c[2]));
                ng++;
            }
    return ng;
}

}  // extern "C"
