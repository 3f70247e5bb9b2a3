// Shared definitions for the gfx950 finite-difference kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/gpaw_amd.h"

namespace gh {

// ---------------------------------------------------------------- errors --
void set_error(const char* fmt, ...);

#define GH_CHECK_HIP(expr)                                                  \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess) {                                             \
            ::gh::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,     \
                            hipGetErrorString(e_));                         \
            return GH_ERR_HIP;                                              \
        }                                                                   \
    } while (0)

#define GH_REQUIRE(cond, code, ...)                                         \
    do {                                                                    \
        if (!(cond)) {                                                      \
            ::gh::set_error(__VA_ARGS__);                                   \
            return (code);                                                  \
        }                                                                   \
    } while (0)

// Launch-error check after a kernel launch.
#define GH_CHECK_LAUNCH() GH_CHECK_HIP(hipGetLastError())

// ------------------------------------------------------------ value types --
// Real and complex element types.  Complex values are stored interleaved
// (re, im) exactly as numpy complex128 / complex64.
template <typename S>
struct cplx {
    S re, im;
};

template <typename V> struct scalar_of { using type = V; };
template <typename S> struct scalar_of<cplx<S>> { using type = S; };

template <typename V> struct is_complex { static constexpr bool value = false; };
template <typename S> struct is_complex<cplx<S>> { static constexpr bool value = true; };

template <typename V> __device__ __forceinline__ V vzero();
template <> __device__ __forceinline__ double vzero<double>() { return 0.0; }
template <> __device__ __forceinline__ float vzero<float>() { return 0.0f; }
template <> __device__ __forceinline__ cplx<double> vzero<cplx<double>>() { return {0.0, 0.0}; }
template <> __device__ __forceinline__ cplx<float> vzero<cplx<float>>() { return {0.0f, 0.0f}; }

// acc + c * a  (c real)
__device__ __forceinline__ double fma_rs(double c, double a, double acc) { return acc + c * a; }
__device__ __forceinline__ float fma_rs(float c, float a, float acc) { return acc + c * a; }
template <typename S>
__device__ __forceinline__ cplx<S> fma_rs(S c, cplx<S> a, cplx<S> acc) {
    return {acc.re + c * a.re, acc.im + c * a.im};
}

// a * p (p complex phase; identity for real types)
template <typename S>
__device__ __forceinline__ S mul_phase(S a, S pre, S pim) { return a; }
template <typename S>
__device__ __forceinline__ cplx<S> mul_phase(cplx<S> a, S pre, S pim) {
    return {a.re * pre - a.im * pim, a.re * pim + a.im * pre};
}

__device__ __forceinline__ double vadd(double a, double b) { return a + b; }
__device__ __forceinline__ float vadd(float a, float b) { return a + b; }
template <typename S>
__device__ __forceinline__ cplx<S> vadd(cplx<S> a, cplx<S> b) { return {a.re + b.re, a.im + b.im}; }

__device__ __forceinline__ double vscale(double a, double s) { return a * s; }
__device__ __forceinline__ float vscale(float a, float s) { return a * s; }
template <typename S>
__device__ __forceinline__ cplx<S> vscale(cplx<S> a, S s) { return {a.re * s, a.im * s}; }

// Streaming (non-temporal) store: the line is not kept in L2, for outputs
// that are written once and not read again by the same kernel.
__device__ __forceinline__ void store_nt(double* p, double v) { __builtin_nontemporal_store(v, p); }
__device__ __forceinline__ void store_nt(float* p, float v) { __builtin_nontemporal_store(v, p); }
template <typename S>
__device__ __forceinline__ void store_nt(cplx<S>* p, cplx<S> v) {
    typedef S pair_t __attribute__((ext_vector_type(2)));
    pair_t w = {v.re, v.im};
    __builtin_nontemporal_store(w, reinterpret_cast<pair_t*>(p));
}

// Host-side Bloch phases: phase_cd[c][d] as (re, im); identity when absent.
struct Phases {
    double re[3][2];
    double im[3][2];
};

inline Phases make_phases(const double* phase_cd) {
    Phases p;
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 2; d++) {
            p.re[c][d] = phase_cd ? phase_cd[(c * 2 + d) * 2 + 0] : 1.0;
            p.im[c][d] = phase_cd ? phase_cd[(c * 2 + d) * 2 + 1] : 0.0;
        }
    return p;
}

// Bands per launch of the kernels that put one band on each grid row
// (gridDim.y); the entry points pass larger band sets in slices of this size.
constexpr int64_t kBandSlice = 32768;

inline int dtype_size(int dtype) {
    switch (dtype) {
        case GH_F64: return 8;
        case GH_C128: return 16;
        case GH_F32: return 4;
        case GH_C64: return 8;
    }
    return 0;
}

// Number of XCDs on MI355X: consecutive workgroup ids are dealt round-robin
// to the XCDs, so neighbouring tiles are given ids that differ by a multiple
// of this to land on the same XCD (shared L2).
constexpr int kNumXCD = 8;

__device__ __forceinline__ int xcd_swizzle(int b, int nblocks) {
    // Map hardware block id b to a logical tile id such that logical tiles
    // [k*nb/8, (k+1)*nb/8) all run on XCD k.  Only valid when nblocks % 8 == 0.
    if (nblocks % kNumXCD) return b;
    int per = nblocks / kNumXCD;
    return (b % kNumXCD) * per + b / kNumXCD;
}

}  // namespace gh
