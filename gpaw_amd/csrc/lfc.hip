// Localized functions (PAW projectors) on gfx950: LFC.integrate / LFC.add.
//
// Semantics follow the reference c/lfc.c: integrate (lfc.c:432-505)
//   c[x, M] = dv * sum_G a[x, G] A_M(G) [* phase_W]
// and add (lfc.c:818-885)
//   a[x, G] += sum_M c[x, M] A_M(G) [* conj(phase_W)]
// where the reference sweeps a sorted event list of sphere z-runs per band.
//
// Layout here (built once per geometry by the host): atom-major entries --
// for atom a, entries e in [entry_ptr[a], entry_ptr[a+1]) carry the grid
// index G_e, the periodic-image index w_e and the atom's nm_a function
// values A[e][0..nm_a) (contiguous rows).
//
// integrate (a gathered skinny GEMM P = psi_a A_a): one block per (atom,
// band tile); inside a wavefront lane l handles function m = l % nm of entry
// group l / nm, so each lane keeps one accumulator per band and a wave reads
// whole A rows with one coalesced load; partial sums are reduced through LDS
// in a fixed order (deterministic).
// add (the transposed product scattered back): add units (an atom's
// entries, split into layers where its sphere overlaps its own periodic
// image) that share a grid point get different colours (greedy colouring on
// the host); one launch per colour, one wavefront per (unit, entry chunk,
// band tile).  Units of one colour touch disjoint points, so every
// read-modify-write is race free without atomics, and the order of the
// contributions to a point is fixed (deterministic).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>
#include <numeric>
#include <vector>

#include "common.h"

namespace gh {

constexpr int kLfcMaxNm = 64;   // functions per atom supported
constexpr int kIntChunk = 256;  // entries staged per integrate step
constexpr int kAddBT = 32;      // bands per add wavefront (at most)
constexpr int kAddGrp = 32;     // bands loaded together in add
constexpr int kAddChunk = 64;   // entries per add wavefront (one per lane)

// FP64 counterpart of a storage type: single-precision data are summed in
// FP64 (the projections are reductions over thousands of sphere points)
template <typename V> struct Wide { using type = V; };
template <> struct Wide<float> { using type = double; };
template <> struct Wide<cplx<float>> { using type = cplx<double>; };
__device__ __forceinline__ double widen(double a) { return a; }
__device__ __forceinline__ double widen(float a) { return (double)a; }
template <typename S>
__device__ __forceinline__ cplx<double> widen(cplx<S> a) {
    return {(double)a.re, (double)a.im};
}
__device__ __forceinline__ void narrow(double a, double* o) { *o = a; }
__device__ __forceinline__ void narrow(double a, float* o) { *o = (float)a; }
template <typename S>
__device__ __forceinline__ void narrow(cplx<double> a, cplx<S>* o) {
    *o = {(S)a.re, (S)a.im};
}

template <typename V> struct IntBT { static constexpr int value = 16; };
template <> struct IntBT<cplx<double>> { static constexpr int value = 8; };
// complex64 sums in complex128: 8 bands keep the LDS of T and the FP64
// reduction buffer at 48 KB
template <> struct IntBT<cplx<float>> { static constexpr int value = 8; };

// integrate: one 256-thread block per (atom, band tile).  Per chunk of 256
// entries every thread gathers a[x, G_e] for its entry and the tile's bands
// (coalesced along the sphere's z-runs; the next chunk is prefetched into
// registers) and stages them in LDS; then lane (g, m) = (entry group,
// function) accumulates sum_e A[e, m] a[x, e] for all bands of the tile,
// reading whole A rows with coalesced loads.  Group partial sums are
// reduced through LDS in a fixed order (deterministic).
template <typename V, typename S>
__global__ void __launch_bounds__(256)
lfc_integrate_kernel(const V* __restrict__ a, V* __restrict__ c,
                     const int64_t* __restrict__ entry_ptr,
                     const int32_t* __restrict__ M_ptr,
                     const int64_t* __restrict__ A_off,
                     const int32_t* __restrict__ G_e,
                     const int32_t* __restrict__ w_e,
                     const S* __restrict__ A, const S* __restrict__ phase_w,
                     int64_t ngpts, int Mmax, int nbands, S dv) {
    constexpr int BT = IntBT<V>::value;
    using W = typename Wide<V>::type;
    constexpr bool kWide = sizeof(W) > sizeof(V);
    __shared__ V T[BT][kIntChunk];
    // reduction buffer: T itself when the sums have the storage type
    __shared__ W Tw[kWide ? BT * kIntChunk : 1];
    const int atom = blockIdx.x;
    const int x0 = blockIdx.y * BT;
    const int M0 = M_ptr[atom];
    const int nm = M_ptr[atom + 1] - M0;
    if (nm == 0) return;
    const int64_t e0 = entry_ptr[atom];
    const int64_t e1 = entry_ptr[atom + 1];
    const S* __restrict__ Aa = A + A_off[atom];
    const int t = threadIdx.x;
    const int nb = min(BT, nbands - x0);
    const int ngrp = kIntChunk / nm;
    const int grp = t / nm;
    const int m = t - grp * nm;
    const bool live = grp < ngrp;

    V pf[BT];
    auto gather = [&](int64_t ebase) {
        const int64_t e = ebase + t;
        if (e < e1) {
            const int64_t G = G_e[e];
            S pre = 1, pim = 0;
            if (is_complex<V>::value && phase_w) {
                int w = w_e[e];
                if (w >= 0) { pre = phase_w[2 * w]; pim = phase_w[2 * w + 1]; }
            }
#pragma unroll
            for (int b = 0; b < BT; b++)
                pf[b] = b < nb ? mul_phase(a[(size_t)(x0 + b) * ngpts + G], pre, pim)
                               : vzero<V>();
        } else {
#pragma unroll
            for (int b = 0; b < BT; b++) pf[b] = vzero<V>();
        }
    };

    W acc[BT];
#pragma unroll
    for (int b = 0; b < BT; b++) acc[b] = vzero<W>();
    gather(e0);
    for (int64_t eb = e0; eb < e1; eb += kIntChunk) {
        __syncthreads();
#pragma unroll
        for (int b = 0; b < BT; b++) T[b][t] = pf[b];
        if (eb + kIntChunk < e1) gather(eb + kIntChunk);
        __syncthreads();
        const int n = (int)min<int64_t>(kIntChunk, e1 - eb);
        if (live) {
            for (int el = grp; el < n; el += ngrp) {
                const S Am = Aa[(eb - e0 + el) * nm + m];
#pragma unroll
                for (int b = 0; b < BT; b++)
                    acc[b] = fma_rs(widen(Am), widen(T[b][el]), acc[b]);
            }
        }
    }
    // reduce over entry groups: R[g][b][m] reuses T
    __syncthreads();
    W* Rb = kWide ? Tw : reinterpret_cast<W*>(&T[0][0]);
    if (live) {
#pragma unroll
        for (int b = 0; b < BT; b++) Rb[(grp * BT + b) * nm + m] = acc[b];
    }
    __syncthreads();
    for (int u = t; u < BT * nm; u += blockDim.x) {
        const int b = u / nm;
        const int mm = u - b * nm;
        W sum = vzero<W>();
        for (int g = 0; g < ngrp; g++) sum = vadd(sum, Rb[(g * BT + b) * nm + mm]);
        if (b < nb) narrow(vscale(sum, (double)dv), c + (size_t)(x0 + b) * Mmax + M0 + mm);
    }
}

// add: one wavefront per (unit of the colour, chunk of 64 entries, band
// tile of bt <= kAddBT bands).  A unit is a contiguous entry range of one atom in
// which every grid point occurs once (an atom whose sphere overlaps its own
// periodic image has several units).  lane = entry: the lane keeps its
// entry's grid index, phase and function values (registers when the count
// is known at compile time) and sweeps the tile's bands, kAddGrp at a time
// (all loads of a group issued before its stores); the coefficients
// c[x, M_a + m] are wave-uniform (scalar loads).
template <typename V, typename S, int NM>
__global__ void __launch_bounds__(64)
lfc_add_kernel(const V* __restrict__ c, V* __restrict__ a,
               const int32_t* __restrict__ units, int nchunks,
               const int32_t* __restrict__ unit_atom,
               const int64_t* __restrict__ unit_beg,
               const int64_t* __restrict__ unit_end,
               const int64_t* __restrict__ entry_ptr,
               const int32_t* __restrict__ M_ptr,
               const int64_t* __restrict__ A_off,
               const int32_t* __restrict__ G_e,
               const int32_t* __restrict__ w_e, const S* __restrict__ A,
               const S* __restrict__ phase_w, int64_t ngpts, int Mmax,
               int nbands, int bt) {
    const int unit = units[blockIdx.x / nchunks];
    const int chunk = blockIdx.x % nchunks;
    const int atom = unit_atom[unit];
    const int x0 = blockIdx.y * bt;
    const int64_t e0 = entry_ptr[atom];
    const int64_t eb = unit_beg[unit] + (int64_t)chunk * kAddChunk;
    const int64_t e1 = min(unit_end[unit], eb + kAddChunk);
    if (eb >= e1) return;
    const int M0 = M_ptr[atom];
    const int nm = NM > 0 ? NM : M_ptr[atom + 1] - M0;
    const int nb = min(bt, nbands - x0);
    const int64_t e = eb + threadIdx.x;
    if (e >= e1) return;
    const int64_t G = G_e[e];
    const S* __restrict__ Ae = A + A_off[atom] + (e - e0) * nm;
    S Ar[NM > 0 ? NM : 1];
    if constexpr (NM > 0) {
#pragma unroll
        for (int m = 0; m < NM; m++) Ar[m] = Ae[m];
    }
    S pre = 1, pim = 0;
    if (is_complex<V>::value && phase_w) {
        int w = w_e[e];
        if (w >= 0) { pre = phase_w[2 * w]; pim = -phase_w[2 * w + 1]; }
    }
    V* __restrict__ ap = a + (size_t)x0 * ngpts + G;
    const V* __restrict__ cp = c + (size_t)x0 * Mmax + M0;
    for (int b0 = 0; b0 < nb; b0 += kAddGrp) {
        V old[kAddGrp];
#pragma unroll
        for (int g = 0; g < kAddGrp; g++)
            if (b0 + g < nb) old[g] = ap[(size_t)(b0 + g) * ngpts];
#pragma unroll
        for (int g = 0; g < kAddGrp; g++) {
            if (b0 + g < nb) {
                const V* __restrict__ cb = cp + (size_t)(b0 + g) * Mmax;
                V s = vzero<V>();
                if constexpr (NM > 0) {
#pragma unroll
                    for (int m = 0; m < NM; m++) s = fma_rs(Ar[m], cb[m], s);
                } else {
                    for (int m = 0; m < nm; m++) s = fma_rs(Ae[m], cb[m], s);
                }
                ap[(size_t)(b0 + g) * ngpts] = vadd(old[g], mul_phase(s, pre, pim));
            }
        }
    }
}

// add, point-major: one read-modify-write per grid point of the spheres'
// union (the slots, in grid order), all atoms in one launch:
//   a[x, G_u] += sum_{entries e at u} sum_m c[x, M_e + m] A_e[m] conj(phase_e)
// One wavefront per (chunk of 64 slots, band tile of BT bands); lane = slot,
// so the loads and stores of a band run along the spheres' z-runs.  The
// chunk's distinct atoms come from a host table; per atom the wave stages
// the atom's coefficients c[x0 .. x0 + BT, M_a ..] in LDS (one coalesced
// load), each lane folds its entries of that atom (several when a sphere
// overlaps its own periodic image) into one coefficient row, and the band
// sums read the staged coefficients as LDS broadcasts.  Every point is
// written by one lane: no colouring, no atomics, and the order of the
// contributions (atoms ascending) is fixed.
// UNIF: every atom has MAXNM functions.
template <typename V, typename S, int MAXNM, int BT, bool UNIF>
__global__ void __launch_bounds__(64)
lfc_add_points_kernel(const V* __restrict__ c, V* __restrict__ a,
                      const int32_t* __restrict__ chunk_ptr,
                      const int32_t* __restrict__ chunk_atom,
                      const int32_t* __restrict__ slot_ptr,
                      const int32_t* __restrict__ slot_G,
                      const int32_t* __restrict__ se_atom,
                      const int64_t* __restrict__ se_A,
                      const int32_t* __restrict__ se_w,
                      const int32_t* __restrict__ M_ptr, const S* __restrict__ A,
                      const S* __restrict__ phase_w, int64_t nslots, int64_t ngpts,
                      int Mmax, int nbands) {
    constexpr bool CPLX = is_complex<V>::value;
    __shared__ V cs[BT * MAXNM];
    const int lane = threadIdx.x;
    const int chunk = blockIdx.x;
    const int64_t u = (int64_t)chunk * 64 + lane;
    const bool live = u < nslots;
    const int x0 = blockIdx.y * BT;
    const int nb = min(BT, nbands - x0);
    int p0 = 0, p1 = 0;
    int64_t G = 0;
    if (live) {
        p0 = slot_ptr[u];
        p1 = slot_ptr[u + 1];
        G = slot_G[u];
    }
    V acc[BT];
#pragma unroll
    for (int b = 0; b < BT; b++) acc[b] = vzero<V>();
    for (int k = chunk_ptr[chunk]; k < chunk_ptr[chunk + 1]; k++) {
        const int at = chunk_atom[k];
        const int M0 = M_ptr[at];
        const int nm = UNIF ? MAXNM : M_ptr[at + 1] - M0;
        // stage c[x0 + b, M0 + m] (b < nb, m < nm) at cs[b * MAXNM + m];
        // one wave: its LDS writes and reads stay in program order
        for (int i = lane; i < nb * MAXNM; i += 64) {
            const int b = i / MAXNM, m = i - b * MAXNM;
            if (UNIF || m < nm) cs[i] = c[(size_t)(x0 + b) * Mmax + M0 + m];
        }
        // this lane's coefficient row for atom `at` (zero without an entry;
        // loads at m >= nm repeat the last function and weigh zero)
        S are[MAXNM], aim[CPLX ? MAXNM : 1];
#pragma unroll
        for (int m = 0; m < MAXNM; m++) {
            are[m] = 0;
            if constexpr (CPLX) aim[m] = 0;
        }
        for (int p = p0; p < p1; p++) {
            if (se_atom[p] != at) continue;
            const S* __restrict__ Ae = A + se_A[p];
            S pr = 1, pi = 0;
            if (CPLX && phase_w) {
                const int w = se_w[p];
                if (w >= 0) { pr = phase_w[2 * w]; pi = -phase_w[2 * w + 1]; }
            }
            S v[MAXNM];
#pragma unroll
            for (int m = 0; m < MAXNM; m++) v[m] = Ae[UNIF || m < nm ? m : nm - 1];
#pragma unroll
            for (int m = 0; m < MAXNM; m++) {
                const S vm = UNIF || m < nm ? v[m] : (S)0;
                are[m] += vm * pr;
                if constexpr (CPLX) aim[m] += vm * pi;
            }
        }
#pragma unroll
        for (int b = 0; b < BT; b++) {
            if (b < nb) {
#pragma unroll
                for (int m = 0; m < MAXNM; m++) {
                    if (!UNIF && m >= nm) break;
                    const V cv = cs[b * MAXNM + m];
                    if constexpr (CPLX) {
                        acc[b].re += are[m] * cv.re - aim[m] * cv.im;
                        acc[b].im += are[m] * cv.im + aim[m] * cv.re;
                    } else {
                        acc[b] += are[m] * cv;
                    }
                }
            }
        }
    }
    if (live) {
        // all loads of the tile issued before its stores
        V* __restrict__ ap = a + (size_t)x0 * ngpts + G;
        V old[BT];
#pragma unroll
        for (int b = 0; b < BT; b++)
            if (b < nb) old[b] = ap[(size_t)b * ngpts];
#pragma unroll
        for (int b = 0; b < BT; b++)
            if (b < nb) ap[(size_t)b * ngpts] = vadd(old[b], acc[b]);
    }
}

// out[x, M0+i] = sum_j P[x, M0+j] dH[j][i]  (- eps[x] sum_j P dS[j][i]),
// summed in FP64 for every storage type V (W: its FP64 counterpart)
template <typename V, typename W>
__global__ void lfc_dH_kernel(const V* __restrict__ P, const double* __restrict__ dH,
                              const int64_t* __restrict__ dH_off,
                              const double* __restrict__ eps, const double* __restrict__ dS,
                              const int32_t* __restrict__ M_ptr, V* __restrict__ out,
                              int natoms, int Mmax, int nbands) {
    using S = double;
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
    W acc = vzero<W>();
    for (int j = 0; j < nm; j++) acc = fma_rs(H[j * nm + i], widen(Px[j]), acc);
    if (eps && dS) {
        const S* D = dS + dH_off[lo];
        W s = vzero<W>();
        for (int j = 0; j < nm; j++) s = fma_rs(D[j * nm + i], widen(Px[j]), s);
        acc = fma_rs(-eps[x], s, acc);
    }
    narrow(acc, out + (size_t)x * Mmax + M);
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
    void* A = nullptr;         // double or float
    void* phase_w = nullptr;   // (re, im) pairs, double or float: nq tables of nw
    int nq = 0;
    int nw = 0;
    int max_nm = 0;                     // most functions on one atom
    int uniform_nm = -1;                // functions per atom if all equal
    int64_t max_unit = 0;               // most entries in one add unit
    int32_t* unit_atom = nullptr;       // add unit -> atom
    int64_t* unit_beg = nullptr;        // add unit -> entry range
    int64_t* unit_end = nullptr;
    int32_t* color_units = nullptr;     // add units grouped by colour
    std::vector<int> color_ptr;         // host: colour c -> [ptr[c], ptr[c+1])
    // point-major add: the slots (grid points of the spheres' union in grid
    // order), their entries and each 64-slot chunk's distinct atoms
    int64_t nslots = 0;
    int nchunks = 0;
    int32_t* slot_ptr = nullptr;        // slot -> its entries
    int32_t* slot_G = nullptr;
    int32_t* se_atom = nullptr;
    int64_t* se_A = nullptr;
    int32_t* se_w = nullptr;
    int32_t* chunk_ptr = nullptr;       // chunk -> distinct atoms, ascending
    int32_t* chunk_atom = nullptr;
    bool points_add = true;             // GPAW_AMD_LFC_ADD=colour: colour launches
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
                    h->A, h->phase_w, h->unit_atom, h->unit_beg, h->unit_end,
                    h->color_units, h->chunk_ptr, h->chunk_atom, h->slot_ptr,
                    h->slot_G, h->se_atom, h->se_A, h->se_w};
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
                  int nq, int nw, const double* phase_qw, double dv, int dtype) {
    GH_REQUIRE(out && entry_ptr && M_ptr && natoms >= 0, GH_ERR_ARG,
               "lfc_create: null argument");
    GH_REQUIRE(dtype >= GH_F64 && dtype <= GH_C64, GH_ERR_ARG, "lfc_create: bad dtype");
    GH_REQUIRE(ngpts > 0 && ngpts < (1ll << 31), GH_ERR_SHAPE, "lfc_create: bad ngpts");
    GH_REQUIRE(nw >= 0 && (!phase_qw || nq >= 1), GH_ERR_ARG,
               "lfc_create: phase tables need nq >= 1 and nw >= 0");
    const int64_t ne = entry_ptr[natoms];
    GH_REQUIRE(ne == 0 || (G_e && A), GH_ERR_ARG, "lfc_create: missing entries");
    // host-side validation and derived tables
    std::vector<int64_t> A_off(natoms + 1, 0), dH_off(natoms + 1, 0);
    for (int a = 0; a < natoms; a++) {
        int nm = M_ptr[a + 1] - M_ptr[a];
        GH_REQUIRE(nm >= 0 && entry_ptr[a + 1] >= entry_ptr[a], GH_ERR_ARG,
                   "lfc_create: pointers not monotonic at atom %d", a);
        GH_REQUIRE(nm <= kLfcMaxNm, GH_ERR_UNSUPPORTED,
                   "lfc_create: %d functions on atom %d (max %d)", nm, a, kLfcMaxNm);
        A_off[a + 1] = A_off[a] + (entry_ptr[a + 1] - entry_ptr[a]) * nm;
        dH_off[a + 1] = dH_off[a] + (int64_t)nm * nm;
        for (int64_t e = entry_ptr[a]; e < entry_ptr[a + 1]; e++) {
            GH_REQUIRE(G_e[e] >= 0 && G_e[e] < ngpts, GH_ERR_ARG,
                       "lfc_create: grid index out of range");
            GH_REQUIRE(!w_e || w_e[e] < nw, GH_ERR_ARG, "lfc_create: image index out of range");
            GH_REQUIRE(!w_e || w_e[e] >= -1, GH_ERR_ARG, "lfc_create: bad image index");
        }
    }
    // Order each atom's entries by layer: the k-th occurrence of a grid
    // point within the atom (a sphere overlapping its own periodic image)
    // goes to layer k.  A layer touches each point once and is one add unit.
    std::vector<int64_t> perm(ne);
    std::vector<int32_t> unit_atom;
    std::vector<int64_t> unit_beg, unit_end;
    for (int a = 0; a < natoms; a++) {
        const int64_t e0 = entry_ptr[a], n = entry_ptr[a + 1] - e0;
        if (n == 0) continue;
        std::vector<int64_t> byG(n);
        std::iota(byG.begin(), byG.end(), 0);
        std::stable_sort(byG.begin(), byG.end(), [&](int64_t x, int64_t y) {
            return G_e[e0 + x] < G_e[e0 + y];
        });
        std::vector<int> occ(n, 0);
        for (int64_t i = 1; i < n; i++)
            if (G_e[e0 + byG[i]] == G_e[e0 + byG[i - 1]]) occ[byG[i]] = occ[byG[i - 1]] + 1;
        std::vector<int64_t> loc(n);
        std::iota(loc.begin(), loc.end(), 0);
        std::stable_sort(loc.begin(), loc.end(),
                         [&](int64_t x, int64_t y) { return occ[x] < occ[y]; });
        for (int64_t i = 0; i < n; i++) {
            perm[e0 + i] = e0 + loc[i];
            if (i == 0 || occ[loc[i]] != occ[loc[i - 1]]) {
                if (i > 0) unit_end.push_back(e0 + i);
                unit_atom.push_back(a);
                unit_beg.push_back(e0 + i);
            }
        }
        unit_end.push_back(e0 + n);
    }
    const int nunits = (int)unit_atom.size();
    std::vector<int32_t> G2(ne), w2(ne, -1), unit_of_e(ne);
    for (int64_t e = 0; e < ne; e++) {
        G2[e] = G_e[perm[e]];
        if (w_e) w2[e] = w_e[perm[e]];
    }
    for (int u = 0; u < nunits; u++)
        for (int64_t e = unit_beg[u]; e < unit_end[u]; e++) unit_of_e[e] = u;
    // colour the units so that units sharing a grid point differ in colour
    std::vector<int64_t> order(ne);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t x, int64_t y) { return G2[x] < G2[y]; });
    std::vector<std::vector<int>> nbr(nunits);
    int64_t nunique = 0;
    for (int64_t i = 0; i < ne;) {
        int64_t j = i;
        while (j < ne && G2[order[j]] == G2[order[i]]) j++;
        nunique++;
        for (int64_t p = i; p < j; p++)
            for (int64_t q = p + 1; q < j; q++) {
                int uu = unit_of_e[order[p]], vv = unit_of_e[order[q]];
                if (uu != vv) { nbr[uu].push_back(vv); nbr[vv].push_back(uu); }
            }
        i = j;
    }
    std::vector<int> color(nunits, -1);
    int ncolors = 0;
    for (int u = 0; u < nunits; u++) {
        std::vector<char> used(ncolors + 1, 0);
        for (int v : nbr[u])
            if (color[v] >= 0) used[color[v]] = 1;
        int c = 0;
        while (used[c]) c++;
        color[u] = c;
        ncolors = std::max(ncolors, c + 1);
    }
    std::vector<int32_t> color_units;
    std::vector<int> color_ptr(1, 0);
    for (int c = 0; c < ncolors; c++) {
        for (int u = 0; u < nunits; u++)
            if (color[u] == c && M_ptr[unit_atom[u] + 1] > M_ptr[unit_atom[u]])
                color_units.push_back(u);
        color_ptr.push_back((int)color_units.size());
    }

    auto* h = new gh_lfc();
    h->dtype = dtype;
    h->natoms = natoms;
    h->Mmax = M_ptr[natoms];
    h->nu = (int)nunique;
    h->color_ptr = color_ptr;
    for (int a = 0; a < natoms; a++) {
        const int nma = M_ptr[a + 1] - M_ptr[a];
        if (a == 0) h->uniform_nm = nma;
        else if (h->uniform_nm != nma) h->uniform_nm = -1;
        h->max_nm = std::max(h->max_nm, M_ptr[a + 1] - M_ptr[a]);
    }
    for (int u = 0; u < nunits; u++)
        h->max_unit = std::max(h->max_unit, unit_end[u] - unit_beg[u]);
    h->ngpts = ngpts;
    h->nentries = ne;
    h->nq = phase_qw && nw > 0 ? nq : 0;
    h->nw = nw;
    h->dv = dv;
    // function values in the permuted entry order
    std::vector<double> A2((size_t)A_off[natoms]);
    for (int a = 0; a < natoms; a++) {
        const int nm = M_ptr[a + 1] - M_ptr[a];
        for (int64_t e = entry_ptr[a]; e < entry_ptr[a + 1]; e++)
            std::copy(A + A_off[a] + (perm[e] - entry_ptr[a]) * nm,
                      A + A_off[a] + (perm[e] - entry_ptr[a] + 1) * nm,
                      A2.begin() + A_off[a] + (e - entry_ptr[a]) * nm);
    }
    int rc;
#define UP(dst, src, n) if ((rc = upload(dst, src, n)) != GH_OK) { free_lfc(h); return rc; }
    UP(&h->entry_ptr, entry_ptr, (size_t)natoms + 1);
    UP(&h->M_ptr, M_ptr, (size_t)natoms + 1);
    UP(&h->A_off, A_off.data(), A_off.size());
    UP(&h->dH_off, dH_off.data(), dH_off.size());
    UP(&h->G_e, G2.data(), (size_t)ne);
    UP(&h->w_e, w2.data(), w2.size());
    UP(&h->unit_atom, unit_atom.data(), unit_atom.size());
    UP(&h->unit_beg, unit_beg.data(), unit_beg.size());
    UP(&h->unit_end, unit_end.data(), unit_end.size());
    UP(&h->color_units, color_units.data(), color_units.size());
    const size_t nA = A2.size();
    if (single(dtype)) {
        std::vector<float> Af(A2.begin(), A2.end());
        float* d = nullptr;
        UP(&d, Af.data(), nA);
        h->A = d;
        if (h->nq > 0) {
            std::vector<float> pf(phase_qw, phase_qw + (size_t)2 * nq * nw);
            float* dp = nullptr;
            UP(&dp, pf.data(), pf.size());
            h->phase_w = dp;
        }
    } else {
        double* d = nullptr;
        UP(&d, A2.data(), nA);
        h->A = d;
        if (h->nq > 0) {
            double* dp = nullptr;
            UP(&dp, phase_qw, (size_t)2 * nq * nw);
            h->phase_w = dp;
        }
    }
    {
        // point-major add tables: entries by grid point, for one point in
        // entry order (atom, then layer); slots = the distinct points
        std::vector<int64_t> order(ne);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(),
                         [&](int64_t x, int64_t y) { return G2[x] < G2[y]; });
        std::vector<int32_t> atom_of(ne);
        for (int a = 0; a < natoms; a++)
            for (int64_t e = entry_ptr[a]; e < entry_ptr[a + 1]; e++) atom_of[e] = a;
        std::vector<int32_t> slot_ptr(1, 0), slot_G, se_atom(ne), se_w(ne);
        std::vector<int64_t> se_A(ne);
        for (int64_t i = 0; i < ne; i++) {
            const int64_t e = order[i];
            const int a = atom_of[e];
            se_atom[i] = a;
            se_w[i] = w2[e];
            se_A[i] = A_off[a] + (e - entry_ptr[a]) * (M_ptr[a + 1] - M_ptr[a]);
            if (i == 0 || G2[e] != G2[order[i - 1]]) {
                if (i > 0) slot_ptr.push_back((int32_t)i);
                slot_G.push_back(G2[e]);
            }
        }
        if (ne > 0) slot_ptr.push_back((int32_t)ne);
        const int64_t ns = (int64_t)slot_G.size();
        const int nch = (int)((ns + 63) / 64);
        std::vector<int32_t> chunk_ptr(1, 0), chunk_atom;
        for (int ch = 0; ch < nch; ch++) {
            std::vector<int32_t> at;
            const int64_t u1 = std::min<int64_t>(ns, (int64_t)(ch + 1) * 64);
            for (int64_t u = (int64_t)ch * 64; u < u1; u++)
                for (int p = slot_ptr[u]; p < slot_ptr[u + 1]; p++) at.push_back(se_atom[p]);
            std::sort(at.begin(), at.end());
            at.erase(std::unique(at.begin(), at.end()), at.end());
            chunk_atom.insert(chunk_atom.end(), at.begin(), at.end());
            chunk_ptr.push_back((int32_t)chunk_atom.size());
        }
        h->nslots = ns;
        h->nchunks = nch;
        UP(&h->slot_ptr, slot_ptr.data(), slot_ptr.size());
        UP(&h->slot_G, slot_G.data(), slot_G.size());
        UP(&h->se_atom, se_atom.data(), se_atom.size());
        UP(&h->se_A, se_A.data(), se_A.size());
        UP(&h->se_w, se_w.data(), se_w.size());
        UP(&h->chunk_ptr, chunk_ptr.data(), chunk_ptr.size());
        UP(&h->chunk_atom, chunk_atom.data(), chunk_atom.size());
        const char* mode = getenv("GPAW_AMD_LFC_ADD");
        h->points_add = !(mode && std::string(mode) == "colour");
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

// Phase table of k-point q (null for real data or no images).
static const void* phase_table(const gh_lfc* h, int q) {
    if (!cplxd(h->dtype) || h->nq == 0) return nullptr;
    const size_t s = single(h->dtype) ? sizeof(float) : sizeof(double);
    return (const char*)h->phase_w + (size_t)q * h->nw * 2 * s;
}

int gh_lfc_integrate(const gh_lfc* h, const void* a_dev, void* c_dev,
                     int64_t nbands, int q, void* stream) {
    GH_REQUIRE(h, GH_ERR_ARG, "lfc_integrate: null handle");
    GH_REQUIRE(nbands >= 0, GH_ERR_ARG, "lfc_integrate: nbands < 0");
    GH_REQUIRE(q >= 0 && (q < h->nq || (h->nq == 0 && q == 0)), GH_ERR_ARG,
               "lfc_integrate: k-point index %d out of range (%d tables)", q, h->nq);
    if (nbands == 0 || h->natoms == 0) return GH_OK;  // nothing to read or write
    GH_REQUIRE(a_dev && c_dev, GH_ERR_ARG, "lfc_integrate: null argument");
    // every (band, M) is written by the block of its atom, also for atoms
    // without grid points, so no clearing pass is needed
    hipStream_t st = (hipStream_t)stream;
    const int bt = h->dtype == GH_C128 ? IntBT<cplx<double>>::value
                   : h->dtype == GH_C64 ? IntBT<cplx<float>>::value
                                        : IntBT<double>::value;
    dim3 grid(h->natoms, (unsigned)((nbands + bt - 1) / bt));
    const void* ph = phase_table(h, q);
#define ARGS(V, S) (const V*)a_dev, (V*)c_dev, h->entry_ptr, h->M_ptr, h->A_off, h->G_e, h->w_e, \
                   (const S*)h->A, (const S*)ph, h->ngpts, h->Mmax, (int)nbands, (S)h->dv
    switch (h->dtype) {
        case GH_F64: hipLaunchKernelGGL((lfc_integrate_kernel<double, double>), grid, dim3(256), 0, st, ARGS(double, double)); break;
        case GH_C128: hipLaunchKernelGGL((lfc_integrate_kernel<cplx<double>, double>), grid, dim3(256), 0, st, ARGS(cplx<double>, double)); break;
        case GH_F32: hipLaunchKernelGGL((lfc_integrate_kernel<float, float>), grid, dim3(256), 0, st, ARGS(float, float)); break;
        case GH_C64: hipLaunchKernelGGL((lfc_integrate_kernel<cplx<float>, float>), grid, dim3(256), 0, st, ARGS(cplx<float>, float)); break;
    }
#undef ARGS
    GH_CHECK_LAUNCH();
    return GH_OK;
}

int gh_lfc_add(const gh_lfc* h, const void* c_dev, void* a_dev,
               int64_t nbands, int q, void* stream) {
    GH_REQUIRE(h, GH_ERR_ARG, "lfc_add: null handle");
    GH_REQUIRE(nbands >= 0, GH_ERR_ARG, "lfc_add: nbands < 0");
    GH_REQUIRE(q >= 0 && (q < h->nq || (h->nq == 0 && q == 0)), GH_ERR_ARG,
               "lfc_add: k-point index %d out of range (%d tables)", q, h->nq);
    if (nbands == 0 || h->nentries == 0) return GH_OK;  // nothing to read or write
    GH_REQUIRE(a_dev && c_dev, GH_ERR_ARG, "lfc_add: null argument");
    hipStream_t st = (hipStream_t)stream;
    const void* ph = phase_table(h, q);
    if (h->points_add && h->max_nm <= 32) {
        const int mx = h->max_nm <= 8 ? 8 : h->max_nm <= 16 ? 16 : 32;
#define PT_ARGS(V, S) (const V*)c_dev, (V*)a_dev, h->chunk_ptr, h->chunk_atom, h->slot_ptr, \
                      h->slot_G, h->se_atom, h->se_A, h->se_w, h->M_ptr, (const S*)h->A,   \
                      (const S*)ph, h->nslots, h->ngpts, h->Mmax, (int)nbands
#define PT_LAUNCH(V, S, BT)                                                                   \
    do {                                                                                      \
        dim3 grid((unsigned)h->nchunks, (unsigned)((nbands + BT - 1) / BT));                  \
        if (h->uniform_nm == 8) hipLaunchKernelGGL((lfc_add_points_kernel<V, S, 8, BT, true>), grid, dim3(64), 0, st, PT_ARGS(V, S)); \
        else if (mx == 8) hipLaunchKernelGGL((lfc_add_points_kernel<V, S, 8, BT, false>), grid, dim3(64), 0, st, PT_ARGS(V, S)); \
        else if (mx == 16) hipLaunchKernelGGL((lfc_add_points_kernel<V, S, 16, BT, false>), grid, dim3(64), 0, st, PT_ARGS(V, S)); \
        else hipLaunchKernelGGL((lfc_add_points_kernel<V, S, 32, BT, false>), grid, dim3(64), 0, st, PT_ARGS(V, S)); \
    } while (0)
        switch (h->dtype) {
            case GH_F64: PT_LAUNCH(double, double, 16); break;
            case GH_C128: PT_LAUNCH(cplx<double>, double, 8); break;
            case GH_F32: PT_LAUNCH(float, float, 32); break;
            case GH_C64: PT_LAUNCH(cplx<float>, float, 16); break;
        }
#undef PT_LAUNCH
#undef PT_ARGS
        GH_CHECK_LAUNCH();
        return GH_OK;
    }
    const bool uniform8 = h->uniform_nm == 8;
    for (size_t col = 0; col + 1 < h->color_ptr.size(); col++) {
        const int n = h->color_ptr[col + 1] - h->color_ptr[col];
        if (n == 0) continue;
        const int nchunks = (int)((h->max_unit + kAddChunk - 1) / kAddChunk);
        // small colours take narrower band tiles, so that every launch
        // keeps the GPU busy (about 16384 wavefronts)
        int bt = kAddBT;
        while (bt > 4 && (int64_t)n * nchunks * ((nbands + bt - 1) / bt) < 16384) bt /= 2;
        const unsigned ntiles = (unsigned)((nbands + bt - 1) / bt);
        dim3 grid(n * nchunks, ntiles);
        const int32_t* units = h->color_units + h->color_ptr[col];
#define ARGS(V, S) (const V*)c_dev, (V*)a_dev, units, nchunks, h->unit_atom, h->unit_beg, \
                   h->unit_end, h->entry_ptr, h->M_ptr, h->A_off, \
                   h->G_e, h->w_e, (const S*)h->A, (const S*)ph, \
                   h->ngpts, h->Mmax, (int)nbands, bt
        switch (h->dtype) {
            case GH_F64:
                if (uniform8) hipLaunchKernelGGL((lfc_add_kernel<double, double, 8>), grid, dim3(64), 0, st, ARGS(double, double));
                else hipLaunchKernelGGL((lfc_add_kernel<double, double, 0>), grid, dim3(64), 0, st, ARGS(double, double));
                break;
            case GH_C128:
                if (uniform8) hipLaunchKernelGGL((lfc_add_kernel<cplx<double>, double, 8>), grid, dim3(64), 0, st, ARGS(cplx<double>, double));
                else hipLaunchKernelGGL((lfc_add_kernel<cplx<double>, double, 0>), grid, dim3(64), 0, st, ARGS(cplx<double>, double));
                break;
            case GH_F32:
                if (uniform8) hipLaunchKernelGGL((lfc_add_kernel<float, float, 8>), grid, dim3(64), 0, st, ARGS(float, float));
                else hipLaunchKernelGGL((lfc_add_kernel<float, float, 0>), grid, dim3(64), 0, st, ARGS(float, float));
                break;
            case GH_C64:
                if (uniform8) hipLaunchKernelGGL((lfc_add_kernel<cplx<float>, float, 8>), grid, dim3(64), 0, st, ARGS(cplx<float>, float));
                else hipLaunchKernelGGL((lfc_add_kernel<cplx<float>, float, 0>), grid, dim3(64), 0, st, ARGS(cplx<float>, float));
                break;
        }
#undef ARGS
        GH_CHECK_LAUNCH();
    }
    return GH_OK;
}

int gh_lfc_apply_dH(const gh_lfc* h, const void* P_dev, const double* dH_dev,
                    const double* eps_dev, const double* dS_dev,
                    void* out_dev, int64_t nbands, void* stream) {
    GH_REQUIRE(h, GH_ERR_ARG, "lfc_apply_dH: null handle");
    GH_REQUIRE(nbands >= 0, GH_ERR_ARG, "lfc_apply_dH: nbands < 0");
    if (nbands == 0 || h->Mmax == 0) return GH_OK;  // nothing to read or write
    GH_REQUIRE(P_dev && dH_dev && out_dev, GH_ERR_ARG, "lfc_apply_dH: null argument");
    GH_REQUIRE(P_dev != out_dev, GH_ERR_ARG, "lfc_apply_dH: P and out must differ");
    if (nbands > kBandSlice) {
        // one grid row per band: pass larger band sets in slices
        const int64_t row = (int64_t)h->Mmax * dtype_size(h->dtype);
        for (int64_t x0 = 0; x0 < nbands; x0 += kBandSlice) {
            const int64_t nx = std::min<int64_t>(kBandSlice, nbands - x0);
            const int rc = gh_lfc_apply_dH(h, (const char*)P_dev + x0 * row, dH_dev,
                                           eps_dev ? eps_dev + x0 : nullptr, dS_dev,
                                           (char*)out_dev + x0 * row, nx, stream);
            if (rc != GH_OK) return rc;
        }
        return GH_OK;
    }
    hipStream_t st = (hipStream_t)stream;
    dim3 grid((h->Mmax + 127) / 128, (unsigned)nbands);
#define DH_ARGS(V) (const V*)P_dev, dH_dev, h->dH_off, eps_dev, dS_dev, h->M_ptr, \
                   (V*)out_dev, h->natoms, h->Mmax, (int)nbands
    switch (h->dtype) {
        case GH_F64:
            hipLaunchKernelGGL((lfc_dH_kernel<double, double>), grid, dim3(128), 0, st,
                               DH_ARGS(double));
            break;
        case GH_C128:
            hipLaunchKernelGGL((lfc_dH_kernel<cplx<double>, cplx<double>>), grid, dim3(128), 0,
                               st, DH_ARGS(cplx<double>));
            break;
        case GH_F32:
            hipLaunchKernelGGL((lfc_dH_kernel<float, double>), grid, dim3(128), 0, st,
                               DH_ARGS(float));
            break;
        case GH_C64:
            hipLaunchKernelGGL((lfc_dH_kernel<cplx<float>, cplx<double>>), grid, dim3(128), 0,
                               st, DH_ARGS(cplx<float>));
            break;
    }
#undef DH_ARGS
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

