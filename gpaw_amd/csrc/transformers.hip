// Grid restriction and interpolation on gfx950 (Transformer.apply).
//
// Semantics follow the reference's Transformer.apply (c/transformers.c:44-
// 164): the input gets ghost layers (c/bc.c; periodic with Bloch phase or
// zero, widths from transformers.py:27-47), then three separable 1-D passes
// along axis 2, 1, 0 (c/bmgs/restrict.c:17-62,136-155;
// c/bmgs/interpolate.c:9-46,70-100).  Per axis and output index o:
//   restrict:    c = 2 o + (K-1) - pad_lo
//                b = 0.5 * (a[c] + sum_j w_j (a[c + 2j+1] + a[c - 2j-1]))
//   interpolate: q = o + skip_lo, i = q / 2, c = i + nn - 1 - pad_lo
//                b = a[c]                                  (q even)
//                b = sum_t w_t (a[c - t] + a[c + 1 + t])   (q odd)
// with index c in interior coordinates; out-of-range indices are ghosts.
//
// One kernel per direction, one pass over HBM: a block owns one band, a
// tile of full output rows (all z) and a run of output x-planes, and marches
// along x.  For every input plane it needs, the block stages the input rows
// of its tile in LDS, filters them along z (LDS -> LDS) and along y (LDS ->
// registers), and keeps the last 2K-1 (restrict) or K (interpolate) filtered
// planes in a register queue that the x-filter reads.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "common.h"
#include "internal.h"

namespace gh {

template <typename S>
struct TPhase {
    S re[3][2];
    S im[3][2];
};

struct TGeom {
    int n[3];      // input grid
    int m[3];      // output grid
    int plo[3];    // ghost width below the interior (pad_cd[:, 0])
    int skip[3];   // interpolation skip_lo (skip_cd[:, 0])
    int pbc[3];
    int tj;        // output rows per tile
    int lx;        // output planes per block
    int nyt, nxc;
    int nry;       // input rows staged per tile (upper bound)
    int zlo;       // ghost columns staged below z = 0
    int zw;        // staged row width: zlo + n[2] + ghost columns above
    int yz_trivial;  // Bloch phases along y and z are all 1 (k along x)
    int vec;         // restriction: stage 16-byte vectors of vec elements
};

template <int K> struct W;
template <> struct W<2> { static constexpr double w[1] = {0.5}; };
template <> struct W<4> { static constexpr double w[2] = {0.5625, -0.0625}; };
template <> struct W<6> { static constexpr double w[3] = {0.58593750, -0.09765625, 0.01171875}; };
template <> struct W<8> { static constexpr double w[4] = {0.59814453125, -0.11962890625, 0.02392578125, -0.00244140625}; };

// Wrap an interior index along axis ax (one period at most); returns false
// for a zero ghost, else sets the phase factor.
template <typename S>
__device__ __forceinline__ bool wrap_idx(int& i, int n, int pbc, int ax,
                                         const TPhase<S>& ph, S& pre, S& pim) {
    if (i >= 0 && i < n) return true;
    if (!pbc) return false;
    int d = i < 0 ? 1 : 0;
    i += i < 0 ? n : -n;
    S r = pre * ph.re[ax][d] - pim * ph.im[ax][d];
    S m = pre * ph.im[ax][d] + pim * ph.re[ax][d];
    pre = r;
    pim = m;
    return true;
}

// (row, column) of the flat index threadIdx.x + t * blockDim.x in a
// row-major array of width w, advanced with t without integer division.
struct RowWalk {
    int r, c, dr, dc, w;
    __device__ __forceinline__ explicit RowWalk(int w_) : w(w_) {
        r = threadIdx.x / w;
        c = threadIdx.x - r * w;
        dr = blockDim.x / w;
        dc = blockDim.x - dr * w;
    }
    __device__ __forceinline__ void next() {
        r += dr;
        c += dc;
        if (c >= w) {
            c -= w;
            r++;
        }
    }
};

// Register prefetch of a staged plane: thread t holds elements
// t + s * blockDim.x (s < SPT) of the nrows x n2 stage, loaded while the
// previous plane is filtered; the part of larger tiles beyond
// SPT * blockDim.x is staged synchronously.
template <typename V> struct StageRegs {
    static constexpr int SPT = 4;
    // loads issued together by the synchronous part of the staging
    static constexpr int kStageBatch = sizeof(V) <= 8 ? 8 : 4;
    V v[SPT];
};

// Staged element i of plane ix (row i / zw, column i % zw, which is
// z = column - zlo: each staged row carries its z ghosts): sets its offset
// in the band and the Bloch phase of its periodic image; false for a zero
// ghost (or, for the plane after the last one a block uses, beyond one
// period: nothing is loaded).
template <typename S>
__device__ __forceinline__ bool stage_elem(int r, int col, int ix, int iy0,
                                           const TGeom& g, const TPhase<S>& ph,
                                           size_t& off, S& pre, S& pim) {
    const int n2 = g.n[2];
    int z = col - g.zlo;
    int y = iy0 + r;
    int x = ix;
    pre = 1;
    pim = 0;
    if (!wrap_idx(x, g.n[0], g.pbc[0], 0, ph, pre, pim) || x < 0 || x >= g.n[0])
        return false;
    if (!wrap_idx(y, g.n[1], g.pbc[1], 1, ph, pre, pim)) return false;
    if (!wrap_idx(z, n2, g.pbc[2], 2, ph, pre, pim) || z < 0 || z >= n2) return false;
    off = ((size_t)x * g.n[1] + y) * n2 + z;
    return true;
}

// A loaded element (zero for ghosts) with its Bloch phase applied; real
// data need no phase.
template <typename V, typename S>
__device__ __forceinline__ V staged_value(V v, int r, int col, int ix, int iy0,
                                          const TGeom& g, const TPhase<S>& ph) {
    if constexpr (!is_complex<V>::value) {
        return v;
    } else {
        if (g.yz_trivial) {
            // only the x image carries a phase: one factor per plane
            if (ix >= 0 && ix < g.n[0]) return v;
            int x = ix;
            S xre = 1, xim = 0;
            wrap_idx(x, g.n[0], g.pbc[0], 0, ph, xre, xim);
            return mul_phase(v, xre, xim);
        }
        size_t off;
        S pre, pim;
        stage_elem(r, col, ix, iy0, g, ph, off, pre, pim);
        return mul_phase(v, pre, pim);
    }
}

// Prefetch: raw values only, so that the loads stay in flight while the
// previous plane is filtered; the phase is applied in stage_store.
template <typename V, typename S>
__device__ __forceinline__ void stage_load(const V* __restrict__ inb, int ix,
                                           int iy0, int nrows,
                                           StageRegs<V>& R, const TGeom& g,
                                           const TPhase<S>& ph, RowWalk w) {
#pragma unroll
    for (int s = 0; s < StageRegs<V>::SPT; s++, w.next()) {
        V v = vzero<V>();
        size_t off;
        S pre, pim;
        if (w.r < nrows && stage_elem(w.r, w.c, ix, iy0, g, ph, off, pre, pim))
            v = inb[off];
        R.v[s] = v;
    }
}

template <typename V, typename S>
__device__ __forceinline__ void stage_store(const V* __restrict__ inb, int ix,
                                            int iy0, int nrows,
                                            const StageRegs<V>& R, V* A,
                                            const TGeom& g,
                                            const TPhase<S>& ph, RowWalk w) {
#pragma unroll
    for (int s = 0; s < StageRegs<V>::SPT; s++, w.next())
        if (w.r < nrows)
            A[w.r * g.zw + w.c] = staged_value(R.v[s], w.r, w.c, ix, iy0, g, ph);
    // the rest: batches of kStageBatch loads in flight before their stores
    constexpr int NB = StageRegs<V>::kStageBatch;
    while (w.r < nrows) {
        V v[NB];
        RowWalk w0 = w;
#pragma unroll
        for (int s = 0; s < NB; s++, w.next()) {
            size_t off;
            S pre, pim;
            v[s] = vzero<V>();
            if (w.r < nrows && stage_elem(w.r, w.c, ix, iy0, g, ph, off, pre, pim))
                v[s] = inb[off];
        }
#pragma unroll
        for (int s = 0; s < NB; s++, w0.next())
            if (w0.r < nrows)
                A[w0.r * g.zw + w0.c] = staged_value(v[s], w0.r, w0.c, ix, iy0, g, ph);
    }
}

// Staging in 16-byte vectors of VEC = 16 / sizeof(V) elements (8-byte and
// 4-byte types): the staged row width and its ghost columns below z = 0 are
// multiples of VEC and so is n2, so a vector never straddles a periodic
// image or the row end, and its global address is 16-byte aligned.  One
// load moves VEC elements (twice or four times the bytes in flight of the
// element-wise staging); the phase of every element is applied at the
// store as before.
template <typename V, int VEC> struct VecT { V e[VEC]; };
typedef float f4x __attribute__((ext_vector_type(4)));

template <typename V, int VEC> struct StageRegsV {
    static constexpr int SPT = 4;
    static constexpr int kStageBatch = 4;
    VecT<V, VEC> v[SPT];
};

template <typename V, typename S, int VEC>
__device__ __forceinline__ VecT<V, VEC> vec_load(const V* __restrict__ inb, int r, int cv,
                                                 int ix, int iy0, int nrows, const TGeom& g,
                                                 const TPhase<S>& ph) {
    VecT<V, VEC> v;
    size_t off;
    S pre, pim;
    if (r < nrows && stage_elem(r, cv * VEC, ix, iy0, g, ph, off, pre, pim)) {
        const f4x t = *reinterpret_cast<const f4x*>(inb + off);
        __builtin_memcpy(&v, &t, 16);
    } else {
#pragma unroll
        for (int e = 0; e < VEC; e++) v.e[e] = vzero<V>();
    }
    return v;
}

template <typename V, typename S, int VEC>
__device__ __forceinline__ void vec_store(const VecT<V, VEC>& v, V* A, int r, int cv, int ix,
                                          int iy0, const TGeom& g, const TPhase<S>& ph) {
    VecT<V, VEC> o;
#pragma unroll
    for (int e = 0; e < VEC; e++)
        o.e[e] = staged_value(v.e[e], r, cv * VEC + e, ix, iy0, g, ph);
    f4x t;
    __builtin_memcpy(&t, &o, 16);
    *reinterpret_cast<f4x*>(A + r * g.zw + cv * VEC) = t;
}

template <typename V, typename S, int VEC>
__device__ __forceinline__ void stage_load_v(const V* __restrict__ inb, int ix, int iy0,
                                             int nrows, StageRegsV<V, VEC>& R,
                                             const TGeom& g, const TPhase<S>& ph,
                                             RowWalk w) {
#pragma unroll
    for (int s = 0; s < StageRegsV<V, VEC>::SPT; s++, w.next())
        R.v[s] = vec_load<V, S, VEC>(inb, w.r, w.c, ix, iy0, nrows, g, ph);
}

template <typename V, typename S, int VEC>
__device__ __forceinline__ void stage_store_v(const V* __restrict__ inb, int ix, int iy0,
                                              int nrows, const StageRegsV<V, VEC>& R, V* A,
                                              const TGeom& g, const TPhase<S>& ph,
                                              RowWalk w) {
#pragma unroll
    for (int s = 0; s < StageRegsV<V, VEC>::SPT; s++, w.next())
        if (w.r < nrows) vec_store<V, S, VEC>(R.v[s], A, w.r, w.c, ix, iy0, g, ph);
    constexpr int NB = StageRegsV<V, VEC>::kStageBatch;
    while (w.r < nrows) {
        VecT<V, VEC> v[NB];
        RowWalk w0 = w;
#pragma unroll
        for (int s = 0; s < NB; s++, w.next())
            v[s] = vec_load<V, S, VEC>(inb, w.r, w.c, ix, iy0, nrows, g, ph);
#pragma unroll
        for (int s = 0; s < NB; s++, w0.next())
            if (w0.r < nrows) vec_store<V, S, VEC>(v[s], A, w0.r, w0.c, ix, iy0, g, ph);
    }
}

// Value of a staged row at z (row points at z = 0; the staged ghost
// columns already carry their phase or zero).
template <typename V, typename S>
__device__ __forceinline__ V row_at(const V* row, int z, const TGeom&,
                                    const TPhase<S>&) {
    return row[z];
}

template <typename V, typename S, int K>
__device__ __forceinline__ V restrict_tap(const V* row, int c, const TGeom& g,
                                          const TPhase<S>& ph) {
    V acc = row_at(row, c, g, ph);
#pragma unroll
    for (int j = 0; j < K / 2; j++) {
        int o = 2 * j + 1;
        acc = fma_rs((S)W<K>::w[j],
                     vadd(row_at(row, c + o, g, ph), row_at(row, c - o, g, ph)),
                     acc);
    }
    return vscale(acc, (S)0.5);
}

// Restriction: output tile rows [J0, J0 + tj), output planes [I0, I1).
template <typename V, typename S, int K, int PPT, int VEC = 1>
__global__ void __launch_bounds__(1024)
restrict_kernel(const V* __restrict__ in, V* __restrict__ out, TGeom g,
                TPhase<S> ph, S alpha) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int n2 = g.n[2], m2 = g.m[2];
    V* A = reinterpret_cast<V*>(smem_raw);   // nry x zw
    V* B = A + g.nry * g.zw;                 // tj x zw (y-filtered)
    // thread's staged element (vector) and y-filtered element
    const RowWalk wz(g.zw / VEC), wy(g.zw);
    const int per_band = g.nyt * g.nxc;
    const int b = xcd_swizzle(blockIdx.x, gridDim.x);
    const int band = b / per_band;
    const int rem = b - band * per_band;
    const int xc = rem / g.nyt;
    const int yt = rem - xc * g.nyt;
    const int J0 = yt * g.tj;
    const int tje = min(g.tj, g.m[1] - J0);
    const int I0 = xc * g.lx;
    const int I1 = min(I0 + g.lx, g.m[0]);
    const V* inb = in + (size_t)band * g.n[0] * g.n[1] * n2;
    V* outb = out + (size_t)band * g.m[0] * g.m[1] * m2;
    // input rows: centre 2J + K-1 - plo, taps +-(K-1)
    const int iy0 = 2 * J0 - g.plo[1];
    const int nrows = 2 * tje + 2 * K - 3;
    const int npts = tje * m2;

    int jl[PPT], kz[PPT];
    bool act[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        int p = threadIdx.x + k * blockDim.x;
        act[k] = p < npts;
        int pp = act[k] ? p : 0;
        jl[k] = pp / m2;
        kz[k] = pp - jl[k] * m2;
    }
    constexpr int QL = 2 * K - 1;
    // register queue of z- and y-filtered planes per point.  For complex64
    // every access uses a compile-time slot (the new planes land in the last
    // slots and the queue shifts): priming it slot by slot put that queue in
    // scratch memory and doubled the kernel's time.  The other types keep
    // the slot-by-slot priming, measured faster for them.
    constexpr bool kShiftIn = std::is_same<V, cplx<float>>::value;
    V q[PPT][QL];
    if constexpr (kShiftIn) {
#pragma unroll
        for (int k = 0; k < PPT; k++)
#pragma unroll
            for (int s = 0; s < QL; s++) q[k][s] = vzero<V>();
    }
    struct ZRow { V v[PPT]; };
    const int base = 2 * I0 - g.plo[0];  // first input plane of output I0

    // plane ix is in registers on entry (loaded during the previous
    // plane's filters); the next plane ix + 1 is loaded during this one's
    typename std::conditional<(VEC > 1), StageRegsV<V, VEC>, StageRegs<V>>::type pre;
    auto sload = [&](int ix) {
        if constexpr (VEC > 1) stage_load_v<V, S, VEC>(inb, ix, iy0, nrows, pre, g, ph, wz);
        else stage_load<V, S>(inb, ix, iy0, nrows, pre, g, ph, wz);
    };
    auto process_plane = [&](int ix) -> ZRow {
        if constexpr (VEC > 1) stage_store_v<V, S, VEC>(inb, ix, iy0, nrows, pre, A, g, ph, wz);
        else stage_store<V, S>(inb, ix, iy0, nrows, pre, A, g, ph, wz);
        __syncthreads();
        sload(ix + 1);
        // y-filter every staged column for the tile's output rows only
        // (rows are already ghost-filled); filtering along z first would
        // also filter the 2K - 3 rows that only feed the y taps
        RowWalk w = wy;
        for (; w.r < tje; w.next()) {
            const int jr = w.r, col = w.c;
            const int c = 2 * jr + (K - 1);
            V acc = A[c * g.zw + col];
#pragma unroll
            for (int j = 0; j < K / 2; j++) {
                const int o = 2 * j + 1;
                acc = fma_rs((S)W<K>::w[j],
                             vadd(A[(c + o) * g.zw + col], A[(c - o) * g.zw + col]), acc);
            }
            B[jr * g.zw + col] = vscale(acc, (S)0.5);
        }
        __syncthreads();
        // z-filter at the own points
        ZRow z;
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            z.v[k] = vzero<V>();
            if (!act[k]) continue;
            const int c = 2 * kz[k] + (K - 1) - g.plo[2];
            z.v[k] = restrict_tap<V, S, K>(B + jl[k] * g.zw + g.zlo, c, g, ph);
        }
        return z;
    };

    // prime the first 2K-3 planes: slots 0 .. QL-3 (shift-in form: they
    // end in slots 2 .. QL-1 and move down at the top of each output plane)
    sload(base);
    if constexpr (kShiftIn) {
        for (int s = 0; s < QL - 2; s++) {
            const ZRow z = process_plane(base + s);
#pragma unroll
            for (int k = 0; k < PPT; k++) {
#pragma unroll
                for (int t = 0; t < QL - 1; t++) q[k][t] = q[k][t + 1];
                q[k][QL - 1] = z.v[k];
            }
        }
    } else {
#pragma unroll
        for (int s = 0; s < QL - 2; s++) {
            const ZRow z = process_plane(base + s);
#pragma unroll
            for (int k = 0; k < PPT; k++) q[k][s] = z.v[k];
        }
    }
    for (int I = I0; I < I1; I++) {
        if constexpr (kShiftIn) {
#pragma unroll
            for (int k = 0; k < PPT; k++)
#pragma unroll
                for (int s = 0; s < QL - 2; s++) q[k][s] = q[k][s + 2];
        }
        int p0 = base + 2 * (I - I0) + QL - 2;
        const ZRow z0 = process_plane(p0);
#pragma unroll
        for (int k = 0; k < PPT; k++) q[k][QL - 2] = z0.v[k];
        const ZRow z1 = process_plane(p0 + 1);
#pragma unroll
        for (int k = 0; k < PPT; k++) q[k][QL - 1] = z1.v[k];
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            if (!act[k]) continue;
            const int c = K - 1;
            V acc = q[k][c];
#pragma unroll
            for (int j = 0; j < K / 2; j++) {
                int o = 2 * j + 1;
                acc = fma_rs((S)W<K>::w[j], vadd(q[k][c + o], q[k][c - o]), acc);
            }
            outb[((size_t)I * g.m[1] + J0 + jl[k]) * m2 + kz[k]] =
                vscale(vscale(acc, (S)0.5), alpha);
        }
        if constexpr (!kShiftIn) {
#pragma unroll
            for (int k = 0; k < PPT; k++)
#pragma unroll
                for (int s = 0; s < QL - 2; s++) q[k][s] = q[k][s + 2];
        }
    }
}

template <typename V, typename S, int K>
__device__ __forceinline__ V interp_tap(const V* row, int o, int skip, int plo,
                                        const TGeom& g, const TPhase<S>& ph) {
    constexpr int nn = K / 2;
    int qq = o + skip;
    int c = (qq >> 1) + nn - 1 - plo;
    if ((qq & 1) == 0) return row_at(row, c, g, ph);
    V acc = vzero<V>();
#pragma unroll
    for (int t = 0; t < nn; t++)
        acc = fma_rs((S)W<K>::w[t],
                     vadd(row_at(row, c - t, g, ph), row_at(row, c + 1 + t, g, ph)),
                     acc);
    return acc;
}

// Interpolation: output (fine) rows [J0, J0 + tj), fine planes [O0, O1).
template <typename V, typename S, int K, int PPT>
__global__ void __launch_bounds__(1024)
interpolate_kernel(const V* __restrict__ in, V* __restrict__ out, TGeom g,
                   TPhase<S> ph, S alpha) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    constexpr int nn = K / 2;
    const int n2 = g.n[2], m2 = g.m[2];
    V* A = reinterpret_cast<V*>(smem_raw);   // nry x zw
    V* B = A + g.nry * g.zw;                 // nry x m2
    const RowWalk wz(g.zw), wm(m2);          // thread's staged / filtered element
    const int per_band = g.nyt * g.nxc;
    const int b = xcd_swizzle(blockIdx.x, gridDim.x);
    const int band = b / per_band;
    const int rem = b - band * per_band;
    const int xc = rem / g.nyt;
    const int yt = rem - xc * g.nyt;
    const int J0 = yt * g.tj;
    const int tje = min(g.tj, g.m[1] - J0);
    const int O0 = xc * g.lx;
    const int O1 = min(O0 + g.lx, g.m[0]);
    const V* inb = in + (size_t)band * g.n[0] * g.n[1] * n2;
    V* outb = out + (size_t)band * g.m[0] * g.m[1] * m2;
    // coarse rows needed by fine rows J0 .. J0 + tje - 1
    const int cy0 = ((J0 + g.skip[1]) >> 1) - g.plo[1];
    const int cy1 = ((J0 + tje - 1 + g.skip[1]) >> 1) + 2 * nn - 1 - g.plo[1];
    const int nrows = cy1 - cy0 + 1;
    const int npts = tje * m2;

    int jl[PPT], kz[PPT];
    bool act[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        int p = threadIdx.x + k * blockDim.x;
        act[k] = p < npts;
        int pp = act[k] ? p : 0;
        jl[k] = pp / m2;
        kz[k] = pp - jl[k] * m2;
    }
    V q[PPT][K];  // coarse planes c - K + 1 .. c after the y/z filters

    // plane ix is in registers on entry; ix + 1 is loaded during its filters
    StageRegs<V> pre;
    auto process_plane = [&](int ix, int slot) {
        stage_store<V, S>(inb, ix, cy0, nrows, pre, A, g, ph, wz);
        __syncthreads();
        stage_load<V, S>(inb, ix + 1, cy0, nrows, pre, g, ph, wz);
        RowWalk w = wm;
        for (; w.r < nrows; w.next()) {
            const int r = w.r, kk = w.c;
            B[r * m2 + kk] = interp_tap<V, S, K>(A + r * g.zw + g.zlo, kk, g.skip[2],
                                                 g.plo[2], g, ph);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            if (!act[k]) continue;
            int J = J0 + jl[k];
            int qq = J + g.skip[1];
            int c = (qq >> 1) + nn - 1 - g.plo[1] - cy0;  // local row
            V v;
            if ((qq & 1) == 0) {
                v = B[c * m2 + kz[k]];
            } else {
                v = vzero<V>();
#pragma unroll
                for (int t = 0; t < nn; t++)
                    v = fma_rs((S)W<K>::w[t],
                               vadd(B[(c - t) * m2 + kz[k]], B[(c + 1 + t) * m2 + kz[k]]),
                               v);
            }
            q[k][slot] = v;
        }
    };

    // fine plane o -> coarse centre c(o) = ((o + skip) >> 1) + nn - 1 - plo;
    // it needs coarse planes c - nn + 1 .. c + nn.
    const int i_first = (O0 + g.skip[0]) >> 1;
    const int i_last = (O1 - 1 + g.skip[0]) >> 1;
    const int cbase = nn - 1 - g.plo[0];
    stage_load<V, S>(inb, i_first + cbase - nn + 1, cy0, nrows, pre, g, ph, wz);
#pragma unroll
    for (int s = 0; s < K - 1; s++) process_plane(i_first + cbase - nn + 1 + s, s);
    for (int i = i_first; i <= i_last; i++) {
        process_plane(i + cbase + nn, K - 1);
        // queue slot s holds coarse plane (c - nn + 1 + s), c = i + cbase
        const int oe = 2 * i - g.skip[0];
#pragma unroll
        for (int par = 0; par < 2; par++) {
            int o = oe + par;
            if (o < O0 || o >= O1) continue;
#pragma unroll
            for (int k = 0; k < PPT; k++) {
                if (!act[k]) continue;
                V v;
                if (par == 0) {
                    v = q[k][nn - 1];
                } else {
                    v = vzero<V>();
#pragma unroll
                    for (int t = 0; t < nn; t++)
                        v = fma_rs((S)W<K>::w[t], vadd(q[k][nn - 1 - t], q[k][nn + t]), v);
                }
                outb[((size_t)o * g.m[1] + J0 + jl[k]) * m2 + kz[k]] = vscale(v, alpha);
            }
        }
#pragma unroll
        for (int k = 0; k < PPT; k++)
#pragma unroll
            for (int s = 0; s < K - 1; s++) q[k][s] = q[k][s + 1];
    }
}

}  // namespace gh

struct gh_transformer {
    int dtype;
    int nn;
    bool interp;
    gh::TGeom g;
};

using namespace gh;

namespace {

template <typename S>
TPhase<S> tphase(const double* phase_cd) {
    Phases p = make_phases(phase_cd);
    TPhase<S> a;
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 2; d++) {
            a.re[c][d] = (S)p.re[c][d];
            a.im[c][d] = (S)p.im[c][d];
        }
    return a;
}

// Output points per thread: two, except for restriction with the long
// filter or complex128 data, where the register queue of 2K - 1 filtered
// planes per point would otherwise limit occupancy (measured at
// 112^3 -> 56^3).  Complex128 interpolation with K >= 6 keeps two points
// although its two K-plane queues spill some registers: one point per
// thread measured slower (1.82 against 1.46 ms at nn = 3).
template <typename V, int K>
constexpr int ppt_of(bool interp) {
    return (!interp && (K >= 6 || sizeof(V) > 8)) ? 1 : 2;
}

// Rows of the output tile: as many as fit 1024 threads and 64 KB LDS.
int choose_rows(const gh_transformer* h, int vsize, int ppt, size_t* lds_out,
                int* nry_out) {
    const TGeom& g = h->g;
    const int K = 2 * h->nn;
    int best = 0;
    size_t best_lds = 0;
    int best_nry = 0;
    // the 2-tap restriction (V-cycle) is fastest with short tiles and more
    // blocks (measured at 112^3 -> 56^3, 256 bands: 4-17 % over 14 rows)
    const int maxrows = (!h->interp && h->nn == 1) ? 7 : g.m[1];
    for (int tj = 1; tj <= std::min(g.m[1], maxrows); tj++) {
        int nry = h->interp ? (tj + 1) / 2 + 2 * h->nn + 1 : 2 * tj + 2 * K - 3;
        // interpolation: staged rows + their z-filtered rows; restriction:
        // staged rows + the tile's y-filtered rows (full staged width)
        size_t lds = h->interp ? (size_t)nry * (g.zw + g.m[2]) * vsize
                               : (size_t)(nry + tj) * g.zw * vsize;
        int thr = (tj * g.m[2] + ppt - 1) / ppt;
        if (lds > 64 * 1024 || thr > 1024) break;
        if (g.m[1] % tj == 0 || tj == g.m[1] || best == 0) {
            best = tj;
            best_lds = lds;
            best_nry = nry;
        }
    }
    *lds_out = best_lds;
    *nry_out = best_nry;
    return best;
}

template <typename V, typename S, int K>
int launch_t(const gh_transformer* h, const void* in, void* out,
             int64_t nbands, const double* phase_cd, double alpha,
             hipStream_t st) {
    TGeom g = h->g;
    size_t lds;
    int nry;
    constexpr int PPT_R = ppt_of<V, K>(false);
    constexpr int PPT_I = ppt_of<V, K>(true);
    const int ppt = h->interp ? PPT_I : PPT_R;
    int tj = choose_rows(h, sizeof(V), ppt, &lds, &nry);
    GH_REQUIRE(tj > 0, GH_ERR_UNSUPPORTED,
               "transformer: grid rows too long for one block");
    g.tj = tj;
    g.nry = nry;
    g.lx = g.m[0];
    g.nyt = (g.m[1] + tj - 1) / tj;
    g.nxc = 1;
    // split x so that there are enough blocks to fill the GPU
    while ((int64_t)nbands * g.nyt * g.nxc < 2048 && g.lx > 8) {
        g.lx = (g.lx + 1) / 2;
        g.nxc = (g.m[0] + g.lx - 1) / g.lx;
    }
    int thr = (tj * g.m[2] + ppt - 1) / ppt;
    thr = ((thr + 63) / 64) * 64;
    int64_t nblocks = nbands * (int64_t)g.nyt * g.nxc;
    GH_REQUIRE(nblocks < (1ll << 31), GH_ERR_SHAPE, "transformer: too many blocks");
    TPhase<S> ph = tphase<S>(phase_cd);
    g.yz_trivial = 1;
    for (int c = 1; c < 3; c++)
        for (int d = 0; d < 2; d++)
            if (ph.re[c][d] != (S)1 || ph.im[c][d] != (S)0) g.yz_trivial = 0;
    if (h->interp)
        hipLaunchKernelGGL((interpolate_kernel<V, S, K, PPT_I>), dim3((unsigned)nblocks),
                           dim3(thr), lds, st, (const V*)in, (V*)out, g, ph, (S)alpha);
    else if (constexpr int VEC = 16 / sizeof(V);
             VEC > 1 && g.vec == VEC && ((uintptr_t)in & 15u) == 0)
        hipLaunchKernelGGL((restrict_kernel<V, S, K, PPT_R, (VEC > 1 ? VEC : 1)>),
                           dim3((unsigned)nblocks), dim3(thr), lds, st, (const V*)in,
                           (V*)out, g, ph, (S)alpha);
    else
        hipLaunchKernelGGL((restrict_kernel<V, S, K, PPT_R>), dim3((unsigned)nblocks),
                           dim3(thr), lds, st, (const V*)in, (V*)out, g, ph, (S)alpha);
    GH_CHECK_LAUNCH();
    return GH_OK;
}

template <typename V, typename S>
int launch_k(const gh_transformer* h, const void* in, void* out, int64_t nb,
             const double* ph, double al, hipStream_t st) {
    switch (h->nn) {
        case 1: return launch_t<V, S, 2>(h, in, out, nb, ph, al, st);
        case 2: return launch_t<V, S, 4>(h, in, out, nb, ph, al, st);
        case 3: return launch_t<V, S, 6>(h, in, out, nb, ph, al, st);
        case 4: return launch_t<V, S, 8>(h, in, out, nb, ph, al, st);
    }
    set_error("transformer: nn=%d unsupported", h->nn);
    return GH_ERR_UNSUPPORTED;
}

}  // namespace

extern "C" {

int gh_transformer_create(gh_transformer** out, const int64_t* nin_c,
                          const int64_t* nout_c, int nn,
                          const int64_t* pad_cd, const int64_t* skip_cd,
                          const int32_t* pbc_c, int interpolate, int dtype) {
    GH_REQUIRE(out && nin_c && nout_c && pad_cd && skip_cd && pbc_c, GH_ERR_ARG,
               "transformer_create: null argument");
    GH_REQUIRE(nn >= 1 && nn <= 4, GH_ERR_ARG, "transformer_create: nn=%d", nn);
    GH_REQUIRE(dtype >= GH_F64 && dtype <= GH_C64, GH_ERR_ARG,
               "transformer_create: bad dtype");
    auto* h = new gh_transformer();
    h->dtype = dtype;
    h->nn = nn;
    h->interp = interpolate != 0;
    for (int c = 0; c < 3; c++) {
        h->g.n[c] = (int)nin_c[c];
        h->g.m[c] = (int)nout_c[c];
        h->g.plo[c] = (int)pad_cd[2 * c];
        h->g.skip[c] = (int)skip_cd[2 * c];
        h->g.pbc[c] = pbc_c[c] ? 1 : 0;
        const int64_t lo = pad_cd[2 * c], hi = pad_cd[2 * c + 1];
        bool ok = lo >= 0 && hi >= 0 && nin_c[c] > 0 && nout_c[c] > 0 &&
                  (!pbc_c[c] || (lo <= nin_c[c] && hi <= nin_c[c]));
        if (ok && !h->interp)
            ok = nout_c[c] == (nin_c[c] + lo + hi - 4 * nn + 3) / 2;
        if (ok && h->interp)
            ok = nout_c[c] == 2 * (nin_c[c] + lo + hi - 2 * nn + 1 + skip_cd[2 * c + 1])
                                  - skip_cd[2 * c] - skip_cd[2 * c + 1];
        if (!ok) {
            delete h;
            set_error("transformer_create: inconsistent sizes/pads on axis %d", c);
            return GH_ERR_SHAPE;
        }
    }
    // z ghost columns staged with each row: the filter taps span
    // [zmin, zmax] in interior z coordinates
    {
        const int n2 = h->g.n[2], m2 = h->g.m[2], p2 = h->g.plo[2], s2 = h->g.skip[2];
        int zmin, zmax;
        if (h->interp) {
            zmin = (s2 >> 1) - p2;
            zmax = ((m2 - 1 + s2) >> 1) + 2 * nn - 1 - p2;
        } else {
            zmin = -p2;
            zmax = 2 * (m2 - 1) + 2 * (2 * nn - 1) - p2;
        }
        const int zlo = std::max(0, -zmin), zhi = std::max(0, zmax - (n2 - 1));
        if (h->g.pbc[2] && (zlo > n2 || zhi > n2)) {
            delete h;
            set_error("transformer_create: z ghosts wider than the grid");
            return GH_ERR_SHAPE;
        }
        h->g.zlo = zlo;
        h->g.zw = zlo + n2 + zhi;
        // restriction of 8- and 4-byte types stages 16-byte vectors: the
        // ghost columns are widened to whole vectors (within one period)
        const int vec = 16 / dtype_size(dtype);
        h->g.vec = 1;
        if (!h->interp && vec > 1 && n2 % vec == 0) {
            const int zl = (zlo + vec - 1) / vec * vec;
            const int zh = (zhi + vec - 1) / vec * vec;
            if (!h->g.pbc[2] || (zl <= n2 && zh <= n2)) {
                h->g.zlo = zl;
                h->g.zw = zl + n2 + zh;
                h->g.vec = vec;
            }
        }
    }
    *out = h;
    return GH_OK;
}

int gh_transformer_destroy(gh_transformer* h) {
    delete h;
    return GH_OK;
}

int gh_transformer_apply(const gh_transformer* h, const void* in_dev,
                         void* out_dev, int64_t nbands,
                         const double* phase_cd, void* stream) {
    return gh::transformer_apply_scaled(h, in_dev, out_dev, nbands, phase_cd,
                                        1.0, (hipStream_t)stream);
}

}  // extern "C"

namespace gh {

int64_t transformer_in_points(const gh_transformer* h) {
    return (int64_t)h->g.n[0] * h->g.n[1] * h->g.n[2];
}
int64_t transformer_out_points(const gh_transformer* h) {
    return (int64_t)h->g.m[0] * h->g.m[1] * h->g.m[2];
}

int transformer_apply_scaled(const gh_transformer* h, const void* in_dev,
                             void* out_dev, int64_t nbands,
                             const double* phase_cd, double al,
                             hipStream_t st) {
    GH_REQUIRE(h, GH_ERR_ARG, "transformer_apply: null handle");
    GH_REQUIRE(nbands >= 0, GH_ERR_ARG, "transformer_apply: nbands < 0");
    if (nbands == 0) return GH_OK;  // empty arrays may carry null pointers
    GH_REQUIRE(in_dev && out_dev, GH_ERR_ARG, "transformer_apply: null argument");
    GH_REQUIRE(in_dev != out_dev, GH_ERR_ARG, "transformer_apply: in and out must differ");
    switch (h->dtype) {
        case GH_F64: return launch_k<double, double>(h, in_dev, out_dev, nbands, phase_cd, al, st);
        case GH_C128: return launch_k<cplx<double>, double>(h, in_dev, out_dev, nbands, phase_cd, al, st);
        case GH_F32: return launch_k<float, float>(h, in_dev, out_dev, nbands, phase_cd, al, st);
        case GH_C64: return launch_k<cplx<float>, float>(h, in_dev, out_dev, nbands, phase_cd, al, st);
    }
    set_error("transformer_apply: bad dtype");
    return GH_ERR_ARG;
}

}  // namespace gh
