// Finite-difference stencils on gfx950: FDOperator.apply as HIP kernels.
//
// Semantics follow the reference's Operator.apply (c/operators.c:250-321):
// a padded copy whose ghosts are periodic (times the Bloch phase) or zero
// (c/bc.c:146-308), then out = sum_p coef_p * in[g + offset_p]
// (c/bmgs/fd.c:9-33).  Here no padded copy is made: ghosts are produced on
// the fly from the interior.
//
// Axis-aligned ("central finite difference") stencils -- every Laplace on an
// orthorhombic cell -- use stencil_cfd_kernel:
//   * a block owns one band, a tile of `ty` full (y, z) rows and a run of
//     `lx` x-planes; it marches along x (the slowest axis);
//   * each thread owns PPT points of the tile and keeps a register queue of
//     the 2R+1 x-planes through them, so the x-neighbours cost no extra loads;
//   * the current plane's tile plus R ghost rows above and below lives in
//     LDS (double buffered, one barrier per plane); y- and z-neighbours are
//     read from LDS, the z-wrap is an index wrap inside the LDS row;
//   * each plane of the tile is a contiguous run of (ty + 2R) * n2 values in
//     HBM, so every load is fully coalesced.
// Other stencils (non-orthogonal cells) use stencil_generic_kernel, one
// thread per point reading its neighbours through the cache hierarchy.
#include <cstdarg>
#include <cstdlib>
#include <type_traits>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <mutex>

#include "common.h"
#include "internal.h"

namespace gh {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

constexpr int kMaxR = 6;        // Laplace n <= 6 (fd_operators.py:21-28)
constexpr int kMaxGeneric = 96; // stencil points of the generic kernel

template <typename S>
struct CfdCoefs {
    S c0;
    const double* eps;  // per band: centre coefficient c0 - eps[band] (or null)
    S cm[3][kMaxR];  // coefficient of offset -(j+1) along axis c
    S cp[3][kMaxR];  // coefficient of offset +(j+1) along axis c
};

template <typename S>
struct PhaseArgs {
    S re[3][2];
    S im[3][2];
};

struct CfdGeom {
    int n0, n1, n2;
    int pbc0, pbc1, pbc2;
    int ty, lx;     // rows per tile, planes per block
    int nyt, nxc;   // tiles along y, chunks along x
};

// Value of in[x, y, z] for an x that may lie in [-n0, 2 n0) (periodic wrap
// with phase, or zero), y and z inside the grid.
template <typename V, typename S>
__device__ __forceinline__ V load_x(const V* __restrict__ inb, int x, int yz,
                                    const CfdGeom& g, const PhaseArgs<S>& ph) {
    if (x >= 0 && x < g.n0) return inb[(size_t)x * g.n1 * g.n2 + yz];
    if (!g.pbc0) return vzero<V>();
    int d = x < 0 ? 1 : 0;
    int xx = x < 0 ? x + g.n0 : x - g.n0;
    V v = inb[(size_t)xx * g.n1 * g.n2 + yz];
    return mul_phase(v, ph.re[0][d], ph.im[0][d]);
}

constexpr int kCfdMaxThreads = 512;

// kGhostSlots: ghost values each thread prefetches per plane (1 when the
// 2R ghost rows fit in one pass of the block, else 4 plus a tail loop).
template <typename V, typename S, int R, int PPT, int kGhostSlots, bool HAS_V,
          bool HAS_SRC>
__global__ void __launch_bounds__(kCfdMaxThreads)
stencil_cfd_kernel(const V* __restrict__ in, V* __restrict__ out,
                   const S* __restrict__ vt, const V* __restrict__ src,
                   S beta, CfdCoefs<S> cf, CfdGeom g, PhaseArgs<S> ph) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V* lds = reinterpret_cast<V*>(smem_raw);

    const int n2 = g.n2;
    const int NT = blockDim.x;
    const int tiles_per_band = g.nyt * g.nxc;
    const int b = xcd_swizzle(blockIdx.x, gridDim.x);
    const int band = b / tiles_per_band;
    const int rem = b - band * tiles_per_band;
    const S c0b = cf.eps ? cf.c0 - (S)cf.eps[band] : cf.c0;
    const int xc = rem / g.nyt;
    const int yt = rem - xc * g.nyt;
    const int y0 = yt * g.ty;
    const int tye = min(g.ty, g.n1 - y0);
    const int x0 = xc * g.lx;
    const int x1 = min(x0 + g.lx, g.n0);
    const size_t plane = (size_t)g.n1 * n2;
    const size_t bstride = plane * g.n0;
    const V* __restrict__ inb = in + band * bstride;
    V* __restrict__ outb = out + band * bstride;
    const V* __restrict__ srcb = HAS_SRC ? src + band * bstride : nullptr;
    const int tile_pts = tye * n2;
    const int bufsz = (g.ty + 2 * R) * n2;
    const int nghost = 2 * R * n2;

    // own points of this thread: tile index p = tid + k * NT
    int yl[PPT], zz[PPT], yz[PPT];
    bool act[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        int p = threadIdx.x + k * NT;
        act[k] = p < tile_pts;
        int pp = act[k] ? p : 0;
        yl[k] = pp / n2;
        zz[k] = pp - yl[k] * n2;
        yz[k] = (y0 + yl[k]) * n2 + zz[k];
    }
    // ghost rows (R above, R below the tile): LDS slot, in-plane source
    // offset and kind (0 interior, 1 wrapped low, 2 wrapped high, 3 zero)
    int gl[kGhostSlots], gsrc[kGhostSlots], gkind[kGhostSlots];
#pragma unroll
    for (int i = 0; i < kGhostSlots; i++) {
        int idx = threadIdx.x + i * NT;
        gl[i] = -1;
        gsrc[i] = 0;
        gkind[i] = 3;
        if (idx < nghost) {
            int hr = idx / n2;
            int z = idx - hr * n2;
            int lrow = hr < R ? hr : hr + tye;
            int y = y0 - R + lrow;
            gl[i] = lrow * n2 + z;
            if (y >= 0 && y < g.n1) { gkind[i] = 0; gsrc[i] = y * n2 + z; }
            else if (g.pbc1) {
                gkind[i] = y < 0 ? 1 : 2;
                gsrc[i] = (y < 0 ? y + g.n1 : y - g.n1) * n2 + z;
            }
        }
    }

    auto ghost_value = [&](const V* __restrict__ pl, int src_off, int kind) {
        if (kind == 0) return pl[src_off];
        if (kind == 3) return vzero<V>();
        V v = pl[src_off];
        return kind == 1 ? mul_phase(v, ph.re[1][1], ph.im[1][1])
                         : mul_phase(v, ph.re[1][0], ph.im[1][0]);
    };

    // register queue: q[k][j] holds plane (x - R + j) at own point k
    V q[PPT][2 * R + 1];
#pragma unroll
    for (int j = 0; j < 2 * R; j++)
#pragma unroll
        for (int k = 0; k < PPT; k++)
            q[k][j] = act[k] ? load_x(inb, x0 - R + j, yz[k], g, ph) : vzero<V>();

    // Two-deep prefetch ring: slot s holds the own points of plane
    // (x + R) and the ghost rows of plane x for the iteration that uses it;
    // the loads for plane x + 2 are issued as soon as slot s is consumed, so
    // two planes of loads are in flight behind each barrier.
    V nq[2][PPT];
    V ng[2][kGhostSlots];
    auto prefetch = [&](int xn, V* nqs, V* ngs) {
#pragma unroll
        for (int k = 0; k < PPT; k++)
            nqs[k] = act[k] ? load_x(inb, xn + R, yz[k], g, ph) : vzero<V>();
        const V* pl = inb + (size_t)xn * plane;
#pragma unroll
        for (int i = 0; i < kGhostSlots; i++)
            ngs[i] = gl[i] >= 0 ? ghost_value(pl, gsrc[i], gkind[i]) : vzero<V>();
    };
    prefetch(x0, nq[0], ng[0]);
    if (x0 + 1 < x1) prefetch(x0 + 1, nq[1], ng[1]);

    auto step = [&](int x, V* nqs, V* ngs) {
        V* buf = lds + (x & 1) * bufsz;
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            q[k][2 * R] = nqs[k];
            if (act[k]) buf[(yl[k] + R) * n2 + zz[k]] = q[k][R];
        }
#pragma unroll
        for (int i = 0; i < kGhostSlots; i++)
            if (gl[i] >= 0) buf[gl[i]] = ngs[i];
        // ghosts beyond the prefetched slots (very small tiles only)
        if (nghost > kGhostSlots * NT) {
            const V* pl = inb + (size_t)x * plane;
            for (int idx = threadIdx.x + kGhostSlots * NT; idx < nghost; idx += NT) {
                int hr = idx / n2;
                int z = idx - hr * n2;
                int lrow = hr < R ? hr : hr + tye;
                int y = y0 - R + lrow;
                V v;
                if (y >= 0 && y < g.n1) v = pl[y * n2 + z];
                else if (g.pbc1) {
                    int d = y < 0 ? 1 : 0;
                    int yy = y < 0 ? y + g.n1 : y - g.n1;
                    v = mul_phase(pl[yy * n2 + z], ph.re[1][d], ph.im[1][d]);
                } else v = vzero<V>();
                buf[lrow * n2 + z] = v;
            }
        }
        if (x + 2 < x1) prefetch(x + 2, nqs, ngs);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            if (!act[k]) continue;
            const int row = yl[k] + R;
            const int z = zz[k];
            V acc = vscale(q[k][R], c0b);
#pragma unroll
            for (int j = 0; j < R; j++) {
                acc = fma_rs(cf.cm[0][j], q[k][R - 1 - j], acc);
                acc = fma_rs(cf.cp[0][j], q[k][R + 1 + j], acc);
            }
#pragma unroll
            for (int j = 0; j < R; j++) {
                acc = fma_rs(cf.cm[1][j], buf[(row - 1 - j) * n2 + z], acc);
                acc = fma_rs(cf.cp[1][j], buf[(row + 1 + j) * n2 + z], acc);
            }
            const V* brow = buf + row * n2;
#pragma unroll
            for (int j = 0; j < R; j++) {
                int zm = z - 1 - j;
                V vm;
                if (zm >= 0) vm = brow[zm];
                else if (g.pbc2) vm = mul_phase(brow[zm + n2], ph.re[2][1], ph.im[2][1]);
                else vm = vzero<V>();
                int zp = z + 1 + j;
                V vp;
                if (zp < n2) vp = brow[zp];
                else if (g.pbc2) vp = mul_phase(brow[zp - n2], ph.re[2][0], ph.im[2][0]);
                else vp = vzero<V>();
                acc = fma_rs(cf.cm[2][j], vm, acc);
                acc = fma_rs(cf.cp[2][j], vp, acc);
            }
            const size_t gi = (size_t)x * plane + yz[k];
            if (HAS_V) acc = fma_rs(vt[gi], q[k][R], acc);
            if (HAS_SRC) acc = fma_rs(beta, srcb[gi], acc);
            outb[gi] = acc;
        }
#pragma unroll
        for (int k = 0; k < PPT; k++)
#pragma unroll
            for (int j = 0; j < 2 * R; j++) q[k][j] = q[k][j + 1];
    };

    for (int x = x0; x < x1; x += 2) {
        step(x, nq[0], ng[0]);
        if (x + 1 < x1) step(x + 1, nq[1], ng[1]);
    }
}

// ------------------------------------------------- LDS-ring stencil --
// The high-throughput form of stencil_cfd_kernel.  The block owns one band,
// a tile of ty full (y, z) rows and a run of x-planes, as above, but plane
// tiles (ty + 2R rows, ghosts included) are streamed HBM -> LDS by
// global_load_lds DMA into a ring of P = R + D + 1 slots, D planes ahead of
// use, by D loader wavefronts: loader w owns the planes xp with
// (xp - x0) % D == w, so it has exactly one plane in flight and waits for it
// with vmcnt(0) just before the barrier of the iteration that first reads
// it.  Compute wavefronts never write LDS: they refill the x register queue
// from the ring, read y/z neighbours from the current plane's slot, apply
// Bloch phases / zero ghosts at read time, and store H psi.  The potential
// (and src) at their own points come from plain loads issued before the
// LDS work of the plane: the potential is shared by every band and stays in
// L2, and a DMA ring for it measured 15% slower than these loads.
constexpr int kRingMaxThreads = 1024;
constexpr int kDmaBytes = 16;              // bytes per lane per DMA op

__device__ __forceinline__ void ring_barrier() {
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_s_barrier();
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

__device__ __forceinline__ void wait_vmcnt0() {
    // s_waitcnt vmcnt(0) expcnt(7) lgkmcnt(15): only vector memory
    __builtin_amdgcn_s_waitcnt(0x0F70);
}

struct RingGeom {
    CfdGeom g;
    int nload;            // loader wavefronts (= prefetch depth D)
    int ncomp;            // compute wavefronts
    int bgroup;           // bands per scheduling group
    int nt_store;         // streaming stores of the output
    int nbands;
    int nlogical;         // logical tiles: nbands * nyt * nxc
    int rounds;           // logical tiles per block (persistent launch)
    int cplx_fast;        // complex data, trivial y / z Bloch phases
};

// Block -> (band, x chunk, y tile).  Logical ids of one XCD are contiguous
// (xcd_swizzle); within them bands come in groups of `bgroup`, all tiles of
// a group being consecutive, so the blocks that run together on an XCD are
// every tile of a few bands: the potential tile of a plane is then shared
// in L2 by the group's bands and the ghost rows of a tile are the rows its
// neighbour tile (same band, running alongside) has just brought in.
__device__ __forceinline__ void ring_block_coords(const RingGeom& rg, int L,
                                                  int& band, int& xc,
                                                  int& yt) {
    const int tiles = rg.g.nyt * rg.g.nxc;
    const int per_group = rg.bgroup * tiles;
    const int grp = L / per_group;
    const int bg = min(rg.bgroup, rg.nbands - grp * rg.bgroup);
    const int rem = L - grp * per_group;
    const int tile = rem / bg;
    band = grp * rg.bgroup + (rem - tile * bg);
    xc = tile / rg.g.nyt;
    yt = tile - xc * rg.g.nyt;
}

// LDS layout of the ring kernel (in elements of V): P psi slots of
// (ty + 2R) rows, then D + 1 slots of ty rows for the potential (as S) and
// D + 1 slots of ty rows for src, when present.
struct RingLayout {
    int psi_slot;               // elements per slot
};

// SYM: the coefficients of +j and -j are equal along every axis (every
// Laplacian); the paired real path then adds the two neighbours first and
// keeps half the coefficients live.
// One logical tile L of the ring stencil (band, x chunk, y tile): the
// block's waves split into rg.ncomp compute and rg.nload loader wavefronts;
// waves beyond those (a batched launch sized for a larger tile) only take
// part in the barriers.
template <typename V, typename S, int R, int PPT, bool HAS_V, bool HAS_SRC,
          bool SYM>
__device__ __forceinline__ void
ring_tile(const V* __restrict__ in, V* __restrict__ out,
          const S* __restrict__ vt, const V* __restrict__ src, S beta,
          const CfdCoefs<S>& cf, const RingGeom& rg, const PhaseArgs<S>& ph,
          int L, V* __restrict__ ring) {
    const CfdGeom& g = rg.g;
    const int D = rg.nload;
    const int P = R + D + 1;
    const int n2 = g.n2;
    int band, xc, yt;
    ring_block_coords(rg, L, band, xc, yt);
    const S c0b = cf.eps ? cf.c0 - (S)cf.eps[band] : cf.c0;
    const int y0 = yt * g.ty;
    const int tye = min(g.ty, g.n1 - y0);
    const int x0 = xc * g.lx;
    const int x1 = min(x0 + g.lx, g.n0);
    const size_t plane = (size_t)g.n1 * n2;
    const size_t bstride = plane * g.n0;
    const V* __restrict__ inb = in + band * bstride;
    const int slot = (g.ty + 2 * R) * n2;   // elements per ring slot
    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int xend = x1 + R;                 // planes [x0, xend) go through LDS

    if (wave >= rg.ncomp + rg.nload) {
        for (int x = x0; x < x1; x++) ring_barrier();
        return;
    }
    if (wave >= rg.ncomp) {
        // ---------------------------------------------------- loader --
        const int w = wave - rg.ncomp;
        constexpr int EPL = kDmaBytes / (int)sizeof(V);   // elements per lane
        constexpr int CH = 64 * EPL;                       // elements per op
        // three row segments of the tile: rows below 0, inside, above n1
        const int ylo = y0 - R, yhi = y0 + tye + R;        // [ylo, yhi)
        int seg_g[3], seg_l[3], seg_n[3];
        {
            int a0 = ylo, a1 = min(0, yhi);                // wrapped low rows
            int b0 = max(ylo, 0), b1 = min(yhi, g.n1);     // interior rows
            int c0 = max(ylo, g.n1), c1 = yhi;             // wrapped high rows
            seg_n[0] = (g.pbc1 && a1 > a0) ? (a1 - a0) * n2 : 0;
            seg_g[0] = (a0 + g.n1) * n2;
            seg_l[0] = 0;
            seg_n[1] = (b1 - b0) * n2;
            seg_g[1] = b0 * n2;
            seg_l[1] = (b0 - ylo) * n2;
            seg_n[2] = (g.pbc1 && c1 > c0) ? (c1 - c0) * n2 : 0;
            seg_g[2] = (c0 - g.n1) * n2;
            seg_l[2] = (c0 - ylo) * n2;
        }
        auto issue = [&](int xp) {
            int xs = xp;
            if (xs >= g.n0) {
                if (!g.pbc0) return;
                xs -= g.n0;
            }
            const V* src_pl = inb + (size_t)xs * plane;
            V* dst = ring + ((xp - x0) % P) * slot;
            for (int sgi = 0; sgi < 3; sgi++) {
                const int n = seg_n[sgi];
                for (int off = 0; off < n; off += CH) {
                    const int e = off + lane * EPL;
                    if (e < n)
                        __builtin_amdgcn_global_load_lds(
                            (const void*)(src_pl + seg_g[sgi] + e),
                            (__attribute__((address_space(3))) void*)(dst + seg_l[sgi] + off),
                            kDmaBytes, 0, 0);
                }
            }
        };
        for (int xp = x0 + w; xp < min(x0 + R + D, xend); xp += D) issue(xp);
        for (int x = x0; x < x1; x++) {
            // plane x+R (queue refill) is first read now; the planes x of
            // the first R iterations were issued in the prologue and are
            // first read now as well
            if (((x + R - x0) % D) == w || (x - x0 < R && ((x - x0) % D) == w))
                wait_vmcnt0();
            ring_barrier();
            const int xp = x + R + D;
            if (xp < xend && ((xp - x0) % D) == w) issue(xp);
        }
        return;
    }

    // ------------------------------------------------------ compute --
    V* __restrict__ outb = out + band * bstride;
    if constexpr (!is_complex<V>::value && PPT == 4) {
        // FP32 data: each thread owns four z-adjacent points (one 16-byte
        // vector) of one row (n2 % 4 == 0).  The same 1024 threads then
        // cover twice the rows of the two-point path, so a tile carries
        // fewer ghost rows per row, and every LDS read, potential / src
        // load and store moves four points.  The z-window z - 4 .. z + 7
        // is three vectors whose index wraps modulo n2 / 4 (R <= 4).
        static_assert(R <= 4, "four-point path: R <= 4");
        typedef V V4 __attribute__((ext_vector_type(4)));
        const int nq = n2 / 4;                       // vectors per row
        const int pidx = threadIdx.x;                // < ncomp * 64
        const bool act = pidx < tye * nq;
        const int pp = act ? pidx : 0;
        const int yl = pp / nq;
        const int zq = pp - yl * nq;                 // vector index in row
        const int yz = (y0 + yl) * n2 + 4 * zq;
        const int zqm = zq - 1, zqp = zq + 1;
        const bool zout_m = zqm < 0 && !g.pbc2;
        const bool zout_p = zqp >= nq && !g.pbc2;
        const int zoff_m = zqm < 0 ? zqm + nq : zqm;
        const int zoff_p = zqp >= nq ? zqp - nq : zqp;
        if (!g.pbc1) {
            for (int r = 0; r < 2 * R; r++) {
                const int lrow = r < R ? r : tye + r;
                const int y = y0 - R + lrow;
                if (y >= 0 && y < g.n1) continue;
                for (int sl = 0; sl < P; sl++)
                    for (int e = pidx; e < n2; e += rg.ncomp * 64)
                        ring[sl * slot + lrow * n2 + e] = (V)0;
            }
        }
        V4 q[2 * R + 1];
#pragma unroll
        for (int j = 0; j < 2 * R; j++)
            q[j] = V4{load_x(inb, x0 - R + j, yz, g, ph), load_x(inb, x0 - R + j, yz + 1, g, ph),
                      load_x(inb, x0 - R + j, yz + 2, g, ph),
                      load_x(inb, x0 - R + j, yz + 3, g, ph)};
        S k0 = c0b;
        S km[3][R], kp[3][R];
        if (SYM) asm volatile("" : "+v"(k0));
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int j = 0; j < R; j++) {
                km[c][j] = cf.cm[c][j];
                kp[c][j] = SYM ? km[c][j] : cf.cp[c][j];
                if (SYM) asm volatile("" : "+v"(km[c][j]));
            }
        auto march = [&](auto zper) {
            constexpr bool ZPER = decltype(zper)::value;
            for (int x = x0; x < x1; x++) {
                ring_barrier();
                {
                    const int xp = x + R;
                    const V4* sl = reinterpret_cast<const V4*>(ring + ((xp - x0) % P) * slot);
                    V4 v = sl[(yl + R) * nq + zq];
                    if (xp >= g.n0 && !g.pbc0) v = V4{0, 0, 0, 0};
                    q[2 * R] = v;
                }
                const size_t gi = (size_t)x * plane + yz;
                V4 vv, ss;
                if (HAS_V) vv = *reinterpret_cast<const V4*>(vt + gi);
                if (HAS_SRC) ss = *reinterpret_cast<const V4*>(src + band * bstride + gi);
                const V4* sl = reinterpret_cast<const V4*>(ring + ((x - x0) % P) * slot);
                V4 acc = q[R] * k0;
#pragma unroll
                for (int j = 0; j < R; j++) {
                    if constexpr (SYM) {
                        acc = (q[R - 1 - j] + q[R + 1 + j]) * km[0][j] + acc;
                    } else {
                        acc = q[R - 1 - j] * km[0][j] + acc;
                        acc = q[R + 1 + j] * kp[0][j] + acc;
                    }
                }
                const int row = yl + R;
#pragma unroll
                for (int j = 0; j < R; j++) {
                    const V4 m = sl[(row - 1 - j) * nq + zq];
                    const V4 pv = sl[(row + 1 + j) * nq + zq];
                    if constexpr (SYM) {
                        acc = (m + pv) * km[1][j] + acc;
                    } else {
                        acc = m * km[1][j] + acc;
                        acc = pv * kp[1][j] + acc;
                    }
                }
                // z-window: w[0] = vector zq - 1, w[1] = zq, w[2] = zq + 1
                V4 w[3];
                const V4* brow = sl + row * nq;
                w[0] = brow[zoff_m];
                w[1] = q[R];
                w[2] = brow[zoff_p];
                if constexpr (!ZPER) {
                    if (zout_m) w[0] = V4{0, 0, 0, 0};
                    if (zout_p) w[2] = V4{0, 0, 0, 0};
                }
                // element e of the window: vector e >> 2, lane e & 3; point
                // i of this thread is element 4 + i
                auto we = [&](int e) { return w[e >> 2][e & 3]; };
                S a4[4] = {acc.x, acc.y, acc.z, acc.w};
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < R; j++) {
                        if constexpr (SYM) {
                            a4[i] = fma_rs(km[2][j], we(3 + i - j) + we(5 + i + j), a4[i]);
                        } else {
                            a4[i] = fma_rs(km[2][j], we(3 + i - j), a4[i]);
                            a4[i] = fma_rs(kp[2][j], we(5 + i + j), a4[i]);
                        }
                    }
                acc = V4{a4[0], a4[1], a4[2], a4[3]};
                if (HAS_V) acc = vv * q[R] + acc;
                if (HAS_SRC) acc = ss * beta + acc;
                if (act) {
                    V4* o = reinterpret_cast<V4*>(outb + gi);
                    if (rg.nt_store) __builtin_nontemporal_store(acc, o);
                    else *o = acc;
                }
#pragma unroll
                for (int j = 0; j < 2 * R; j++) q[j] = q[j + 1];
            }
        };
        if (g.pbc2) march(std::true_type{});
        else march(std::false_type{});
        return;
    }
    if constexpr (!is_complex<V>::value && PPT == 2) {
        // Real data: each thread owns two z-adjacent points (z, z + 1) of
        // one row (n2 is even: rows are whole 16-byte multiples), so every
        // LDS read, potential / src load and store moves both points at
        // once.  Wrapped neighbours need no phase; those beyond a
        // non-periodic face are zero.  The z-window z - 2P .. z + 1 + 2P is
        // read as pairs whose index wraps modulo n2 / 2.
        typedef V V2 __attribute__((ext_vector_type(2)));
        const int np = n2 / 2;                       // pairs per row
        const int pidx = threadIdx.x;                // < ncomp * 64
        const bool act = pidx < tye * np;
        const int pp = act ? pidx : 0;
        const int yl = pp / np;
        const int zp = pp - yl * np;                 // pair index in row
        const int yz = (y0 + yl) * n2 + 2 * zp;
        constexpr int NPW = (R + 1) / 2;             // window pairs per side
        int zoff[2 * NPW];                           // wrapped pair index
        bool zout[2 * NPW];                          // beyond a non-pbc face
#pragma unroll
        for (int i = 0; i < NPW; i++) {
            int zm = zp - 1 - i, zq = zp + 1 + i;
            zout[i] = zm < 0 && !g.pbc2;
            zout[NPW + i] = zq >= np && !g.pbc2;
            zoff[i] = zm < 0 ? zm + np : zm;
            zoff[NPW + i] = zq >= np ? zq - np : zq;
        }
        // rows of the tile beyond a non-periodic y face are never written by
        // the loaders: zero them once in every slot, so the loop reads them
        // without a test (the first barrier of the loop orders this)
        if (!g.pbc1) {
            for (int r = 0; r < 2 * R; r++) {
                const int lrow = r < R ? r : tye + r;
                const int y = y0 - R + lrow;
                if (y >= 0 && y < g.n1) continue;
                for (int sl = 0; sl < P; sl++)
                    for (int e = pidx; e < n2; e += rg.ncomp * 64)
                        ring[sl * slot + lrow * n2 + e] = (V)0;
            }
        }
        V2 q[2 * R + 1];
#pragma unroll
        for (int j = 0; j < 2 * R; j++) {
            q[j].x = load_x(inb, x0 - R + j, yz, g, ph);
            q[j].y = load_x(inb, x0 - R + j, yz + 1, g, ph);
        }
        // the coefficients live in VGPRs (uniform values; the scalar file
        // is full in this loop and would spill them through v_readlane)
        // (symmetric stencils only: with both signs pinned the register
        // file overflows)
        S k0 = c0b;
        S km[3][R], kp[3][R];
        if (SYM) asm volatile("" : "+v"(k0));
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int j = 0; j < R; j++) {
                km[c][j] = cf.cm[c][j];
                kp[c][j] = SYM ? km[c][j] : cf.cp[c][j];
                if (SYM) asm volatile("" : "+v"(km[c][j]));
            }
        auto march = [&](auto zper) {
            constexpr bool ZPER = decltype(zper)::value;
            for (int x = x0; x < x1; x++) {
                ring_barrier();
                {
                    const int xp = x + R;
                    const V2* sl = reinterpret_cast<const V2*>(ring + ((xp - x0) % P) * slot);
                    V2 v = sl[(yl + R) * np + zp];
                    if (xp >= g.n0 && !g.pbc0) v = V2{(V)0, (V)0};
                    q[2 * R] = v;
                }
                const size_t gi = (size_t)x * plane + yz;
                V2 vv, ss;
                if (HAS_V) vv = *reinterpret_cast<const V2*>(vt + gi);
                if (HAS_SRC) ss = *reinterpret_cast<const V2*>(src + band * bstride + gi);
                const V2* sl = reinterpret_cast<const V2*>(ring + ((x - x0) % P) * slot);
                V2 acc = q[R] * k0;
#pragma unroll
                for (int j = 0; j < R; j++) {
                    if constexpr (SYM) {
                        acc = (q[R - 1 - j] + q[R + 1 + j]) * km[0][j] + acc;
                    } else {
                        acc = q[R - 1 - j] * km[0][j] + acc;
                        acc = q[R + 1 + j] * kp[0][j] + acc;
                    }
                }
                const int row = yl + R;
#pragma unroll
                for (int j = 0; j < R; j++) {
                    const V2 m = sl[(row - 1 - j) * np + zp];
                    const V2 pv = sl[(row + 1 + j) * np + zp];
                    if constexpr (SYM) {
                        acc = (m + pv) * km[1][j] + acc;
                    } else {
                        acc = m * km[1][j] + acc;
                        acc = pv * kp[1][j] + acc;
                    }
                }
                // z-window: w[0] = pair zp - NPW ... w[2 NPW] = pair zp + NPW
                V2 w[2 * NPW + 1];
                w[NPW] = q[R];
                const V2* brow = sl + row * np;
#pragma unroll
                for (int i = 0; i < NPW; i++) {
                    V2 lo = brow[zoff[i]], hi = brow[zoff[NPW + i]];
                    if constexpr (!ZPER) {
                        if (zout[i]) lo = V2{(V)0, (V)0};
                        if (zout[NPW + i]) hi = V2{(V)0, (V)0};
                    }
                    w[NPW - 1 - i] = lo;
                    w[NPW + 1 + i] = hi;
                }
                // element e of the window: pair e >> 1, lane e & 1; element
                // 2 NPW (pair NPW, lane 0) is the point z
                auto we = [&](int e) { return (e & 1) ? w[e >> 1].y : w[e >> 1].x; };
#pragma unroll
                for (int j = 0; j < R; j++) {
                    const int c = 2 * NPW;
                    if constexpr (SYM) {
                        acc.x = fma_rs(km[2][j], we(c - 1 - j) + we(c + 1 + j), acc.x);
                        acc.y = fma_rs(km[2][j], we(c - j) + we(c + 2 + j), acc.y);
                    } else {
                        acc.x = fma_rs(km[2][j], we(c - 1 - j), acc.x);
                        acc.x = fma_rs(kp[2][j], we(c + 1 + j), acc.x);
                        acc.y = fma_rs(km[2][j], we(c - j), acc.y);
                        acc.y = fma_rs(kp[2][j], we(c + 2 + j), acc.y);
                    }
                }
                if (HAS_V) acc = vv * q[R] + acc;
                if (HAS_SRC) acc = ss * beta + acc;
                if (act) {
                    V2* o = reinterpret_cast<V2*>(outb + gi);
                    if (rg.nt_store) __builtin_nontemporal_store(acc, o);
                    else *o = acc;
                }
#pragma unroll
                for (int j = 0; j < 2 * R; j++) q[j] = q[j + 1];
            }
        };
        if (g.pbc2) march(std::true_type{});
        else march(std::false_type{});
        return;
    }
    if constexpr (is_complex<V>::value && PPT == 2) {
      {
        // Complex data whose y and z Bloch phases are 1: the stencil's real
        // coefficients act on real and imaginary parts alike, so this is the
        // paired path above on (re, im) pairs -- a thread owns two
        // z-adjacent points as one 4-vector (re0, im0, re1, im1).  Only the
        // x-wrap phase (uniform per plane) is applied, at the queue refill.
        typedef S V4 __attribute__((ext_vector_type(4)));
        typedef S S2 __attribute__((ext_vector_type(2)));
        const int np = n2 / 2;
        const int pidx = threadIdx.x;
        const bool act = pidx < tye * np;
        const int pp = act ? pidx : 0;
        const int yl = pp / np;
        const int zp = pp - yl * np;
        const int yz = (y0 + yl) * n2 + 2 * zp;
        constexpr int NPW = (R + 1) / 2;
        int zoff[2 * NPW];
        bool zout[2 * NPW];
#pragma unroll
        for (int i = 0; i < NPW; i++) {
            int zm = zp - 1 - i, zq = zp + 1 + i;
            zout[i] = zm < 0 && !g.pbc2;
            zout[NPW + i] = zq >= np && !g.pbc2;
            zoff[i] = zm < 0 ? zm + np : zm;
            zoff[NPW + i] = zq >= np ? zq - np : zq;
        }
        if (!g.pbc1) {
            for (int r = 0; r < 2 * R; r++) {
                const int lrow = r < R ? r : tye + r;
                const int y = y0 - R + lrow;
                if (y >= 0 && y < g.n1) continue;
                for (int sl = 0; sl < P; sl++)
                    for (int e = pidx; e < n2; e += rg.ncomp * 64)
                        ring[sl * slot + lrow * n2 + e] = vzero<V>();
            }
        }
        // (symmetric stencils only: with both signs pinned the register
        // file overflows)
        S k0 = c0b;
        S km[3][R], kp[3][R];
        if (SYM) asm volatile("" : "+v"(k0));
#pragma unroll
        for (int c = 0; c < 3; c++)
#pragma unroll
            for (int j = 0; j < R; j++) {
                km[c][j] = cf.cm[c][j];
                kp[c][j] = SYM ? km[c][j] : cf.cp[c][j];
                if (SYM) asm volatile("" : "+v"(km[c][j]));
            }
        auto pack = [&](V a, V b) { return V4{a.re, a.im, b.re, b.im}; };
        V4 q[2 * R + 1];
#pragma unroll
        for (int j = 0; j < 2 * R; j++)
            q[j] = pack(load_x(inb, x0 - R + j, yz, g, ph),
                        load_x(inb, x0 - R + j, yz + 1, g, ph));
        const S pr = ph.re[0][0], pi = ph.im[0][0];   // x-wrap high -> low
        for (int x = x0; x < x1; x++) {
            ring_barrier();
            {
                const int xp = x + R;
                const V4* sl = reinterpret_cast<const V4*>(ring + ((xp - x0) % P) * slot);
                V4 v = sl[(yl + R) * np + zp];
                if (xp >= g.n0) {
                    if (!g.pbc0) v = V4{0, 0, 0, 0};
                    else v = V4{v.x * pr - v.y * pi, v.x * pi + v.y * pr,
                                v.z * pr - v.w * pi, v.z * pi + v.w * pr};
                }
                q[2 * R] = v;
            }
            const size_t gi = (size_t)x * plane + yz;
            S2 vv;
            V4 ss;
            if (HAS_V) vv = *reinterpret_cast<const S2*>(vt + gi);
            if (HAS_SRC) ss = *reinterpret_cast<const V4*>(src + band * bstride + gi);
            const V4* sl = reinterpret_cast<const V4*>(ring + ((x - x0) % P) * slot);
            V4 acc = q[R] * k0;
#pragma unroll
            for (int j = 0; j < R; j++) {
                if constexpr (SYM) {
                    acc = (q[R - 1 - j] + q[R + 1 + j]) * km[0][j] + acc;
                } else {
                    acc = q[R - 1 - j] * km[0][j] + acc;
                    acc = q[R + 1 + j] * kp[0][j] + acc;
                }
            }
            const int row = yl + R;
#pragma unroll
            for (int j = 0; j < R; j++) {
                const V4 m = sl[(row - 1 - j) * np + zp];
                const V4 pv = sl[(row + 1 + j) * np + zp];
                if constexpr (SYM) {
                    acc = (m + pv) * km[1][j] + acc;
                } else {
                    acc = m * km[1][j] + acc;
                    acc = pv * kp[1][j] + acc;
                }
            }
            V4 w[2 * NPW + 1];
            w[NPW] = q[R];
            const V4* brow = sl + row * np;
#pragma unroll
            for (int i = 0; i < NPW; i++) {
                V4 lo = brow[zoff[i]], hi = brow[zoff[NPW + i]];
                if (zout[i]) lo = V4{0, 0, 0, 0};
                if (zout[NPW + i]) hi = V4{0, 0, 0, 0};
                w[NPW - 1 - i] = lo;
                w[NPW + 1 + i] = hi;
            }
            // complex element e of the window (pair e >> 1, point e & 1)
            auto we = [&](int e) {
                return (e & 1) ? S2{w[e >> 1].z, w[e >> 1].w} : S2{w[e >> 1].x, w[e >> 1].y};
            };
            S2 a0 = S2{acc.x, acc.y}, a1 = S2{acc.z, acc.w};
#pragma unroll
            for (int j = 0; j < R; j++) {
                const int c = 2 * NPW;
                if constexpr (SYM) {
                    a0 = (we(c - 1 - j) + we(c + 1 + j)) * km[2][j] + a0;
                    a1 = (we(c - j) + we(c + 2 + j)) * km[2][j] + a1;
                } else {
                    a0 = we(c - 1 - j) * km[2][j] + a0;
                    a0 = we(c + 1 + j) * kp[2][j] + a0;
                    a1 = we(c - j) * km[2][j] + a1;
                    a1 = we(c + 2 + j) * kp[2][j] + a1;
                }
            }
            acc = V4{a0.x, a0.y, a1.x, a1.y};
            if (HAS_V) acc = V4{vv.x, vv.x, vv.y, vv.y} * q[R] + acc;
            if (HAS_SRC) acc = ss * beta + acc;
            if (act) {
                V4* o = reinterpret_cast<V4*>(outb + gi);
                if (rg.nt_store) __builtin_nontemporal_store(acc, o);
                else *o = acc;
            }
#pragma unroll
            for (int j = 0; j < 2 * R; j++) q[j] = q[j + 1];
        }
        return;
      }
    } else {
    // general path: complex data with Bloch phases (one point per thread)
    const int NT = rg.ncomp * 64;
    const int tile_pts = tye * n2;
    int yl[PPT], zz[PPT], yz[PPT];
    bool act[PPT];
#pragma unroll
    for (int k = 0; k < PPT; k++) {
        int p = threadIdx.x + k * NT;
        act[k] = p < tile_pts;
        int pp = act[k] ? p : 0;
        yl[k] = pp / n2;
        zz[k] = pp - yl[k] * n2;
        yz[k] = (y0 + yl[k]) * n2 + zz[k];
    }
    // y-ghost rows of the slot: R below (local rows 0..R-1) and R above
    // (local rows R+tye..); kind per row offset j: 0 interior, 1 low wrap,
    // 2 high wrap, 3 zero
    auto row_kind = [&](int lrow) {
        int y = y0 - R + lrow;
        if (y >= 0 && y < g.n1) return 0;
        if (!g.pbc1) return 3;
        return y < 0 ? 1 : 2;
    };
    auto slot_val = [&](const V* sl, int lrow, int z) {
        int kind = row_kind(lrow);
        if (kind == 0) return sl[lrow * n2 + z];
        if (kind == 3) return vzero<V>();
        V v = sl[lrow * n2 + z];
        return kind == 1 ? mul_phase(v, ph.re[1][1], ph.im[1][1])
                         : mul_phase(v, ph.re[1][0], ph.im[1][0]);
    };

    V q[PPT][2 * R + 1];
#pragma unroll
    for (int j = 0; j < 2 * R; j++)
#pragma unroll
        for (int k = 0; k < PPT; k++)
            q[k][j] = act[k] ? load_x(inb, x0 - R + j, yz[k], g, ph) : vzero<V>();
    for (int x = x0; x < x1; x++) {
        ring_barrier();
        // queue refill: plane x + R (wrapped with phase, or zero)
        {
            const int xp = x + R;
            const V* sl = ring + ((xp - x0) % P) * slot;
            const bool wrap = xp >= g.n0;
            const bool zero = wrap && !g.pbc0;
#pragma unroll
            for (int k = 0; k < PPT; k++) {
                V v = sl[(yl[k] + R) * n2 + zz[k]];
                if (zero) v = vzero<V>();
                else if (wrap) v = mul_phase(v, ph.re[0][0], ph.im[0][0]);
                q[k][2 * R] = act[k] ? v : vzero<V>();
            }
        }
        // potential and src at the own points of plane x: plain loads
        // (L2-resident for the potential), issued before the LDS work
        S vv[PPT];
        V ss[PPT];
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            const size_t gi = (size_t)x * plane + yz[k];
            vv[k] = HAS_V ? vt[gi] : (S)0;
            ss[k] = HAS_SRC ? src[band * bstride + gi] : vzero<V>();
        }
        const V* sl = ring + ((x - x0) % P) * slot;
#pragma unroll
        for (int k = 0; k < PPT; k++) {
            if (!act[k]) continue;
            const int row = yl[k] + R;
            const int z = zz[k];
            V acc = vscale(q[k][R], c0b);
#pragma unroll
            for (int j = 0; j < R; j++) {
                acc = fma_rs(cf.cm[0][j], q[k][R - 1 - j], acc);
                acc = fma_rs(cf.cp[0][j], q[k][R + 1 + j], acc);
            }
#pragma unroll
            for (int j = 0; j < R; j++) {
                acc = fma_rs(cf.cm[1][j], slot_val(sl, row - 1 - j, z), acc);
                acc = fma_rs(cf.cp[1][j], slot_val(sl, row + 1 + j, z), acc);
            }
            const V* brow = sl + row * n2;
#pragma unroll
            for (int j = 0; j < R; j++) {
                int zm = z - 1 - j;
                V vm;
                if (zm >= 0) vm = brow[zm];
                else if (g.pbc2) vm = mul_phase(brow[zm + n2], ph.re[2][1], ph.im[2][1]);
                else vm = vzero<V>();
                int zp = z + 1 + j;
                V vp;
                if (zp < n2) vp = brow[zp];
                else if (g.pbc2) vp = mul_phase(brow[zp - n2], ph.re[2][0], ph.im[2][0]);
                else vp = vzero<V>();
                acc = fma_rs(cf.cm[2][j], vm, acc);
                acc = fma_rs(cf.cp[2][j], vp, acc);
            }
            if (HAS_V) acc = fma_rs(vv[k], q[k][R], acc);
            if (HAS_SRC) acc = fma_rs(beta, ss[k], acc);
            if (rg.nt_store) store_nt(outb + (size_t)x * plane + yz[k], acc);
            else outb[(size_t)x * plane + yz[k]] = acc;
        }
#pragma unroll
        for (int k = 0; k < PPT; k++)
#pragma unroll
            for (int j = 0; j < 2 * R; j++) q[k][j] = q[k][j + 1];
    }
    }  // general path
}

template <typename V, typename S, int R, int PPT, bool HAS_V, bool HAS_SRC,
          bool SYM = false>
__global__ void __launch_bounds__(kRingMaxThreads)
stencil_ring_kernel(const V* __restrict__ in, V* __restrict__ out,
                    const S* __restrict__ vt, const V* __restrict__ src,
                    S beta, CfdCoefs<S> cf, RingGeom rg, PhaseArgs<S> ph,
                    RingLayout lay) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    V* ring = reinterpret_cast<V*>(smem_raw);
    const int Lbase = xcd_swizzle(blockIdx.x, gridDim.x);
    // persistent blocks take logical tiles Lbase, Lbase + gridDim.x, ...;
    // every wave of the block sees the same sequence
    for (int rd = 0; rd < rg.rounds; rd++) {
        const int L = Lbase + rd * (int)gridDim.x;
        if (L >= rg.nlogical) break;
        if (rd > 0) ring_barrier();   // last reads of the previous tile done
        ring_tile<V, S, R, PPT, HAS_V, HAS_SRC, SYM>(in, out, vt, src, beta, cf, rg, ph, L,
                                                      ring);
    }
}

// Several independent grids (cells) in one persistent launch.  Cell c owns
// the logical tiles [lbeg_c, lbeg_c + nlogical_c) of a global numbering;
// its tile geometry (rows, prefetch depth, x chunks, wave split) comes from
// the cell table, its arrays from the pointer table.  The block is sized
// for the largest tile; LDS for the largest ring.
constexpr int kMaxBatchCells = 8;

template <typename S>
struct RingCell {
    CfdCoefs<S> cf;
    RingGeom rg;
    PhaseArgs<S> ph;
    int lbeg;
};

template <typename V, typename S>
struct BatchPtrs {
    const V* in[kMaxBatchCells];
    V* out[kMaxBatchCells];
    const S* vt[kMaxBatchCells];
};

// Tiles are handed out in order (largest first) from a device counter, so
// blocks that draw short tiles simply draw more of them.  The last block
// to finish resets the counters for the next launch (no memset node, so
// the launch can be captured in a graph); launches of one batch must not
// overlap in time.
template <typename V, typename S, int R, int PPT, bool HAS_V, bool SYM>
__global__ void __launch_bounds__(kRingMaxThreads)
stencil_ring_batched_kernel(const RingCell<S>* __restrict__ cells, BatchPtrs<V, S> ptr,
                            int ncells, int ntotal, int* __restrict__ counters) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    __shared__ int s_tile;
    V* ring = reinterpret_cast<V*>(smem_raw);
    for (int rd = 0;; rd++) {
        if (rd > 0) ring_barrier();   // the previous tile's LDS reads are done
        if (threadIdx.x == 0) s_tile = atomicAdd(&counters[0], 1);
        __syncthreads();
        const int L = s_tile;
        if (L >= ntotal) break;
        int c = 0;
        for (int k = 1; k < ncells; k++)
            if (L >= cells[k].lbeg) c = k;
        const RingCell<S>& cell = cells[c];
        ring_tile<V, S, R, PPT, HAS_V, false, SYM>(ptr.in[c], ptr.out[c], ptr.vt[c], nullptr,
                                                    (S)0, cell.cf, cell.rg, cell.ph,
                                                    L - cell.lbeg, ring);
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&counters[1], 1) == (int)gridDim.x - 1) {
            atomicExch(&counters[0], 0);
            atomicExch(&counters[1], 0);
        }
    }
}

// ------------------------------------------------------ generic stencil --
template <typename S>
struct GenericArgs {
    int np;
    int n0, n1, n2;
    int pbc0, pbc1, pbc2;
    S coef[kMaxGeneric];
    signed char off[kMaxGeneric][3];
};

template <typename V, typename S>
__device__ __forceinline__ V fetch_wrapped(const V* __restrict__ inb, int x,
                                           int y, int z,
                                           const GenericArgs<S>& a,
                                           const PhaseArgs<S>& ph) {
    S pre = 1, pim = 0;
    int c[3] = {x, y, z};
    const int n[3] = {a.n0, a.n1, a.n2};
    const int pbc[3] = {a.pbc0, a.pbc1, a.pbc2};
    for (int ax = 0; ax < 3; ax++) {
        int d = -1;
        if (c[ax] < 0) { c[ax] += n[ax]; d = 1; }
        else if (c[ax] >= n[ax]) { c[ax] -= n[ax]; d = 0; }
        if (d >= 0) {
            if (!pbc[ax]) return vzero<V>();
            S r = pre * ph.re[ax][d] - pim * ph.im[ax][d];
            S i = pre * ph.im[ax][d] + pim * ph.re[ax][d];
            pre = r;
            pim = i;
        }
    }
    V v = inb[((size_t)c[0] * a.n1 + c[1]) * a.n2 + c[2]];
    return mul_phase(v, pre, pim);
}

template <typename V, typename S, bool HAS_V, bool HAS_SRC>
__global__ void __launch_bounds__(256)
stencil_generic_kernel(const V* __restrict__ in, V* __restrict__ out,
                       const S* __restrict__ vt, const V* __restrict__ src,
                       S beta, S gamma, GenericArgs<S> a, PhaseArgs<S> ph,
                       int64_t total) {
    const size_t npts = (size_t)a.n0 * a.n1 * a.n2;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        size_t band = t / npts;
        size_t gi = t - band * npts;
        int z = gi % a.n2;
        int y = (gi / a.n2) % a.n1;
        int x = gi / ((size_t)a.n1 * a.n2);
        const V* inb = in + band * npts;
        V acc = vzero<V>();
        for (int p = 0; p < a.np; p++)
            acc = fma_rs(a.coef[p],
                         fetch_wrapped(inb, x + a.off[p][0], y + a.off[p][1],
                                       z + a.off[p][2], a, ph),
                         acc);
        if (gamma != (S)0) acc = fma_rs(gamma, inb[gi], acc);
        if (HAS_V) acc = fma_rs(vt[gi], inb[gi], acc);
        if (HAS_SRC) acc = fma_rs(beta, src[t], acc);
        out[t] = acc;
    }
}

// Gauss-Seidel sweep (relax_method 1) for stencils whose offsets lie on
// the grid axes: a point's neighbours before it in C order lie on earlier
// hyperplanes i0 + i1 + i2 = const and those after it on later ones, so
// updating the hyperplanes in order, each in one launch, reproduces the
// reference's in-place C-order sweep; wrapped neighbours read `f0`, the
// sweep's starting values (the reference fills its ghosts before the
// sweep).  The coefficients are summed in stencil order.
template <typename S>
__global__ void __launch_bounds__(256)
gs_plane_kernel(S* __restrict__ f, const S* __restrict__ f0, const S* __restrict__ src,
                GenericArgs<S> a, int icentre, int plane, S inv) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    const int i0 = t / a.n1;
    const int i1 = t - i0 * a.n1;
    if (i0 >= a.n0) return;
    const int i2 = plane - i0 - i1;
    if (i2 < 0 || i2 >= a.n2) return;
    S x = 0;
    for (int p = 0; p < a.np; p++) {
        if (p == icentre) continue;
        int q0 = i0 + a.off[p][0], q1 = i1 + a.off[p][1], q2 = i2 + a.off[p][2];
        const bool in = q0 >= 0 && q0 < a.n0 && q1 >= 0 && q1 < a.n1 && q2 >= 0 && q2 < a.n2;
        S v;
        if (in) {
            v = f[((size_t)q0 * a.n1 + q1) * a.n2 + q2];
        } else {
            const bool ok = (q0 >= 0 && q0 < a.n0 ? true : a.pbc0) &&
                            (q1 >= 0 && q1 < a.n1 ? true : a.pbc1) &&
                            (q2 >= 0 && q2 < a.n2 ? true : a.pbc2);
            if (ok) {
                q0 = q0 < 0 ? q0 + a.n0 : q0 >= a.n0 ? q0 - a.n0 : q0;
                q1 = q1 < 0 ? q1 + a.n1 : q1 >= a.n1 ? q1 - a.n1 : q1;
                q2 = q2 < 0 ? q2 + a.n2 : q2 >= a.n2 ? q2 - a.n2 : q2;
                v = f0[((size_t)q0 * a.n1 + q1) * a.n2 + q2];
            } else {
                v = 0;
            }
        }
        x += v * a.coef[p];
    }
    const size_t i = ((size_t)i0 * a.n1 + i1) * a.n2 + i2;
    f[i] = (src[i] - x) * inv;
}

}  // namespace gh

// ------------------------------------------------------------- handle --
// (declared before use by the internal launcher below)
struct gh_stencil {
    int dtype;
    int64_t n[3];
    int pbc[3];
    int R;          // max |offset| per axis (padding of the reference)
    bool cfd;
    double diag;
    // cfd representation
    double c0;
    double cm[3][gh::kMaxR];
    double cp[3][gh::kMaxR];
    // generic representation
    std::vector<double> coefs;
    std::vector<int64_t> offs;
    int ty_req = 0, lx_req = 0, depth_req = 0;
    int bgroup_req = 0;      // bands per scheduling group (GPAW_AMD_BGROUP)
    bool reg_only = false;   // GPAW_AMD_STENCIL=reg selects the register kernel
    int persist = 0;         // persistent ring launch: 1 one block per CU,
                             // n >= 8 a grid of n blocks (GPAW_AMD_PERSIST)
};

using namespace gh;

namespace {

template <typename S>
PhaseArgs<S> phase_args(const double* phase_cd) {
    Phases p = make_phases(phase_cd);
    PhaseArgs<S> a;
    for (int c = 0; c < 3; c++)
        for (int d = 0; d < 2; d++) {
            a.re[c][d] = (S)p.re[c][d];
            a.im[c][d] = (S)p.im[c][d];
        }
    return a;
}

struct TileChoice {
    int ty, lx, ppt, threads;
    size_t lds;
};

TileChoice choose_tile(const gh_stencil* h, int vsize, int ppt) {
    const int n1 = (int)h->n[1], n2 = (int)h->n[2], n0 = (int)h->n[0];
    const int R = h->R;
    const size_t lds_cap = 64 * 1024;
    TileChoice t{};
    t.ppt = ppt;
    int ty = h->ty_req > 0 ? h->ty_req : 0;
    if (ty == 0) {
        // largest divisor-friendly row count that fits the LDS and 1024
        // threads; prefer divisors of n1 to avoid a ragged last tile
        int best = 1;
        for (int c = 1; c <= n1; c++) {
            size_t lds = 2 * (size_t)(c + 2 * R) * n2 * vsize;
            int thr = (c * n2 + ppt - 1) / ppt;
            if (lds > lds_cap || thr > kCfdMaxThreads) break;
            if (n1 % c == 0 || c == n1) best = c;
        }
        ty = best;
        // ragged tiles are fine if no divisor >= 8 exists
        if (ty < 8) {
            for (int c = n1; c >= 1; c--) {
                size_t lds = 2 * (size_t)(c + 2 * R) * n2 * vsize;
                int thr = (c * n2 + ppt - 1) / ppt;
                if (lds <= lds_cap && thr <= kCfdMaxThreads) { ty = c; break; }
            }
        }
    }
    ty = std::min(ty, n1);
    t.ty = ty;
    t.lx = h->lx_req > 0 ? std::min<int>(h->lx_req, n0) : n0;
    int thr = (ty * n2 + ppt - 1) / ppt;
    t.threads = ((thr + 63) / 64) * 64;
    t.lds = 2 * (size_t)(ty + 2 * R) * n2 * vsize;
    return t;
}

template <typename V, typename S, int R, int PPT>
int launch_cfd_r(const gh_stencil* h, const void* in, void* out,
                 int64_t nbands, const double* phase_cd, const void* vt,
                 const void* src, double beta, double alpha, double gamma,
                 const double* eps, hipStream_t st) {
    TileChoice t = choose_tile(h, sizeof(V), PPT);
    GH_REQUIRE(t.threads <= kCfdMaxThreads && t.lds <= 64 * 1024, GH_ERR_UNSUPPORTED,
               "stencil: grid rows of %lld points do not fit one block",
               (long long)h->n[2]);
    CfdGeom g;
    g.n0 = (int)h->n[0]; g.n1 = (int)h->n[1]; g.n2 = (int)h->n[2];
    g.pbc0 = h->pbc[0]; g.pbc1 = h->pbc[1]; g.pbc2 = h->pbc[2];
    g.ty = t.ty; g.lx = t.lx;
    g.nyt = (g.n1 + t.ty - 1) / t.ty;
    g.nxc = (g.n0 + t.lx - 1) / t.lx;
    CfdCoefs<S> cf;
    cf.c0 = (S)(alpha * h->c0 + gamma);
    cf.eps = eps;
    for (int c = 0; c < 3; c++)
        for (int j = 0; j < kMaxR; j++) {
            cf.cm[c][j] = (S)(alpha * h->cm[c][j]);
            cf.cp[c][j] = (S)(alpha * h->cp[c][j]);
        }
    PhaseArgs<S> ph = phase_args<S>(phase_cd);
    int64_t nblocks = nbands * (int64_t)g.nyt * g.nxc;
    GH_REQUIRE(nblocks < (1ll << 31), GH_ERR_SHAPE, "stencil: too many blocks");
    dim3 grid((unsigned)nblocks), block(t.threads);
    const S* vtp = (const S*)vt;
    const V* srcp = (const V*)src;
    const bool gs1 = 2 * R * g.n2 <= t.threads;
#define GH_LAUNCH_CFD(GS, HV, HS)                                              \
    hipLaunchKernelGGL((stencil_cfd_kernel<V, S, R, PPT, GS, HV, HS>), grid,   \
                       block, t.lds, st, (const V*)in, (V*)out, vtp, srcp,     \
                       (S)beta, cf, g, ph)
    if (gs1) {
        if (vt && src) GH_LAUNCH_CFD(1, true, true);
        else if (vt) GH_LAUNCH_CFD(1, true, false);
        else if (src) GH_LAUNCH_CFD(1, false, true);
        else GH_LAUNCH_CFD(1, false, false);
    } else {
        if (vt && src) GH_LAUNCH_CFD(4, true, true);
        else if (vt) GH_LAUNCH_CFD(4, true, false);
        else if (src) GH_LAUNCH_CFD(4, false, true);
        else GH_LAUNCH_CFD(4, false, false);
    }
#undef GH_LAUNCH_CFD
    GH_CHECK_LAUNCH();
    return GH_OK;
}

// LDS-ring launch: returns GH_ERR_UNSUPPORTED when the shape does not fit
// (rows not 16-byte multiples, or no tile fits the LDS), so the caller can
// fall back to the register kernel.
constexpr size_t kLdsBytes = 160 * 1024;

// Compute units of the current device (queried once per device).
int device_cu_count(int* out) {
    static int cached[64] = {0};
    int dev = 0;
    GH_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) dev = 0;
    if (!cached[dev])
        GH_CHECK_HIP(hipDeviceGetAttribute(&cached[dev], hipDeviceAttributeMultiprocessorCount,
                                           dev));
    *out = cached[dev];
    return GH_OK;
}

// Raise a kernel's dynamic LDS limit to the whole 160 KB once per process
// (not per launch).
std::mutex g_attr_mutex;
std::vector<const void*> g_attr_done;

int ensure_max_lds(const void* fn, int bytes) {
    std::lock_guard<std::mutex> lock(g_attr_mutex);
    if (std::find(g_attr_done.begin(), g_attr_done.end(), fn) != g_attr_done.end())
        return GH_OK;
    GH_CHECK_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
    g_attr_done.push_back(fn);
    return GH_OK;
}

template <typename V, typename S, int R, int PPT>
int launch_ring_r(const gh_stencil* h, const void* in, void* out,
                  int64_t nbands, const double* phase_cd, const void* vt,
                  const void* src, double beta, double alpha, double gamma,
                  const double* eps, hipStream_t st) {
    const int n0 = (int)h->n[0], n1 = (int)h->n[1], n2 = (int)h->n[2];
    const int vs = (int)sizeof(V);
    if ((n2 * vs) % kDmaBytes != 0) return GH_ERR_UNSUPPORTED;
    // the plane DMA and the paired loads / stores move 16-byte vectors: a
    // pointer off that alignment (a view at an odd element offset) takes the
    // register kernel instead
    auto aligned16 = [](const void* p) { return ((uintptr_t)p & 15u) == 0; };
    if (!aligned16(in) || !aligned16(out) || (vt && !aligned16(vt)) ||
        (src && !aligned16(src)))
        return GH_ERR_UNSUPPORTED;
    auto lds_of = [&](int c, int d) {
        return (size_t)(R + d + 1) * (c + 2 * R) * n2 * vs;
    };
    // Largest tile (rows) that fits for each prefetch depth.  Explicit
    // rows/depth (autotuner) override the automatic choice.
    auto max_rows = [&](int d) {
        int best = 0;
        for (int c = 1; c <= n1; c++) {
            int thr = (c * n2 + PPT - 1) / PPT;
            int ncomp = (thr + 63) / 64;
            if (lds_of(c, d) > kLdsBytes - 1024 || (ncomp + d) * 64 > kRingMaxThreads) break;
            best = c;
        }
        return best;
    };
    auto divisor_rows = [&](int c) {   // prefer a divisor of n1 >= c / 2
        for (int t = c; t >= std::max(1, c / 2); t--)
            if (n1 % t == 0) return t;
        return c;
    };
    int ty = 0, D = 0;
    if (h->depth_req >= 2) {
        D = h->depth_req;
        int mr = max_rows(D);
        ty = h->ty_req > 0 ? std::min(h->ty_req, mr) : divisor_rows(mr);
    } else {
        // depth 2 with the largest tile: measured faster than depth 3 with
        // fewer rows (256 x 112^3 FP64: 16 rows x 2 planes 1.84 ms,
        // 14 rows x 3 planes 1.96 ms)
        D = 2;
        ty = h->ty_req > 0 ? std::min(h->ty_req, max_rows(D)) : divisor_rows(max_rows(D));
    }
    if (ty <= 0) return GH_ERR_UNSUPPORTED;
    const size_t lds = lds_of(ty, D);
    RingLayout lay;
    lay.psi_slot = (ty + 2 * R) * n2;
    RingGeom rg;
    CfdGeom& g = rg.g;
    g.n0 = n0; g.n1 = n1; g.n2 = n2;
    g.pbc0 = h->pbc[0]; g.pbc1 = h->pbc[1]; g.pbc2 = h->pbc[2];
    g.ty = ty;
    g.nyt = (n1 + ty - 1) / ty;
    if (h->lx_req > 0) {
        g.lx = std::min(h->lx_req, n0);
    } else {
        // one block per CU: when the bands x y-tiles do not cover the CUs,
        // split the x-march into chunks (each chunk re-reads 2R planes),
        // keeping chunks at least 4R planes long
        int ncu = 0;
        int rc = device_cu_count(&ncu);
        if (rc != GH_OK) return rc;
        const int64_t base = nbands * (int64_t)g.nyt;
        int nxc = 1;
        while (base * nxc < ncu && (n0 + nxc) / (nxc + 1) >= 4 * R) nxc++;
        g.lx = (n0 + nxc - 1) / nxc;
    }
    g.nxc = (n0 + g.lx - 1) / g.lx;
    rg.nload = D;
    rg.ncomp = ((ty * n2 + PPT - 1) / PPT + 63) / 64;
    rg.nbands = (int)nbands;
    const int tiles = g.nyt * g.nxc;
    const int64_t nlogical = nbands * (int64_t)tiles;
    GH_REQUIRE(nlogical < (1ll << 31), GH_ERR_SHAPE, "stencil: too many blocks");
    rg.nlogical = (int)nlogical;
    rg.rounds = 1;
    int64_t nblocks = nlogical;
    int per_xcd = 32;                     // blocks running together per XCD
    if (h->persist) {
        // one resident block per CU walks its XCD's tiles band group by band
        // group, so all blocks of an XCD stay on the same planes
        int64_t n = h->persist;
        if (n < kNumXCD) {
            int ncu = 0;
            int rc = device_cu_count(&ncu);
            if (rc != GH_OK) return rc;
            const int per_cu = std::max<int>(1, (int)(kLdsBytes / std::max<size_t>(lds, 1)));
            n = (int64_t)ncu * per_cu;
        }
        n -= n % kNumXCD;
        if (n >= kNumXCD && n < nlogical) {
            nblocks = n;
            rg.rounds = (int)((nlogical + n - 1) / n);
            per_xcd = (int)(n / kNumXCD);
        }
    }
    {
        int bgr = h->bgroup_req > 0 ? h->bgroup_req : std::max(1, per_xcd / tiles);
        rg.bgroup = std::min<int>(bgr, (int)nbands);
        const char* nt = getenv("GPAW_AMD_NTSTORE");
        rg.nt_store = nt ? atoi(nt) : 1;
    }
    CfdCoefs<S> cf;
    cf.c0 = (S)(alpha * h->c0 + gamma);
    cf.eps = eps;
    for (int c = 0; c < 3; c++)
        for (int j = 0; j < kMaxR; j++) {
            cf.cm[c][j] = (S)(alpha * h->cm[c][j]);
            cf.cp[c][j] = (S)(alpha * h->cp[c][j]);
        }
    PhaseArgs<S> ph = phase_args<S>(phase_cd);
    rg.cplx_fast = 0;
    if (is_complex<V>::value) {
        bool triv = true;
        for (int c = 1; c < 3; c++)
            for (int d = 0; d < 2; d++)
                triv = triv && ph.re[c][d] == (S)1 && ph.im[c][d] == (S)0;
        rg.cplx_fast = triv ? 1 : 0;
    }
    dim3 grid((unsigned)nblocks), block((rg.ncomp + rg.nload) * 64);
    const S* vtp = (const S*)vt;
    const V* srcp = (const V*)src;
    // symmetric coefficients: the paired real path keeps half of them live
    bool sym = true;
    for (int c = 0; c < 3; c++)
        for (int j = 0; j < R; j++) sym = sym && h->cm[c][j] == h->cp[c][j];
    constexpr bool kPaired = PPT == 2 || PPT == 4;   // the vector paths take SYM
#define GH_LAUNCH_RING_S(HV, HS, SY)                                           \
    do {                                                                       \
        auto kfn = stencil_ring_kernel<V, S, R, PPT, HV, HS, SY>;              \
        int arc = ensure_max_lds((const void*)kfn, (int)kLdsBytes);            \
        if (arc != GH_OK) return arc;                                          \
        hipLaunchKernelGGL(kfn, grid, block, lds, st, (const V*)in, (V*)out,   \
                           vtp, srcp, (S)beta, cf, rg, ph, lay);               \
    } while (0)
#define GH_LAUNCH_RING(HV, HS)                                                 \
    do {                                                                       \
        if constexpr (kPaired) {                                               \
            if (sym) GH_LAUNCH_RING_S(HV, HS, true);                           \
            else GH_LAUNCH_RING_S(HV, HS, false);                              \
        } else {                                                               \
            GH_LAUNCH_RING_S(HV, HS, false);                                   \
        }                                                                      \
    } while (0)
    if (vt && src) GH_LAUNCH_RING(true, true);
    else if (vt) GH_LAUNCH_RING(true, false);
    else if (src) GH_LAUNCH_RING(false, true);
    else GH_LAUNCH_RING(false, false);
#undef GH_LAUNCH_RING
#undef GH_LAUNCH_RING_S
    GH_CHECK_LAUNCH();
    return GH_OK;
}

// points per compute thread of the ring kernel
template <typename V> struct RingPPT { static constexpr int value = 2; };
template <> struct RingPPT<cplx<double>> { static constexpr int value = 1; };
template <> struct RingPPT<float> { static constexpr int value = 4; };
template <> struct RingPPT<cplx<float>> { static constexpr int value = 2; };

template <typename V, typename S, int PPT>
int launch_cfd(const gh_stencil* h, const void* in, void* out, int64_t nb,
               const double* ph, const void* vt, const void* src, double beta,
               double al, double ga, const double* eps, hipStream_t st) {
    if (!h->reg_only) {
        constexpr int RP = RingPPT<V>::value;
        int rc = GH_ERR_UNSUPPORTED;
        // complex data takes the two-point path only when the y and z Bloch
        // phases are 1; otherwise the one-point path with phases
        bool two = RP > 1;
        if (is_complex<V>::value && ph)
            for (int c = 1; c < 3; c++)
                for (int d = 0; d < 2; d++)
                    two = two && ph[(c * 2 + d) * 2] == 1.0 && ph[(c * 2 + d) * 2 + 1] == 0.0;
        if (two) {
            switch (h->R) {
                case 1: rc = launch_ring_r<V, S, 1, RP>(h, in, out, nb, ph, vt, src, beta, al, ga, eps, st); break;
                case 2: rc = launch_ring_r<V, S, 2, RP>(h, in, out, nb, ph, vt, src, beta, al, ga, eps, st); break;
                case 3: rc = launch_ring_r<V, S, 3, RP>(h, in, out, nb, ph, vt, src, beta, al, ga, eps, st); break;
            }
        } else {
            switch (h->R) {
                case 1: rc = launch_ring_r<V, S, 1, 1>(h, in, out, nb, ph, vt, src, beta, al, ga, eps, st); break;
                case 2: rc = launch_ring_r<V, S, 2, 1>(h, in, out, nb, ph, vt, src, beta, al, ga, eps, st); break;
                case 3: rc = launch_ring_r<V, S, 3, 1>(h, in, out, nb, ph, vt, src, beta, al, ga, eps, st); break;
            }
        }
        if (rc != GH_ERR_UNSUPPORTED) return rc;
    }
    switch (h->R) {
        case 1: return launch_cfd_r<V, S, 1, PPT>(h, in, out, nb, ph, vt, src, beta, al, ga, eps, st);
        case 2: return launch_cfd_r<V, S, 2, PPT>(h, in, out, nb, ph, vt, src, beta, al, ga, eps, st);
        case 3: return launch_cfd_r<V, S, 3, PPT>(h, in, out, nb, ph, vt, src, beta, al, ga, eps, st);
        case 4: return launch_cfd_r<V, S, 4, PPT>(h, in, out, nb, ph, vt, src, beta, al, ga, eps, st);
        case 5: return launch_cfd_r<V, S, 5, PPT>(h, in, out, nb, ph, vt, src, beta, al, ga, eps, st);
        case 6: return launch_cfd_r<V, S, 6, PPT>(h, in, out, nb, ph, vt, src, beta, al, ga, eps, st);
    }
    set_error("stencil: radius %d unsupported", h->R);
    return GH_ERR_UNSUPPORTED;
}

template <typename V, typename S>
int launch_generic(const gh_stencil* h, const void* in, void* out,
                   int64_t nbands, const double* phase_cd, const void* vt,
                   const void* src, double beta, double alpha, double gamma,
                   hipStream_t st) {
    GenericArgs<S> a;
    a.np = (int)h->coefs.size();
    a.n0 = (int)h->n[0]; a.n1 = (int)h->n[1]; a.n2 = (int)h->n[2];
    a.pbc0 = h->pbc[0]; a.pbc1 = h->pbc[1]; a.pbc2 = h->pbc[2];
    for (int p = 0; p < a.np; p++) {
        a.coef[p] = (S)(alpha * h->coefs[p]);
        for (int c = 0; c < 3; c++) a.off[p][c] = (signed char)h->offs[p * 3 + c];
    }
    PhaseArgs<S> ph = phase_args<S>(phase_cd);
    int64_t total = nbands * h->n[0] * h->n[1] * h->n[2];
    int64_t nblk = std::min<int64_t>((total + 255) / 256, 65536);
    const S* vtp = (const S*)vt;
    const V* srcp = (const V*)src;
    if (vt && src)
        hipLaunchKernelGGL((stencil_generic_kernel<V, S, true, true>), dim3(nblk), dim3(256), 0, st,
                           (const V*)in, (V*)out, vtp, srcp, (S)beta, (S)gamma, a, ph, total);
    else if (vt)
        hipLaunchKernelGGL((stencil_generic_kernel<V, S, true, false>), dim3(nblk), dim3(256), 0, st,
                           (const V*)in, (V*)out, vtp, srcp, (S)beta, (S)gamma, a, ph, total);
    else if (src)
        hipLaunchKernelGGL((stencil_generic_kernel<V, S, false, true>), dim3(nblk), dim3(256), 0, st,
                           (const V*)in, (V*)out, vtp, srcp, (S)beta, (S)gamma, a, ph, total);
    else
        hipLaunchKernelGGL((stencil_generic_kernel<V, S, false, false>), dim3(nblk), dim3(256), 0, st,
                           (const V*)in, (V*)out, vtp, srcp, (S)beta, (S)gamma, a, ph, total);
    GH_CHECK_LAUNCH();
    return GH_OK;
}


// ------------------------------------------------------------ batches --
// Launch constants of a batch of cells for the persistent ring launch.
struct BatchPlan {
    int dtype = 0;
    int ncells = 0;
    int R = 0;
    bool has_v = false;
    bool sym = true;
    int order[kMaxBatchCells];      // table slot k holds cell order[k]
    int threads = 0;
    size_t lds = 0;
    int nblocks = 0;
    int rounds = 0;
    int ntotal = 0;
    void* table = nullptr;          // device RingCell<S>[ncells]
    int* counters = nullptr;        // device: next tile, finished blocks
    int rows[kMaxBatchCells];
    int depth[kMaxBatchCells];
};


template <typename V, typename S, int R>
int plan_batch(BatchPlan& bp, const gh_stencil* const* ops, const int64_t* nbands,
               const int32_t* rows, const int32_t* depths) {
    constexpr int PPT = RingPPT<V>::value;
    const int vs = (int)sizeof(V);
    std::vector<RingCell<S>> cells(bp.ncells);
    std::vector<double> cost(bp.ncells);
    std::vector<size_t> ldsv(bp.ncells);
    int maxthr = 0;
    size_t maxlds = 0;
    for (int c = 0; c < bp.ncells; c++) {
        const gh_stencil* h = ops[c];
        const int n0 = (int)h->n[0], n1 = (int)h->n[1], n2 = (int)h->n[2];
        GH_REQUIRE((n2 * vs) % kDmaBytes == 0 && n2 % 2 == 0, GH_ERR_UNSUPPORTED,
                   "stencil_batch: cell %d rows of %d elements are not 16-byte multiples", c,
                   n2);
        auto lds_of = [&](int ty, int d) {
            return (size_t)(R + d + 1) * (ty + 2 * R) * n2 * vs;
        };
        const int D = depths && depths[c] >= 2 ? depths[c] : 2;
        int best = 0;
        for (int ty = 1; ty <= n1; ty++) {
            const int ncomp = ((ty * n2 + PPT - 1) / PPT + 63) / 64;
            if (lds_of(ty, D) > kLdsBytes - 1024 || (ncomp + D) * 64 > kRingMaxThreads) break;
            best = ty;
        }
        GH_REQUIRE(best > 0, GH_ERR_UNSUPPORTED, "stencil_batch: cell %d has no tile", c);
        int ty = rows && rows[c] > 0 ? std::min<int>(rows[c], best) : best;
        RingCell<S>& rc = cells[c];
        CfdGeom& g = rc.rg.g;
        g.n0 = n0; g.n1 = n1; g.n2 = n2;
        g.pbc0 = h->pbc[0]; g.pbc1 = h->pbc[1]; g.pbc2 = h->pbc[2];
        g.ty = ty;
        g.nyt = (n1 + ty - 1) / ty;
        g.lx = n0;
        g.nxc = 1;
        rc.rg.nload = D;
        rc.rg.ncomp = ((ty * n2 + PPT - 1) / PPT + 63) / 64;
        rc.rg.nbands = (int)nbands[c];
        rc.rg.nlogical = (int)(nbands[c] * g.nyt);
        rc.rg.rounds = 1;
        rc.rg.nt_store = 1;
        rc.rg.cplx_fast = 0;
        rc.cf.c0 = (S)h->c0;
        rc.cf.eps = nullptr;
        for (int a = 0; a < 3; a++)
            for (int j = 0; j < kMaxR; j++) {
                rc.cf.cm[a][j] = (S)h->cm[a][j];
                rc.cf.cp[a][j] = (S)h->cp[a][j];
            }
        rc.ph = phase_args<S>(nullptr);
        for (int a = 0; a < 3; a++)
            for (int j = 0; j < R; j++) bp.sym = bp.sym && h->cm[a][j] == h->cp[a][j];
        ldsv[c] = lds_of(ty, D);
        maxlds = std::max(maxlds, ldsv[c]);
        maxthr = std::max(maxthr, (rc.rg.ncomp + D) * 64);
        cost[c] = (double)n0 * ty * n2;
        bp.rows[c] = ty;
        bp.depth[c] = D;
    }
    // largest tiles first: the small ones fill the last rounds
    for (int k = 0; k < bp.ncells; k++) bp.order[k] = k;
    std::stable_sort(bp.order, bp.order + bp.ncells,
                     [&](int a, int b) { return cost[a] > cost[b]; });
    int ncu = 0;
    int rc = device_cu_count(&ncu);
    if (rc != GH_OK) return rc;
    const int per_cu = std::max(1, std::min((int)(kLdsBytes / maxlds), 2048 / maxthr));
    int nblocks = ncu * per_cu;
    nblocks -= nblocks % kNumXCD;
    int64_t ntotal = 0;
    std::vector<RingCell<S>> sorted(bp.ncells);
    for (int k = 0; k < bp.ncells; k++) {
        RingCell<S> cl = cells[bp.order[k]];
        cl.lbeg = (int)ntotal;
        const int tiles = cl.rg.g.nyt * cl.rg.g.nxc;
        cl.rg.bgroup = std::min(cl.rg.nbands, std::max(1, (nblocks / kNumXCD) / tiles));
        ntotal += cl.rg.nlogical;
        sorted[k] = cl;
    }
    GH_REQUIRE(ntotal < (1ll << 31), GH_ERR_SHAPE, "stencil_batch: too many tiles");
    bp.ntotal = (int)ntotal;
    bp.nblocks = (int)std::max<int64_t>(1, std::min<int64_t>(nblocks, ntotal));
    bp.rounds = (int)((ntotal + bp.nblocks - 1) / bp.nblocks);   // on average
    bp.threads = maxthr;
    bp.lds = maxlds;
    GH_CHECK_HIP(hipMalloc((void**)&bp.counters, 2 * sizeof(int)));
    GH_CHECK_HIP(hipMemset(bp.counters, 0, 2 * sizeof(int)));
    GH_CHECK_HIP(hipMalloc(&bp.table, sizeof(RingCell<S>) * bp.ncells));
    GH_CHECK_HIP(hipMemcpy(bp.table, sorted.data(), sizeof(RingCell<S>) * bp.ncells,
                           hipMemcpyHostToDevice));
    // the dynamic LDS limit is a property of the kernel: set once here
#define GH_BATCH_ATTR(HV, SY)                                                   \
    GH_CHECK_HIP(hipFuncSetAttribute(                                           \
        (const void*)stencil_ring_batched_kernel<V, S, R, PPT, HV, SY>,         \
        hipFuncAttributeMaxDynamicSharedMemorySize, (int)(kLdsBytes - 1024)))
    GH_BATCH_ATTR(true, true);
    GH_BATCH_ATTR(true, false);
    GH_BATCH_ATTR(false, true);
    GH_BATCH_ATTR(false, false);
#undef GH_BATCH_ATTR
    return GH_OK;
}

template <typename V, typename S, int R>
int launch_batch(const BatchPlan& bp, const void* const* in, void* const* out,
                 const void* const* vt, hipStream_t st) {
    constexpr int PPT = RingPPT<V>::value;
    BatchPtrs<V, S> p;
    for (int k = 0; k < kMaxBatchCells; k++) {
        p.in[k] = nullptr;
        p.out[k] = nullptr;
        p.vt[k] = nullptr;
    }
    for (int k = 0; k < bp.ncells; k++) {
        const int c = bp.order[k];
        p.in[k] = (const V*)in[c];
        p.out[k] = (V*)out[c];
        p.vt[k] = vt ? (const S*)vt[c] : nullptr;
    }
    const RingCell<S>* tab = (const RingCell<S>*)bp.table;
    dim3 grid((unsigned)bp.nblocks), block((unsigned)bp.threads);
#define GH_BATCH_LAUNCH(HV, SY)                                                 \
    hipLaunchKernelGGL((stencil_ring_batched_kernel<V, S, R, PPT, HV, SY>), grid, block, \
                       bp.lds, st, tab, p, bp.ncells, bp.ntotal, bp.counters)
    if (bp.has_v) {
        if (bp.sym) GH_BATCH_LAUNCH(true, true);
        else GH_BATCH_LAUNCH(true, false);
    } else {
        if (bp.sym) GH_BATCH_LAUNCH(false, true);
        else GH_BATCH_LAUNCH(false, false);
    }
#undef GH_BATCH_LAUNCH
    GH_CHECK_LAUNCH();
    return GH_OK;
}

}  // namespace

extern "C" {

const char* gh_last_error(void) { return gh::g_err; }

const char* gh_version(void) { return "gpaw_amd 0.1 gfx950"; }

int gh_stencil_create(gh_stencil** out, int ncoefs, const double* coefs,
                      const int64_t* offsets_pc, const int64_t* n_c,
                      const int32_t* pbc_c, int dtype) {
    GH_REQUIRE(out && coefs && offsets_pc && n_c && pbc_c, GH_ERR_ARG,
               "stencil_create: null argument");
    GH_REQUIRE(ncoefs > 0 && ncoefs <= kMaxGeneric, GH_ERR_ARG,
               "stencil_create: ncoefs=%d out of range", ncoefs);
    GH_REQUIRE(dtype >= GH_F64 && dtype <= GH_C64, GH_ERR_ARG,
               "stencil_create: bad dtype %d", dtype);
    for (int c = 0; c < 3; c++)
        GH_REQUIRE(n_c[c] > 0 && n_c[c] < (1 << 20), GH_ERR_SHAPE,
                   "stencil_create: bad grid size");
    auto* h = new gh_stencil();
    h->dtype = dtype;
    for (int c = 0; c < 3; c++) {
        h->n[c] = n_c[c];
        h->pbc[c] = pbc_c[c] ? 1 : 0;
    }
    h->coefs.assign(coefs, coefs + ncoefs);
    {
        const char* e = getenv("GPAW_AMD_STENCIL");
        h->reg_only = e && std::string(e) == "reg";
        const char* bg = getenv("GPAW_AMD_BGROUP");
        h->bgroup_req = bg ? atoi(bg) : 0;
        const char* pe = getenv("GPAW_AMD_PERSIST");
        h->persist = pe ? std::max(0, atoi(pe)) : 0;
    }
    h->offs.assign(offsets_pc, offsets_pc + 3 * ncoefs);
    h->R = 0;
    h->cfd = true;
    h->diag = 0.0;
    h->c0 = 0.0;
    std::memset(h->cm, 0, sizeof(h->cm));
    std::memset(h->cp, 0, sizeof(h->cp));
    for (int p = 0; p < ncoefs; p++) {
        const int64_t* o = offsets_pc + 3 * p;
        int nz = (o[0] != 0) + (o[1] != 0) + (o[2] != 0);
        for (int c = 0; c < 3; c++) {
            int64_t a = o[c] < 0 ? -o[c] : o[c];
            if (a > h->R) h->R = (int)a;
            if (a > n_c[c]) {
                delete h;
                set_error("stencil_create: stencil longer than domain");
                return GH_ERR_SHAPE;
            }
        }
        if (nz == 0) { h->c0 += coefs[p]; h->diag = coefs[p]; }
        if (nz >= 2) h->cfd = false;
    }
    if (h->R > 127) { delete h; set_error("stencil_create: offset too large"); return GH_ERR_UNSUPPORTED; }
    if (h->cfd && h->R >= 1 && h->R <= kMaxR) {
        for (int p = 0; p < ncoefs; p++) {
            const int64_t* o = offsets_pc + 3 * p;
            for (int c = 0; c < 3; c++) {
                if (o[c] > 0) h->cp[c][o[c] - 1] += coefs[p];
                if (o[c] < 0) h->cm[c][-o[c] - 1] += coefs[p];
            }
        }
    } else {
        h->cfd = false;
    }
    *out = h;
    return GH_OK;
}

int gh_stencil_destroy(gh_stencil* h) {
    delete h;
    return GH_OK;
}

double gh_stencil_diagonal(const gh_stencil* h) { return h ? h->diag : 0.0; }

int gh_stencil_set_tile(gh_stencil* h, int rows, int planes, int depth) {
    GH_REQUIRE(h && rows >= 0 && planes >= 0 && depth >= 0 && depth <= 3 &&
                   depth != 1,
               GH_ERR_ARG, "set_tile: bad args");
    h->ty_req = rows;
    h->lx_req = planes;
    h->depth_req = depth;
    return GH_OK;
}

int gh_stencil_apply(const gh_stencil* h, const void* in_dev, void* out_dev,
                     int64_t nbands, const double* phase_cd,
                     const void* vt_dev, const void* src_dev, double beta,
                     void* stream) {
    return gh::stencil_apply_scaled(h, in_dev, out_dev, nbands, phase_cd,
                                    vt_dev, src_dev, beta, 1.0, 0.0,
                                    (hipStream_t)stream);
}

int gh_stencil_apply_shifted(const gh_stencil* h, const void* in_dev,
                             void* out_dev, int64_t nbands,
                             const double* phase_cd, const void* vt_dev,
                             const double* eps_dev, void* stream) {
    GH_REQUIRE(eps_dev || nbands == 0, GH_ERR_ARG, "stencil_apply_shifted: null eps");
    return gh::stencil_apply_full(h, in_dev, out_dev, nbands, phase_cd, vt_dev,
                                  nullptr, 0.0, 1.0, 0.0, eps_dev,
                                  (hipStream_t)stream);
}

int gh_stencil_relax(const gh_stencil* h, int relax_method, void* f_dev,
                     const void* s_dev, void* work_dev, int nrelax, double w,
                     void* stream) {
    GH_REQUIRE(h && f_dev && s_dev && work_dev, GH_ERR_ARG, "relax: null argument");
    GH_REQUIRE(nrelax >= 0, GH_ERR_ARG, "relax: nrelax < 0");
    GH_REQUIRE(h->dtype == GH_F64 || h->dtype == GH_F32, GH_ERR_ARG,
               "relax: real grids only");
    GH_REQUIRE(relax_method == 1 || relax_method == 2, GH_ERR_ARG,
               "relax: method %d (1 Gauss-Seidel, 2 Jacobi)", relax_method);
    GH_REQUIRE(h->diag != 0.0, GH_ERR_ARG, "relax: zero diagonal");
    GH_REQUIRE(f_dev != work_dev && s_dev != work_dev && f_dev != s_dev, GH_ERR_ARG,
               "relax: f, s and work must be distinct");
    hipStream_t st = (hipStream_t)stream;
    if (relax_method == 1) {
        GH_REQUIRE(h->cfd, GH_ERR_UNSUPPORTED,
                   "relax: Gauss-Seidel needs a stencil along the grid axes");
        GH_REQUIRE((int)h->coefs.size() <= kMaxGeneric, GH_ERR_UNSUPPORTED,
                   "relax: too many stencil points");
        const int n0 = (int)h->n[0], n1 = (int)h->n[1], n2 = (int)h->n[2];
        int icentre = -1;
        for (size_t p = 0; p < h->coefs.size(); p++)
            if (h->offs[3 * p] == 0 && h->offs[3 * p + 1] == 0 && h->offs[3 * p + 2] == 0)
                icentre = (int)p;
        GH_REQUIRE(icentre >= 0, GH_ERR_ARG, "relax: stencil without a centre point");
        const size_t bytes = (size_t)n0 * n1 * n2 * dtype_size(h->dtype);
        const unsigned nblk = (unsigned)(((int64_t)n0 * n1 + 255) / 256);
        const int nplanes = n0 + n1 + n2 - 2;
#define GS_SWEEPS(S)                                                                   \
        do {                                                                           \
            GenericArgs<S> ga;                                                         \
            ga.np = (int)h->coefs.size();                                              \
            ga.n0 = n0; ga.n1 = n1; ga.n2 = n2;                                        \
            ga.pbc0 = h->pbc[0]; ga.pbc1 = h->pbc[1]; ga.pbc2 = h->pbc[2];             \
            for (int p = 0; p < ga.np; p++) {                                          \
                ga.coef[p] = (S)h->coefs[p];                                           \
                for (int c = 0; c < 3; c++) ga.off[p][c] = (signed char)h->offs[3 * p + c]; \
            }                                                                          \
            const S inv = (S)(1.0 / h->coefs[icentre]);                                \
            for (int n = 0; n < nrelax; n++) {                                         \
                GH_CHECK_HIP(hipMemcpyAsync(work_dev, f_dev, bytes, hipMemcpyDeviceToDevice, st)); \
                for (int pl = 0; pl < nplanes; pl++)                                   \
                    hipLaunchKernelGGL((gs_plane_kernel<S>), dim3(nblk), dim3(256), 0, st, \
                                       (S*)f_dev, (const S*)work_dev, (const S*)s_dev, ga, \
                                       icentre, pl, inv);                              \
                GH_CHECK_LAUNCH();                                                     \
            }                                                                          \
        } while (0)
        if (h->dtype == GH_F64) GS_SWEEPS(double);
        else GS_SWEEPS(float);
#undef GS_SWEEPS
        return GH_OK;
    }
    // f' = (1 - w) f + (w / c0) (s - (L f - c0 f)) = -(w / c0) L f + f + (w / c0) s
    const double a = -w / h->diag;
    const void* cur = f_dev;
    void* nxt = work_dev;
    for (int n = 0; n < nrelax; n++) {
        int rc = gh::stencil_apply_scaled(h, cur, nxt, 1, nullptr, nullptr, s_dev,
                                          -a, a, 1.0, st);
        if (rc != GH_OK) return rc;
        void* done = nxt;
        nxt = (void*)cur;
        cur = done;
    }
    if (cur != f_dev) {
        const size_t bytes = (size_t)(h->n[0] * h->n[1] * h->n[2]) * dtype_size(h->dtype);
        GH_CHECK_HIP(hipMemcpyAsync(f_dev, cur, bytes, hipMemcpyDeviceToDevice, st));
    }
    return GH_OK;
}


struct gh_stencil_batch {
    BatchPlan plan;
};

int gh_stencil_batch_create(gh_stencil_batch** out, int ncells, const gh_stencil* const* ops,
                            const int64_t* nbands, const int32_t* rows,
                            const int32_t* depths, int with_potential) {
    GH_REQUIRE(out && ops && nbands, GH_ERR_ARG, "stencil_batch_create: null argument");
    GH_REQUIRE(ncells >= 1 && ncells <= kMaxBatchCells, GH_ERR_ARG,
               "stencil_batch_create: %d cells (1..%d)", ncells, kMaxBatchCells);
    const int dt = ops[0] ? ops[0]->dtype : -1;
    GH_REQUIRE(dt == GH_F32 || dt == GH_F64, GH_ERR_UNSUPPORTED,
               "stencil_batch_create: real float32 / float64 cells only");
    const int R = ops[0]->R;
    for (int c = 0; c < ncells; c++) {
        GH_REQUIRE(ops[c] && ops[c]->dtype == dt && ops[c]->cfd && ops[c]->R == R,
                   GH_ERR_ARG,
                   "stencil_batch_create: cell %d: every cell needs an axis-aligned stencil "
                   "of the same dtype and radius", c);
        GH_REQUIRE(nbands[c] >= 1 && nbands[c] < (1 << 24), GH_ERR_ARG,
                   "stencil_batch_create: cell %d: bad band count", c);
    }
    GH_REQUIRE(R >= 1 && R <= 3, GH_ERR_UNSUPPORTED, "stencil_batch_create: radius %d", R);
    auto* b = new gh_stencil_batch();
    BatchPlan& bp = b->plan;
    bp.dtype = dt;
    bp.ncells = ncells;
    bp.R = R;
    bp.has_v = with_potential != 0;
    int rc = GH_ERR_UNSUPPORTED;
#define GH_PLAN(V, S)                                                           \
    switch (R) {                                                               \
        case 1: rc = plan_batch<V, S, 1>(bp, ops, nbands, rows, depths); break; \
        case 2: rc = plan_batch<V, S, 2>(bp, ops, nbands, rows, depths); break; \
        case 3: rc = plan_batch<V, S, 3>(bp, ops, nbands, rows, depths); break; \
    }
    if (dt == GH_F32) { GH_PLAN(float, float) } else { GH_PLAN(double, double) }
#undef GH_PLAN
    if (rc != GH_OK) {
        if (bp.table) (void)hipFree(bp.table);
        if (bp.counters) (void)hipFree(bp.counters);
        delete b;
        return rc;
    }
    *out = b;
    return GH_OK;
}

int gh_stencil_batch_destroy(gh_stencil_batch* b) {
    if (!b) return GH_OK;
    if (b->plan.table) (void)hipFree(b->plan.table);
    if (b->plan.counters) (void)hipFree(b->plan.counters);
    delete b;
    return GH_OK;
}

int gh_stencil_batch_info(const gh_stencil_batch* b, int64_t* info, int32_t* rows,
                          int32_t* depths) {
    GH_REQUIRE(b && info, GH_ERR_ARG, "stencil_batch_info: null argument");
    const BatchPlan& bp = b->plan;
    info[0] = bp.nblocks;
    info[1] = bp.threads;
    info[2] = (int64_t)bp.lds;
    info[3] = bp.ntotal;
    info[4] = bp.rounds;
    for (int c = 0; c < bp.ncells; c++) {
        if (rows) rows[c] = bp.rows[c];
        if (depths) depths[c] = bp.depth[c];
    }
    return GH_OK;
}

int gh_stencil_batch_apply(const gh_stencil_batch* b, const void* const* in_dev,
                           void* const* out_dev, const void* const* vt_dev, void* stream) {
    GH_REQUIRE(b && in_dev && out_dev, GH_ERR_ARG, "stencil_batch_apply: null argument");
    const BatchPlan& bp = b->plan;
    GH_REQUIRE(!bp.has_v || vt_dev, GH_ERR_ARG,
               "stencil_batch_apply: the batch was made with potentials");
    for (int c = 0; c < bp.ncells; c++) {
        GH_REQUIRE(in_dev[c] && out_dev[c] && in_dev[c] != out_dev[c], GH_ERR_ARG,
                   "stencil_batch_apply: cell %d: null or aliased arrays", c);
        auto al = [](const void* p) { return ((uintptr_t)p & 15u) == 0; };
        GH_REQUIRE(al(in_dev[c]) && al(out_dev[c]) && (!bp.has_v || (vt_dev[c] && al(vt_dev[c]))),
                   GH_ERR_ARG, "stencil_batch_apply: cell %d: arrays must be 16-byte aligned",
                   c);
    }
    hipStream_t st = (hipStream_t)stream;
    const void* const* vt = bp.has_v ? vt_dev : nullptr;
#define GH_RUN(V, S)                                                            \
    switch (bp.R) {                                                            \
        case 1: return launch_batch<V, S, 1>(bp, in_dev, out_dev, vt, st);      \
        case 2: return launch_batch<V, S, 2>(bp, in_dev, out_dev, vt, st);      \
        case 3: return launch_batch<V, S, 3>(bp, in_dev, out_dev, vt, st);      \
    }
    if (bp.dtype == GH_F32) { GH_RUN(float, float) } else { GH_RUN(double, double) }
#undef GH_RUN
    return GH_ERR_UNSUPPORTED;
}

}  // extern "C"

namespace gh {

int64_t stencil_points(const gh_stencil* h) { return h->n[0] * h->n[1] * h->n[2]; }
int stencil_dtype(const gh_stencil* h) { return h->dtype; }

int stencil_apply_scaled(const gh_stencil* h, const void* in_dev, void* out_dev,
                         int64_t nbands, const double* phase_cd,
                         const void* vt_dev, const void* src_dev, double beta,
                         double al, double ga, hipStream_t st) {
    return stencil_apply_full(h, in_dev, out_dev, nbands, phase_cd, vt_dev, src_dev,
                              beta, al, ga, nullptr, st);
}

int stencil_apply_full(const gh_stencil* h, const void* in_dev, void* out_dev,
                       int64_t nbands, const double* phase_cd,
                       const void* vt_dev, const void* src_dev, double beta,
                       double al, double ga, const double* eps, hipStream_t st) {
    GH_REQUIRE(h, GH_ERR_ARG, "stencil_apply: null handle");
    GH_REQUIRE(nbands >= 0, GH_ERR_ARG, "stencil_apply: nbands < 0");
    if (nbands == 0) return GH_OK;  // empty arrays may carry null pointers
    GH_REQUIRE(in_dev && out_dev, GH_ERR_ARG, "stencil_apply: null argument");
    GH_REQUIRE(in_dev != out_dev, GH_ERR_ARG, "stencil_apply: in and out must differ");
    if (h->cfd) {
        switch (h->dtype) {
            case GH_F64: return launch_cfd<double, double, 4>(h, in_dev, out_dev, nbands, phase_cd, vt_dev, src_dev, beta, al, ga, eps, st);
            case GH_C128: return launch_cfd<cplx<double>, double, 2>(h, in_dev, out_dev, nbands, phase_cd, vt_dev, src_dev, beta, al, ga, eps, st);
            case GH_F32: return launch_cfd<float, float, 8>(h, in_dev, out_dev, nbands, phase_cd, vt_dev, src_dev, beta, al, ga, eps, st);
            case GH_C64: return launch_cfd<cplx<float>, float, 4>(h, in_dev, out_dev, nbands, phase_cd, vt_dev, src_dev, beta, al, ga, eps, st);
        }
    } else {
        GH_REQUIRE(!eps, GH_ERR_UNSUPPORTED,
                   "stencil_apply: per-band shift needs an orthogonal stencil");
        switch (h->dtype) {
            case GH_F64: return launch_generic<double, double>(h, in_dev, out_dev, nbands, phase_cd, vt_dev, src_dev, beta, al, ga, st);
            case GH_C128: return launch_generic<cplx<double>, double>(h, in_dev, out_dev, nbands, phase_cd, vt_dev, src_dev, beta, al, ga, st);
            case GH_F32: return launch_generic<float, float>(h, in_dev, out_dev, nbands, phase_cd, vt_dev, src_dev, beta, al, ga, st);
            case GH_C64: return launch_generic<cplx<float>, float>(h, in_dev, out_dev, nbands, phase_cd, vt_dev, src_dev, beta, al, ga, st);
        }
    }
    set_error("stencil_apply: bad dtype");
    return GH_ERR_ARG;
}

}  // namespace gh
