// Learned preconditioner: the layers of a small 3D U-Net on gfx950 matrix
// cores.  It replaces the multigrid V-cycle of the eigensolver
// (gpaw/preconditioner.py:9-93, called from gpaw/eigensolvers/rmmdiis.py:
// 138-143): d = f(R, v), a correction from the residual R and the local
// potential v.
//
// Activations are BF16, channel-last: act[b][x][y][z][C] (C = 8, 16 or 32)
// on a periodic grid.  Weights are BF16; every product is summed in FP32
// (v_mfma_f32_16x16x16bf16_1k).  A 3x3x3 convolution is a GEMM of
// (points) x (27 taps * C_in) against (27 taps * C_in) x (C_out): a
// wavefront owns 16 points (flattened (y, z) of one x-plane) and all output
// channels; for each k-chunk of 16 (tap, channel) pairs each lane gathers
// the 4 channels of one neighbour of its point (8 bytes) and the matching
// 4 x 16 weights, pre-packed on the host in the lane order of the
// instruction.  Periodic neighbours wrap by index.  Epilogue: bias, ReLU,
// BF16 store.
//
// Input modes of the convolution:
//   FIRST  channel 0 = s_b R[b][p], channel 1 = v_scale v[p] (FP32 or FP64
//          inputs, band-major grids), no other channels (C_in = 2);
//   PLAIN  act_in[b][p][:];
//   UPADD  act_in[b][p / 2][:] + skip[b][p][:] (nearest upsampling of a
//          coarse tensor plus a skip connection, summed in FP32).
// The remaining layers (2x2x2 average pooling, 1x1 channel maps, the final
// 1x1 map to the FP32 / FP64 correction with the per-band scale) are
// memory-bound element kernels.
#include <algorithm>
#include <cstring>

#include "common.h"

namespace gh {

namespace {

typedef short bf4 __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(unsigned short h) {
    return __uint_as_float((unsigned)h << 16);
}
// round to nearest even
__device__ __forceinline__ unsigned short f2bf(float f) {
    unsigned u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)(u >> 16);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

enum { kFirst = 0, kPlain = 1, kUpAdd = 2 };

struct ConvGeom {
    int nx, ny, nz;      // output (and input, except the coarse UPADD input) grid
    int nbands;
};

template <typename T>
struct FirstIn {
    const T* R;          // (nbands, nx, ny, nz)
    const T* v;          // (nx, ny, nz)
    const float* scale;  // per band
    float vscale;
};

// 3x3x3 periodic convolution, C_in -> C_out (C_out multiple of 8, <= 32).
// Grid: (ceil(ny * nz / 64), nx, nbands); block 256 (4 waves x 16 points).
// W: packed bf16 [ct][kc][lane][4], NK = ceil(27 C_in / 16) chunks,
// NCT = ceil(C_out / 16) column tiles.
template <int CIN, int COUT, int MODE, typename T>
__global__ void __launch_bounds__(256)
conv3_kernel(ConvGeom g, const unsigned short* __restrict__ in,
             const unsigned short* __restrict__ skip, FirstIn<T> fin,
             const unsigned short* __restrict__ W, const float* __restrict__ bias,
             int relu, unsigned short* __restrict__ out) {
    constexpr int NK = (27 * CIN + 15) / 16;
    constexpr int NCT = (COUT + 15) / 16;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = lane >> 4, row = lane & 15;
    const int b = blockIdx.z, x = blockIdx.y;
    const int nyz = g.ny * g.nz;
    const int p0 = (blockIdx.x * 4 + wave) * 16;          // first point of the wave
    if (p0 >= nyz) return;                                  // whole wave idle
    const int p = p0 + row;                                 // this lane's A row
    const bool pv = p < nyz;
    const int y = pv ? p / g.nz : 0, z = pv ? p - (p / g.nz) * g.nz : 0;
    const int64_t plane = (int64_t)nyz;
    const int cnx = g.nx / 2, cny = g.ny / 2, cnz = g.nz / 2;   // UPADD coarse grid

    f4v acc[NCT];
#pragma unroll
    for (int t = 0; t < NCT; t++) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};

    for (int kc = 0; kc < NK; kc++) {
        bf4 a = bf4{0, 0, 0, 0};
        const int k0 = 16 * kc + 4 * grp;
        if (MODE == kFirst) {
            // C_in = 2: the 4 k of a lane span two taps
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int k = k0 + e;
                const int tap = k >> 1, ci = k & 1;
                float v = 0.f;
                if (pv && tap < 27) {
                    const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
                    int xx = x + dx, yy = y + dy, zz = z + dz;
                    xx += xx < 0 ? g.nx : (xx >= g.nx ? -g.nx : 0);
                    yy += yy < 0 ? g.ny : (yy >= g.ny ? -g.ny : 0);
                    zz += zz < 0 ? g.nz : (zz >= g.nz ? -g.nz : 0);
                    const int64_t q = ((int64_t)xx * g.ny + yy) * g.nz + zz;
                    v = ci == 0 ? (float)fin.R[(int64_t)b * g.nx * plane + q] * fin.scale[b]
                                : (float)fin.v[q] * fin.vscale;
                }
                a[e] = (short)f2bf(v);
            }
        } else {
            const int tap = k0 / CIN, ci = k0 - tap * CIN;
            if (pv && tap < 27) {
                const int dx = tap / 9 - 1, dy = (tap / 3) % 3 - 1, dz = tap % 3 - 1;
                int xx = x + dx, yy = y + dy, zz = z + dz;
                xx += xx < 0 ? g.nx : (xx >= g.nx ? -g.nx : 0);
                yy += yy < 0 ? g.ny : (yy >= g.ny ? -g.ny : 0);
                zz += zz < 0 ? g.nz : (zz >= g.nz ? -g.nz : 0);
                if (MODE == kPlain) {
                    const int64_t q = (((int64_t)b * g.nx + xx) * g.ny + yy) * g.nz + zz;
                    a = *reinterpret_cast<const bf4*>(in + q * CIN + ci);
                } else {
                    const int64_t q = (((int64_t)b * g.nx + xx) * g.ny + yy) * g.nz + zz;
                    const int64_t qc =
                        (((int64_t)b * cnx + (xx >> 1)) * cny + (yy >> 1)) * cnz + (zz >> 1);
                    const bf4 s = *reinterpret_cast<const bf4*>(skip + q * CIN + ci);
                    const bf4 c = *reinterpret_cast<const bf4*>(in + qc * CIN + ci);
#pragma unroll
                    for (int e = 0; e < 4; e++)
                        a[e] = (short)f2bf(bf2f((unsigned short)s[e]) +
                                           bf2f((unsigned short)c[e]));
                }
            }
        }
#pragma unroll
        for (int t = 0; t < NCT; t++) {
            const bf4 w = *reinterpret_cast<const bf4*>(W + (((int64_t)t * NK + kc) * 64 + lane) * 4);
            acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w, acc[t], 0, 0, 0);
        }
    }
    // D[i][j]: point p0 + 4 grp + r, channel t * 16 + row
#pragma unroll
    for (int t = 0; t < NCT; t++) {
        const int co = t * 16 + row;
        if (co >= COUT) continue;
        const float bb = bias[co];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int pp = p0 + 4 * grp + r;
            if (pp >= nyz) continue;
            float v = acc[t][r] + bb;
            if (relu) v = fmaxf(v, 0.f);
            out[(((int64_t)b * g.nx + x) * plane + pp) * COUT + co] = f2bf(v);
        }
    }
}

// The same convolution with the input slab staged in LDS: a block owns
// TY full rows (all z) of one x-plane and stages the 3 x-planes x (TY + 2)
// rows x (nz + 2) points (periodic wrap applied while staging) of its
// input once -- for UPADD the upsampled sum and for FIRST the scaled R and
// v (as 4 channels: R, v, 0, 0) are formed there, once per input point
// instead of once per use -- then every wave walks its 16-point tiles of
// the block reading the A operand from LDS.  CL: staged channels (CIN, or
// 4 for FIRST).
template <int CIN, int COUT, int MODE, typename T>
__global__ void __launch_bounds__(256)
conv3_lds_kernel(ConvGeom g, int ty, const unsigned short* __restrict__ in,
                 const unsigned short* __restrict__ skip, FirstIn<T> fin,
                 const unsigned short* __restrict__ W, const float* __restrict__ bias,
                 int relu, unsigned short* __restrict__ out) {
    constexpr int CL = MODE == kFirst ? 4 : CIN;
    constexpr int NK = (27 * CL + 15) / 16;
    constexpr int NCT = (COUT + 15) / 16;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned short* sIn = reinterpret_cast<unsigned short*>(smem_raw);
    const int b = blockIdx.z, x = blockIdx.y;
    const int y0 = blockIdx.x * ty;
    const int tye = min(ty, g.ny - y0);
    const int nz = g.nz, nzp = nz + 2, rows = tye + 2;
    const int64_t plane = (int64_t)g.ny * nz;
    const int cnx = g.nx / 2, cny = g.ny / 2, cnz = g.nz / 2;
    // ---- stage: (3, rows, nzp) points x CL channels, 4 channels per item
    const int nitem = 3 * rows * nzp * (CL / 4);
    for (int it = threadIdx.x; it < nitem; it += blockDim.x) {
        const int c4 = it % (CL / 4);
        int q = it / (CL / 4);
        const int zl = q % nzp; q /= nzp;
        const int yl = q % rows;
        const int dx = q / rows;
        int xx = x + dx - 1, yy = y0 + yl - 1, zz = zl - 1;
        xx += xx < 0 ? g.nx : (xx >= g.nx ? -g.nx : 0);
        yy += yy < 0 ? g.ny : (yy >= g.ny ? -g.ny : 0);
        zz += zz < 0 ? g.nz : (zz >= g.nz ? -g.nz : 0);
        const int64_t pt = ((int64_t)xx * g.ny + yy) * nz + zz;   // in the band
        bf4 a;
        if (MODE == kFirst) {
            const float r = (float)fin.R[(int64_t)b * g.nx * plane + pt] * fin.scale[b];
            const float v = (float)fin.v[pt] * fin.vscale;
            a = bf4{(short)f2bf(r), (short)f2bf(v), 0, 0};
        } else if (MODE == kPlain) {
            a = *reinterpret_cast<const bf4*>(in + ((int64_t)b * g.nx * plane + pt) * CIN +
                                              4 * c4);
        } else {
            const int64_t qc =
                (((int64_t)b * cnx + (xx >> 1)) * cny + (yy >> 1)) * cnz + (zz >> 1);
            const bf4 sv = *reinterpret_cast<const bf4*>(
                skip + ((int64_t)b * g.nx * plane + pt) * CIN + 4 * c4);
            const bf4 cv = *reinterpret_cast<const bf4*>(in + qc * CIN + 4 * c4);
#pragma unroll
            for (int e = 0; e < 4; e++)
                a[e] = (short)f2bf(bf2f((unsigned short)sv[e]) + bf2f((unsigned short)cv[e]));
        }
        *reinterpret_cast<bf4*>(sIn + (((dx * rows + yl) * nzp + zl) * CL + 4 * c4)) = a;
    }
    __syncthreads();
    // ---- products: 16-point tiles of the block's rows x nz points
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int grp = lane >> 4, row = lane & 15;
    const int npts = tye * nz;
    // small layers keep their B operands in registers for all the wave's
    // tiles (the fine levels, where the tiles are many)
    constexpr bool WREG = NK * NCT <= 28;
    bf4 wr[WREG ? NK * NCT : 1];
    if constexpr (WREG) {
#pragma unroll
        for (int i = 0; i < NK * NCT; i++)
            wr[i] = *reinterpret_cast<const bf4*>(W + ((int64_t)i * 64 + lane) * 4);
    }
    for (int p0 = wave * 16; p0 < npts; p0 += 64) {
        const int p = p0 + row;
        const bool pv = p < npts;
        const int yl = pv ? p / nz : 0, z = pv ? p - (p / nz) * nz : 0;
        f4v acc[NCT];
#pragma unroll
        for (int t = 0; t < NCT; t++) acc[t] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kc = 0; kc < NK; kc++) {
            const int k0 = 16 * kc + 4 * grp;
            const int tap = k0 / CL, ci = k0 - tap * CL;
            bf4 a = bf4{0, 0, 0, 0};
            if (pv && tap < 27) {
                const int dx = tap / 9, dy = (tap / 3) % 3, dz = tap % 3;
                a = *reinterpret_cast<const bf4*>(
                    sIn + (((dx * rows + yl + dy) * nzp + z + dz) * CL + ci));
            }
#pragma unroll
            for (int t = 0; t < NCT; t++) {
                bf4 w;
                if constexpr (WREG) w = wr[t * NK + kc];
                else w = *reinterpret_cast<const bf4*>(W + (((int64_t)t * NK + kc) * 64 + lane) * 4);
                acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w, acc[t], 0, 0, 0);
            }
        }
#pragma unroll
        for (int t = 0; t < NCT; t++) {
            const int co = t * 16 + row;
            if (co >= COUT) continue;
            const float bb = bias[co];
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int pp = p0 + 4 * grp + r;
                if (pp >= npts) continue;
                float v = acc[t][r] + bb;
                if (relu) v = fmaxf(v, 0.f);
                out[(((int64_t)b * g.nx + x) * plane + (int64_t)y0 * nz + pp) * COUT + co] =
                    f2bf(v);
            }
        }
    }
}

// 2x2x2 average of a channel-last BF16 tensor (fine grid nx, ny, nz even).
template <int C>
__global__ void __launch_bounds__(256)
pool2_kernel(const unsigned short* __restrict__ in, unsigned short* __restrict__ out,
             int nx, int ny, int nz, int64_t total) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int c = (int)(t % C);
    int64_t q = t / C;
    const int cz = nz / 2, cy = ny / 2, cx = nx / 2;
    const int Z = (int)(q % cz); q /= cz;
    const int Y = (int)(q % cy); q /= cy;
    const int X = (int)(q % cx);
    const int64_t b = q / cx;
    float s = 0.f;
#pragma unroll
    for (int d = 0; d < 8; d++) {
        const int xx = 2 * X + (d >> 2), yy = 2 * Y + ((d >> 1) & 1), zz = 2 * Z + (d & 1);
        s += bf2f(in[(((b * nx + xx) * ny + yy) * nz + zz) * C + c]);
    }
    out[t] = f2bf(0.125f * s);
}

// 1x1 channel map C_in -> C_out (FP32 weights W[ci][co], bias), BF16 in/out.
template <int CIN, int COUT>
__global__ void __launch_bounds__(256)
map1_kernel(const unsigned short* __restrict__ in, const float* __restrict__ W,
            const float* __restrict__ bias, unsigned short* __restrict__ out, int64_t npts) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= npts * COUT) return;
    const int co = (int)(t % COUT);
    const int64_t p = t / COUT;
    float s = bias[co];
#pragma unroll
    for (int ci = 0; ci < CIN; ci++) s += W[ci * COUT + co] * bf2f(in[p * CIN + ci]);
    out[t] = f2bf(s);
}

// final 1x1 map C -> 1 into the correction, times the inverse band scale
template <int C, typename T>
__global__ void __launch_bounds__(256)
head_kernel(const unsigned short* __restrict__ in, const float* __restrict__ W, float bias,
            const float* __restrict__ scale, int64_t npb, int64_t total, T* __restrict__ out) {
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int64_t b = t / npb;
    float s = bias;
#pragma unroll
    for (int ci = 0; ci < C; ci++) s += W[ci] * bf2f(in[t * C + ci]);
    out[t] = (T)(s / scale[b]);
}

constexpr int kConvLds = 60 * 1024;   // LDS budget of the staged slab

template <int CIN, int COUT, int MODE, typename T>
int launch_conv3(const ConvGeom& g, const void* in, const void* skip, const FirstIn<T>& fin,
                 const void* W, const float* bias, int relu, void* out, hipStream_t st) {
    constexpr int CL = MODE == kFirst ? 4 : CIN;
    const int64_t per_row = 3LL * (g.nz + 2) * CL * 2;
    const int ty = (int)std::min<int64_t>(g.ny, kConvLds / per_row - 2);
    if (ty >= 1) {
        const size_t lds = (size_t)per_row * (ty + 2);
        dim3 grid((unsigned)((g.ny + ty - 1) / ty), (unsigned)g.nx, (unsigned)g.nbands);
        hipLaunchKernelGGL((conv3_lds_kernel<CIN, COUT, MODE, T>), grid, dim3(256), lds, st,
                           g, ty, (const unsigned short*)in, (const unsigned short*)skip, fin,
                           (const unsigned short*)W, bias, relu, (unsigned short*)out);
        GH_CHECK_LAUNCH();
        return GH_OK;
    }
    if (MODE == kFirst) {
        set_error("unet_conv3: rows of %d points do not fit the LDS slab", g.nz);
        return GH_ERR_UNSUPPORTED;
    }
    const int nyz = g.ny * g.nz;
    dim3 grid((unsigned)((nyz + 63) / 64), (unsigned)g.nx, (unsigned)g.nbands);
    hipLaunchKernelGGL((conv3_kernel<CIN, COUT, MODE, T>), grid, dim3(256), 0, st, g,
                       (const unsigned short*)in, (const unsigned short*)skip, fin,
                       (const unsigned short*)W, bias, relu, (unsigned short*)out);
    GH_CHECK_LAUNCH();
    return GH_OK;
}

}  // namespace

}  // namespace gh

using namespace gh;

extern "C" {

int64_t gh_unet_conv3_weight_elems(int cin, int cout) {
    if (cin < 1 || cout < 1) return -1;
    const int64_t nk = (27 * cin + 15) / 16, nct = (cout + 15) / 16;
    return nct * nk * 64 * 4;
}

// Pack W[co][ci][kx][ky][kz] (float, torch's conv3d layout) into the lane
// order of the matrix-core instruction (bf16 bits), k = tap * cin + ci with
// tap = (kx * 3 + ky) * 3 + kz.
int gh_unet_pack_conv3(int cin, int cout, const float* W, uint16_t* packed) {
    GH_REQUIRE(W && packed && cin >= 1 && cout >= 1, GH_ERR_ARG, "unet_pack: bad argument");
    const int nk = (27 * cin + 15) / 16, nct = (cout + 15) / 16;
    for (int t = 0; t < nct; t++)
        for (int kc = 0; kc < nk; kc++)
            for (int lane = 0; lane < 64; lane++)
                for (int e = 0; e < 4; e++) {
                    const int k = 16 * kc + 4 * (lane >> 4) + e;
                    const int co = t * 16 + (lane & 15);
                    float v = 0.f;
                    if (k < 27 * cin && co < cout) {
                        const int tap = k / cin, ci = k % cin;
                        v = W[((int64_t)co * cin + ci) * 27 + tap];
                    }
                    uint32_t u;
                    std::memcpy(&u, &v, 4);
                    if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
                    packed[(((int64_t)t * nk + kc) * 64 + lane) * 4 + e] = (uint16_t)(u >> 16);
                }
    return GH_OK;
}

int gh_unet_conv3(int cin, int cout, int mode, int dtype, int nbands, const int64_t* n_c,
                  const void* in_dev, const void* skip_dev, const void* R_dev,
                  const void* v_dev, const float* scale_dev, double vscale,
                  const void* W_dev, const float* bias_dev, int relu, void* out_dev,
                  void* stream) {
    GH_REQUIRE(n_c && W_dev && bias_dev && out_dev && nbands >= 0, GH_ERR_ARG,
               "unet_conv3: null argument");
    ConvGeom g;
    g.nx = (int)n_c[0]; g.ny = (int)n_c[1]; g.nz = (int)n_c[2];
    g.nbands = nbands;
    GH_REQUIRE(g.nx >= 1 && g.ny >= 1 && g.nz >= 1 && (int64_t)g.ny * g.nz < (1 << 30),
               GH_ERR_SHAPE, "unet_conv3: bad grid");
    if (nbands == 0) return GH_OK;
    if (mode == kUpAdd)
        GH_REQUIRE(g.nx % 2 == 0 && g.ny % 2 == 0 && g.nz % 2 == 0 && skip_dev && in_dev,
                   GH_ERR_SHAPE, "unet_conv3: upsampling needs an even grid and a skip");
    hipStream_t st = (hipStream_t)stream;
    FirstIn<float> f32{};
    FirstIn<double> f64{};
    if (mode == kFirst) {
        GH_REQUIRE(R_dev && v_dev && scale_dev && cin == 2, GH_ERR_ARG,
                   "unet_conv3: the first layer takes R, v and the band scales "
                   "(weights packed for 4 input channels, the last two zero)");
        GH_REQUIRE(dtype == GH_F32 || dtype == GH_F64, GH_ERR_UNSUPPORTED,
                   "unet_conv3: real residuals only");
        f32 = FirstIn<float>{(const float*)R_dev, (const float*)v_dev, scale_dev, (float)vscale};
        f64 = FirstIn<double>{(const double*)R_dev, (const double*)v_dev, scale_dev,
                              (float)vscale};
    } else {
        GH_REQUIRE(in_dev, GH_ERR_ARG, "unet_conv3: null input");
    }
#define GH_CONV(CI, CO, MD)                                                          \
    if (cin == CI && cout == CO && mode == MD) {                                     \
        if (MD == kFirst && dtype == GH_F64)                                         \
            return launch_conv3<CI, CO, MD, double>(g, in_dev, skip_dev, f64, W_dev, \
                                                    bias_dev, relu, out_dev, st);    \
        return launch_conv3<CI, CO, MD, float>(g, in_dev, skip_dev, f32, W_dev,      \
                                               bias_dev, relu, out_dev, st);         \
    }
    GH_CONV(2, 8, kFirst)
    GH_CONV(8, 16, kPlain)
    GH_CONV(16, 32, kPlain)
    GH_CONV(32, 32, kPlain)
    GH_CONV(16, 16, kUpAdd)
    GH_CONV(8, 8, kUpAdd)
    GH_CONV(8, 8, kPlain)
    GH_CONV(16, 16, kPlain)
#undef GH_CONV
    set_error("unet_conv3: layer %d -> %d (mode %d) not built", cin, cout, mode);
    return GH_ERR_UNSUPPORTED;
}

int gh_unet_pool2(int c, int nbands, const int64_t* n_c, const void* in_dev, void* out_dev,
                  void* stream) {
    GH_REQUIRE(n_c && in_dev && out_dev, GH_ERR_ARG, "unet_pool2: null argument");
    const int nx = (int)n_c[0], ny = (int)n_c[1], nz = (int)n_c[2];
    GH_REQUIRE(nx % 2 == 0 && ny % 2 == 0 && nz % 2 == 0, GH_ERR_SHAPE,
               "unet_pool2: odd grid");
    const int64_t total = (int64_t)nbands * (nx / 2) * (ny / 2) * (nz / 2) * c;
    if (total == 0) return GH_OK;
    hipStream_t st = (hipStream_t)stream;
    const unsigned nblk = (unsigned)((total + 255) / 256);
    switch (c) {
        case 8: hipLaunchKernelGGL((pool2_kernel<8>), dim3(nblk), dim3(256), 0, st, (const unsigned short*)in_dev, (unsigned short*)out_dev, nx, ny, nz, total); break;
        case 16: hipLaunchKernelGGL((pool2_kernel<16>), dim3(nblk), dim3(256), 0, st, (const unsigned short*)in_dev, (unsigned short*)out_dev, nx, ny, nz, total); break;
        case 32: hipLaunchKernelGGL((pool2_kernel<32>), dim3(nblk), dim3(256), 0, st, (const unsigned short*)in_dev, (unsigned short*)out_dev, nx, ny, nz, total); break;
        default: set_error("unet_pool2: %d channels", c); return GH_ERR_UNSUPPORTED;
    }
    GH_CHECK_LAUNCH();
    return GH_OK;
}

int gh_unet_map1(int cin, int cout, int64_t npts, const void* in_dev, const float* W_dev,
                 const float* bias_dev, void* out_dev, void* stream) {
    GH_REQUIRE(in_dev && W_dev && bias_dev && out_dev && npts >= 0, GH_ERR_ARG,
               "unet_map1: bad argument");
    if (npts == 0) return GH_OK;
    hipStream_t st = (hipStream_t)stream;
    const unsigned nblk = (unsigned)((npts * cout + 255) / 256);
    const unsigned short* in = (const unsigned short*)in_dev;
    unsigned short* out = (unsigned short*)out_dev;
    if (cin == 32 && cout == 16)
        hipLaunchKernelGGL((map1_kernel<32, 16>), dim3(nblk), dim3(256), 0, st, in, W_dev, bias_dev, out, npts);
    else if (cin == 16 && cout == 8)
        hipLaunchKernelGGL((map1_kernel<16, 8>), dim3(nblk), dim3(256), 0, st, in, W_dev, bias_dev, out, npts);
    else {
        set_error("unet_map1: %d -> %d not built", cin, cout);
        return GH_ERR_UNSUPPORTED;
    }
    GH_CHECK_LAUNCH();
    return GH_OK;
}

int gh_unet_head(int cin, int dtype, int nbands, int64_t npts_per_band, const void* in_dev,
                 const float* W_dev, double bias, const float* scale_dev, void* out_dev,
                 void* stream) {
    GH_REQUIRE(in_dev && W_dev && scale_dev && out_dev && nbands >= 0, GH_ERR_ARG,
               "unet_head: bad argument");
    GH_REQUIRE(cin == 8, GH_ERR_UNSUPPORTED, "unet_head: %d channels not built", cin);
    const int64_t total = (int64_t)nbands * npts_per_band;
    if (total == 0) return GH_OK;
    hipStream_t st = (hipStream_t)stream;
    const unsigned nblk = (unsigned)((total + 255) / 256);
    const unsigned short* in = (const unsigned short*)in_dev;
    if (dtype == GH_F32)
        hipLaunchKernelGGL((head_kernel<8, float>), dim3(nblk), dim3(256), 0, st, in, W_dev,
                           (float)bias, scale_dev, npts_per_band, total, (float*)out_dev);
    else if (dtype == GH_F64)
        hipLaunchKernelGGL((head_kernel<8, double>), dim3(nblk), dim3(256), 0, st, in, W_dev,
                           (float)bias, scale_dev, npts_per_band, total, (double*)out_dev);
    else {
        set_error("unet_head: real outputs only");
        return GH_ERR_UNSUPPORTED;
    }
    GH_CHECK_LAUNCH();
    return GH_OK;
}

}  // extern "C"
