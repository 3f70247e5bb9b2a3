/* gpaw_amd.h -- C ABI of the MI355X (gfx950) finite-difference H-psi kernels.
 *
 * Plain pointers and sizes only.  Array arguments named *_dev are device
 * pointers (HBM); everything else is host memory.  Grids are C-contiguous
 * (nbands, n0, n1, n2) with n2 fastest; complex data is interleaved
 * (re, im) like numpy complex128/complex64.  Bloch phases phase_cd are host
 * arrays of 3x2 complex numbers (12 doubles: [c][d][re, im]) with
 * phase_cd[c][0] applied to data that wraps from the low face of axis c into
 * the high ghost and phase_cd[c][1] to data wrapping from the high face into
 * the low ghost (reference c/bc.c:242-257, gpaw/kpt_descriptor.py:283-284);
 * NULL means all phases are 1.
 *
 * Every function returns GH_OK (0) or a negative error code; gh_last_error()
 * then returns a message.  Work is enqueued on the given HIP stream (NULL =
 * default stream) and is asynchronous unless stated otherwise.  Handles are
 * immutable after creation and may be used from several streams at once;
 * scratch memory is owned by the caller or allocated per handle as stated.
 * A call with nbands == 0 (an empty band set, as the reference's loops over
 * the leading axis allow) returns GH_OK without touching memory; its array
 * pointers may then be NULL, as empty device arrays' pointers are.
 * gh_lfc_apply_dH, gh_multi_dotc, gh_multi_gram and gh_multi_lincomb launch
 * one grid row per band; they split band sets above 32768 into slices of
 * that size themselves (one launch sequence per slice, in stream order).
 */
#ifndef GPAW_AMD_H
#define GPAW_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GH_OK 0
#define GH_ERR_ARG (-1)
#define GH_ERR_SHAPE (-2)
#define GH_ERR_HIP (-3)
#define GH_ERR_UNSUPPORTED (-4)
#define GH_ERR_ALLOC (-5)

/* element types */
#define GH_F64 0  /* float64            */
#define GH_C128 1 /* complex128         */
#define GH_F32 2  /* float32            */
#define GH_C64 3  /* complex64          */

typedef struct gh_stencil gh_stencil;
typedef struct gh_transformer gh_transformer;
typedef struct gh_lfc gh_lfc;
typedef struct gh_stencil_batch gh_stencil_batch;

/* Last error message of the calling thread ("" if none). */
const char* gh_last_error(void);

/* Library version string and the gfx target it was built for. */
const char* gh_version(void);

/* ----------------------------------------------------------- stencils --
 * Replaces cgpaw.Operator (reference c/operators.c:384-425, created by
 * gpaw/fd_operators.py:83-86) and its apply (c/operators.c:250-321,
 * bc.c:146-308 halo fill + bmgs/fd.c:9-33).
 *
 * coefs[ncoefs], offsets_pc[ncoefs*3]: stencil points as in FDOperator
 * (fd_operators.py:36-91), offsets in grid units per axis.  n_c: grid shape;
 * pbc_c: 1 = periodic (Bloch) halo, 0 = zero halo.  dtype: GH_*.
 * The largest |offset| along an axis must not exceed n_c on that axis
 * (fd_operators.py:53).
 */
int gh_stencil_create(gh_stencil** out, int ncoefs, const double* coefs,
                      const int64_t* offsets_pc, const int64_t* n_c,
                      const int32_t* pbc_c, int dtype);
int gh_stencil_destroy(gh_stencil* h);

/* Coefficient at offset (0,0,0) -- Operator.get_diagonal_element
 * (operators.c:324-337). */
double gh_stencil_diagonal(const gh_stencil* h);

/* out = L in  (+ vt .* in if vt_dev) (+ beta * src if src_dev).
 * in_dev/out_dev: (nbands, n0, n1, n2) of the stencil's dtype; in != out.
 * The fastest kernel needs in, out, vt and src 16-byte aligned and rows of
 * a multiple of 16 bytes; other pointers and shapes take a slower kernel.
 * vt_dev: local potential (n0, n1, n2), real, same precision as the data
 * (double for F64/C128, float for F32/C64) -- fused
 * Hamiltonian.apply_local_potential (hamiltonian.py:437-443).
 * src_dev: optional (nbands, n0, n1, n2) array of the data type. */
int gh_stencil_apply(const gh_stencil* h, const void* in_dev, void* out_dev,
                     int64_t nbands, const double* phase_cd,
                     const void* vt_dev, const void* src_dev, double beta,
                     void* stream);

/* out[x] = (L + vt - eps[x]) in[x] per band: the local part of the
 * residual R = (H - eps S) psi in one stencil pass -- replaces
 * apply_pseudo_hamiltonian (wavefunctions/fd.py:82-88) followed by the
 * -eps psi axpy of Eigensolver.calculate_residuals (eigensolver.py:153-172;
 * the non-local part is added by gh_lfc_add).  eps_dev: nbands doubles on
 * the device.  vt_dev may be null.  Orthogonal-cell stencils only
 * (GH_ERR_UNSUPPORTED otherwise). */
int gh_stencil_apply_shifted(const gh_stencil* h, const void* in_dev,
                             void* out_dev, int64_t nbands,
                             const double* phase_cd, const void* vt_dev,
                             const double* eps_dev, void* stream);

/* Tile choice of the stencil kernel: rows of the (y, z) tile per block,
 * planes per block along x, and the prefetch depth in planes (2 or 3) of
 * the LDS-ring kernel; 0 = automatic.  Used by the autotuner. */
int gh_stencil_set_tile(gh_stencil* h, int rows, int planes, int depth);

/* Jacobi relaxation of L f = s in place (replaces Operator.relax,
 * c/operators.c:55-95 with c/bmgs/relax.c, relax_method 2): nrelax sweeps
 *   f <- (1 - w) f + w (s - sum_{c != 0} coef_c f[shift_c]) / coef_0
 * with the boundary conditions of gh_stencil_apply (no Bloch phases).
 * One real grid (n0, n1, n2); work_dev: a device array of the same size.
 * relax_method 1 (Gauss-Seidel) returns GH_ERR_UNSUPPORTED. */
int gh_stencil_relax(const gh_stencil* h, int relax_method, void* f_dev,
                     const void* s_dev, void* work_dev, int nrelax, double w,
                     void* stream);

/* ------------------------------------------------------ batched cells --
 * Several independent grids ("cells") of different sizes and band counts in
 * one persistent launch of the LDS-ring stencil: out_c = L_c in_c
 * (+ vt_c .* in_c).  No single reference counterpart: each cell is one
 * Operator.apply (c/operators.c:250-321; the reference's device-pointer
 * form c/gpu/operators.c:350-391 launches one grid per call); the batch
 * issues all of them as one kernel whose blocks walk a global tile list,
 * largest tiles first.  Cells: 1..8 real float32 or float64 axis-aligned
 * stencils (radius <= 3) with rows of a multiple of 16 bytes.  rows[c],
 * depths[c]: tile rows and prefetch depth per cell (0 / NULL: the largest
 * tile at depth 2), e.g. from an autotuner.  with_potential: apply takes
 * one potential per cell.  The plan (tile table and a tile counter) lives
 * on the device; the apply call only passes pointers, so it can be captured
 * in a HIP graph.  Blocks draw tiles from the counter (largest tiles first)
 * and the last block resets it: launches of one batch must not overlap in
 * time (one stream, or ordered streams). */
int gh_stencil_batch_create(gh_stencil_batch** out, int ncells,
                            const gh_stencil* const* ops, const int64_t* nbands,
                            const int32_t* rows, const int32_t* depths,
                            int with_potential);
int gh_stencil_batch_destroy(gh_stencil_batch* b);
/* info[5] = {blocks, threads per block, LDS bytes, tiles, tiles per block};
 * rows / depths (may be NULL) receive the tile chosen per cell. */
int gh_stencil_batch_info(const gh_stencil_batch* b, int64_t* info,
                          int32_t* rows, int32_t* depths);
/* in_dev[c], out_dev[c]: (nbands_c, n0, n1, n2) of cell c, 16-byte aligned;
 * vt_dev[c]: (n0, n1, n2) potential of cell c (with_potential). */
int gh_stencil_batch_apply(const gh_stencil_batch* b, const void* const* in_dev,
                           void* const* out_dev, const void* const* vt_dev,
                           void* stream);

/* -------------------------------------------------------- transformers --
 * Replaces cgpaw.Transformer (reference c/transformers.c:202-252, created
 * by gpaw/transformers.py:68-76) and its apply (transformers.c:118-164 ->
 * bmgs/restrict.c, bmgs/interpolate.c).
 *
 * nin_c/nout_c: local grid shapes; nn: 1..4 (K = 2 nn); pad_cd[6]: ghost
 * widths (low, high) per axis of the padded input; skip_cd[6]: interpolation
 * skips; pbc_c: periodic axes; interpolate: 0 restrict / 1 interpolate.
 * These are exactly the values transformers.py:27-47 computes.
 */
int gh_transformer_create(gh_transformer** out, const int64_t* nin_c,
                          const int64_t* nout_c, int nn,
                          const int64_t* pad_cd, const int64_t* skip_cd,
                          const int32_t* pbc_c, int interpolate, int dtype);
int gh_transformer_destroy(gh_transformer* h);
int gh_transformer_apply(const gh_transformer* h, const void* in_dev,
                         void* out_dev, int64_t nbands,
                         const double* phase_cd, void* stream);

/* ----------------------------------------------------------------- LFC --
 * Replaces cgpaw.LFC (reference c/lfc.c:146-257, built by lfc.py:273-369)
 * with its integrate (lfc.c:432-505) and add (lfc.c:818-885).
 *
 * The caller describes each atom's functions on its grid points as
 * "entries": for atom a, entries [entry_ptr[a], entry_ptr[a+1]) give the
 * flattened grid index G_e (into n0*n1*n2), the periodic-image index w_e
 * (into the phase tables, -1 = no phase) and nm_a = M_ptr[a+1]-M_ptr[a]
 * function values A[e][m] (A_e stored row-major with stride nm_a, atoms
 * concatenated in order).  M_ptr[natoms+1] is the global function numbering
 * (lfc.py:277-331: atom-major, spline-major, m-minor).  phase_qw[nq][nw]
 * complex (2 doubles each) is exp(2 pi i k_q . sdisp_w) for each of the nq
 * k-points (phase_qW, lfc.py:339-340), or NULL for real data.  dv is the
 * grid volume element.  All inputs are host arrays; the handle keeps device
 * copies.
 */
int gh_lfc_create(gh_lfc** out, int64_t ngpts, int natoms,
                  const int64_t* entry_ptr, const int32_t* M_ptr,
                  const int32_t* G_e, const int32_t* w_e, const double* A,
                  int nq, int nw, const double* phase_qw, double dv,
                  int dtype);
int gh_lfc_destroy(gh_lfc* h);
int gh_lfc_nfunctions(const gh_lfc* h);

/* Per-geometry setup on the host (no device work):
 * gh_spline_coefs -- replaces cgpaw.Spline construction (c/spline.c:151-165
 * -> bmgs/spline.c:11-47): f[nbins+1] tabulated on [0, rcut] with
 * dr = rcut / nbins -> out[4 * (nbins + 1)] cubic coefficients per bin.
 * gh_spline_to_grid -- replaces cgpaw.spline_to_grid (lfc.c:887-977): for
 * the box [beg_c, end_c) (grid indices), the points with |r - pos_v| < rcut
 * in C order get A_out[p*(2l+1) + m] = spline(r) * r^l Y_lm (l <= 2, m order
 * of bmgs/spherical_harmonics.h) and G_out[p] = flattened index relative to
 * gdcorner_c in a grid of shape n_c.  h_cv is the 3x3 grid-spacing matrix.
 * A_out/G_out need room for every box point.  Returns the point count. */
int gh_spline_coefs(int nbins, double dr, const double* f, double* out);
int64_t gh_spline_to_grid(int l, int nbins, double dr, const double* coefs,
                          const int64_t* beg_c, const int64_t* end_c,
                          const double* pos_v, const double* h_cv,
                          const int64_t* n_c, const int64_t* gdcorner_c,
                          double* A_out, int32_t* G_out);

/* c[x, M] = dv * sum_G a[x, G] A_M(G) [* phase_q]   (written, not added)
 * a_dev: (nbands, ngpts); c_dev: (nbands, Mmax), same dtype.  q: k-point
 * index into the phase tables (0 when there are none), as the q argument of
 * the reference's LFC.integrate_gpu (c/gpu/lfc.cpp:465-503). */
int gh_lfc_integrate(const gh_lfc* h, const void* a_dev, void* c_dev,
                     int64_t nbands, int q, void* stream);

/* a[x, G] += sum_M c[x, M] A_M(G) [* conj(phase_q)]  (LFC.add_gpu,
 * lfc.py:416-419).  One launch over the grid points of the spheres' union:
 * one read-modify-write per point, the contributions of several atoms (or
 * periodic images) summed in a fixed order (atoms ascending).  With
 * GPAW_AMD_LFC_ADD=colour in the environment at gh_lfc_create (or more
 * than 32 functions on one atom), per-atom launches instead, coloured so
 * that no two launches' atoms share a point. */
int gh_lfc_add(const gh_lfc* h, const void* c_dev, void* a_dev,
               int64_t nbands, int q, void* stream);

/* dst = src, nbytes (a multiple of 16, 16-byte aligned pointers) as a
 * plain streaming copy in 16-byte vectors over nblocks x 256 threads
 * (nblocks <= 0: 2048).  No counterpart in the reference: the measurement
 * of the achievable HBM rate that bench.py reports next to the kernels. */
int gh_stream_copy(const void* src_dev, void* dst_dev, int64_t nbytes, int nblocks,
                   void* stream);

/* --------------------------------------------- non-local Hamiltonian --
 * out[x, M_a + i] = sum_j P[x, M_a + j] dH_a[j][i] for every atom
 * (hamiltonian.py:23-44, apply_non_local_hamilton).  dH_dev holds the
 * atoms' nm_a x nm_a real matrices concatenated (row-major); P and out have
 * the LFC's dtype (single-precision P is summed in FP64 and rounded once).
 * If eps_dev (real, nbands) and dS_dev are given, also subtracts
 * eps[x] * sum_j P[x, j] dS_a[j][i] (eigensolver.py:161-165). */
int gh_lfc_apply_dH(const gh_lfc* h, const void* P_dev, const double* dH_dev,
                    const double* eps_dev, const double* dS_dev,
                    void* out_dev, int64_t nbands, void* stream);

/* ----------------------------------------------------- preconditioner --
 * The whole 14-step multigrid V-cycle of reference
 * c/fd_preconditioner.c:108-353 / gpaw/preconditioner.py:35-93 in one call:
 * d0 = P(R).  kin0/kin1/kin2: stencils of the three levels; rest0/rest1:
 * restrictions gd0->gd1->gd2; intp2/intp1: interpolations gd2->gd1->gd0;
 * step = 0.66666666 / kin0 diagonal.  scratch_dev: caller-owned device
 * buffer of at least gh_precond_scratch_bytes() bytes.  R_dev and d0_dev
 * are (nbands, n0, n1, n2) of the stencils' dtype. */
int64_t gh_precond_scratch_bytes(const gh_stencil* kin0,
                                 const gh_stencil* kin1,
                                 const gh_stencil* kin2, int64_t nbands);
int gh_precond_apply(const gh_stencil* kin0, const gh_stencil* kin1,
                     const gh_stencil* kin2, const gh_transformer* rest0,
                     const gh_transformer* rest1,
                     const gh_transformer* intp2,
                     const gh_transformer* intp1, double step,
                     const void* R_dev, void* d0_dev, int64_t nbands,
                     const double* phase_cd, void* scratch_dev, void* stream);

/* ------------------------------------------------- band-wise BLAS-1 --
 * multi_axpy (gpaw/utilities/mblas.py:13-32): y[x] += a[x] * x[x] with a
 * per-band real coefficient array a_dev (nbands doubles).  Data dtype any
 * of GH_*; n = points per band. */
int gh_multi_axpy(int dtype, int64_t nbands, int64_t n, const double* a_dev,
                  const void* x_dev, void* y_dev, void* stream);

/* out[x] = Re sum_G conj(a[x,G]) b[x,G] * scale  (per band, FP64 result
 * and FP64 accumulation for every input precision, fixed reduction order;
 * the band-wise integrals of rmmdiis.py:124-130,158-164).  b_dev may equal
 * a_dev.  work_dev: caller-owned device scratch of GH_DOT_CHUNKS * nbands
 * doubles. */
#define GH_DOT_CHUNKS 64
int gh_multi_dotc(int dtype, int64_t nbands, int64_t n, const void* a_dev,
                  const void* b_dev, double scale, double* out_dev,
                  double* work_dev, void* stream);

/* Per-band Gram matrices of k vectors (the DIIS residual matrix R_nn of
 * rmmdiis.py:207-212, wfs.integrate(R_i, R_j) = dv <R_i|R_j>):
 *   out[x][i][j] = scale * sum_G conj(v_i[x, G]) v_j[x, G],  1 <= k <= 6,
 * for x < nbands, each v_i a device array (nbands, n) of `dtype`.
 * vecs_dev: host array of k device pointers.  out_dev: device
 * (nbands, k, k) doubles for real dtypes, (nbands, k, k, 2) (re, im) for
 * complex ones (Hermitian).  work_dev: gh_multi_gram_work(k, nbands)
 * doubles.  FP64 accumulation in a fixed order. */
int64_t gh_multi_gram_work(int k, int64_t nbands);
int gh_multi_gram(int dtype, int64_t nbands, int64_t n, int k,
                  const void* const* vecs_dev, double scale, double* out_dev,
                  double* work_dev, void* stream);

/* gh_multi_gram without a workspace argument and without any allocation:
 * no counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The partial sums go to a fixed 16 MiB device array
 * per GPU (held until the process ends), so gh_multi_gram_work(k, nbands)
 * doubles must fit in 16 MiB (e.g. k = 6: up to 780 bands; k = 2: 5461);
 * larger calls return GH_ERR_ARG.  Every call on one GPU shares that array:
 * calls that overlap in time on one GPU (several streams or threads) race.
 * Results are identical to gh_multi_gram's. */
int gh_multi_gram_noalloc(int dtype, int64_t nbands, int64_t n, int k,
                          const void* const* vecs_dev, double scale,
                          double* out_dev, void* stream);

/* Per-band linear combination of k vectors, one pass over the data:
 *   out[x, G] = sum_{i<k} c[x, i] v_i[x, G],  1 <= k <= 8,
 * accumulated in the order given.  Replaces the DIIS update of
 * rmmdiis.py:226-237 (psitb = alpha_nit psi_nit, then axpy of the older
 * iterates) and the out-of-place form of the per-band axpy of
 * mblas.py:13-32 (v_0 + lam v_1).  vecs_dev: host array of k device
 * pointers (nbands, n) of `dtype`; coef_dev: device (nbands, k) doubles for
 * real dtypes, (nbands, k, 2) (re, im) for complex ones.  out_dev may be
 * one of the inputs. */
int gh_multi_lincomb(int dtype, int64_t nbands, int64_t n, int k,
                     const void* const* vecs_dev, const double* coef_dev,
                     void* out_dev, void* stream);

/* Pseudo-density accumulation, replaces the per-band calls of
 * cgpaw.add_to_density (c/utilities.c:313-356; wavefunctions/fd.py:95-100):
 *   nt[G] += sum_{x < nbands} f[x] |psit[x, G]|^2   (FP64 nt, any dtype)
 * f_dev: nbands doubles on the device; bands are added in order. */
int gh_add_to_density(int dtype, int64_t nbands, int64_t n, const double* f_dev,
                      const void* psit_dev, double* nt_dev, void* stream);

/* ------------------------------------------------- band overlap matrices --
 * Replaces the rk / r2k / mmm(..., 'N', ..., 'C') contractions of
 * gpaw/utilities/blas.py:39,273,314 (c/blas.c:48,101,140) as used for the
 * subspace matrices (gpaw/matrix.py:229-327, wavefunctions/arrays.py:262-340,
 * GridDescriptor.integrate grid_descriptor.py:275-326):
 *   C[i][j] = alpha * sum_g A[i][g] * conj(B[j][g])   (conj_out: conj of it)
 * A_dev: m rows, B_dev: n rows, each of k elements of dtype with row pitch
 * lda / ldb elements (k * element size, lda and ldb must be multiples of 16
 * bytes; A and B 16-byte aligned).  C_dev: m x n row-major float64 (real
 * dtypes) or complex128 (complex dtypes).  hermitian = 1 (m == n): only the
 * lower triangle is formed from the data and the upper one is its conjugate
 * (for products known to be Hermitian, like rk's lower triangle).
 * FP32/complex64 on v_mfma_f32_32x32x2f32 and FP64/complex128 on
 * v_mfma_f64_16x16x4f64; products are summed in the element precision within
 * k-chunks and the chunks in FP64 (fixed order: deterministic).
 * work_dev: device scratch of gh_gram_work_bytes(...) bytes. */
int64_t gh_gram_work_bytes(int dtype, int64_t m, int64_t n, int64_t k, int hermitian);
int gh_gram(int dtype, int64_t m, int64_t n, int64_t k, const void* A_dev,
            int64_t lda, const void* B_dev, int64_t ldb, double alpha,
            int hermitian, int conj_out, double* C_dev, void* work_dev,
            int64_t work_bytes, void* stream);

/* ---------------------------------------------- learned preconditioner --
 * Layers of the 3D U-Net that replaces the multigrid V-cycle
 * (gpaw/preconditioner.py:9-93; used by rmmdiis.py:138-143): d = f(R, v).
 * Activations: BF16 bits, channel-last (nbands, n0, n1, n2, C); periodic
 * grids.  Convolution weights: BF16, packed by gh_unet_pack_conv3 from
 * torch's conv3d layout W[co][ci][3][3][3] (gh_unet_conv3_weight_elems
 * uint16 values); sums in FP32 on the matrix cores.
 * gh_unet_conv3 mode 0 (first layer, cin = 2): channel 0 = scale[b] R[b],
 * channel 1 = vscale v (R, v: dtype GH_F32 / GH_F64, band-major grids),
 * staged as 4 channels (the last two zero): its weights are packed with
 * gh_unet_pack_conv3(4, cout, W) from W[co][4][3][3][3];
 * mode 1: act_in; mode 2: nearest 2x upsampling of act_in (the grid halved)
 * plus skip.  The input slab of a block (3 x-planes, its rows plus halo)
 * is staged in LDS once.  relu: apply max(0, .) after the bias.  Layers built:
 * 2->8 (mode 0), 8->16, 16->32, 32->32, 8->8, 16->16 (mode 1), 16->16,
 * 8->8 (mode 2). */
int64_t gh_unet_conv3_weight_elems(int cin, int cout);
int gh_unet_pack_conv3(int cin, int cout, const float* W, uint16_t* packed);
int gh_unet_conv3(int cin, int cout, int mode, int dtype, int nbands,
                  const int64_t* n_c, const void* in_dev, const void* skip_dev,
                  const void* R_dev, const void* v_dev, const float* scale_dev,
                  double vscale, const void* W_dev, const float* bias_dev,
                  int relu, void* out_dev, void* stream);
/* 2x2x2 average pooling (fine grid n_c, even), c = 8, 16 or 32. */
int gh_unet_pool2(int c, int nbands, const int64_t* n_c, const void* in_dev,
                  void* out_dev, void* stream);
/* 1x1 channel map with FP32 weights W[ci][co] (32->16, 16->8). */
int gh_unet_map1(int cin, int cout, int64_t npts, const void* in_dev,
                 const float* W_dev, const float* bias_dev, void* out_dev,
                 void* stream);
/* out[b][p] = (sum_c W[c] act[b][p][c] + bias) / scale[b] (cin = 8; out of
 * dtype GH_F32 or GH_F64). */
int gh_unet_head(int cin, int dtype, int nbands, int64_t npts_per_band,
                 const void* in_dev, const float* W_dev, double bias,
                 const float* scale_dev, void* out_dev, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* GPAW_AMD_H */
