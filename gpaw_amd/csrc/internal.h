// Library-internal entry points shared between translation units.
#pragma once

#include "common.h"

namespace gh {

// out = alpha * (L in) + gamma * in [+ vt .* in] [+ beta * src]
int stencil_apply_scaled(const gh_stencil* h, const void* in, void* out,
                         int64_t nbands, const double* phase_cd,
                         const void* vt, const void* src, double beta,
                         double alpha, double gamma, hipStream_t st);

// as stencil_apply_scaled, with the centre coefficient alpha c0 + gamma -
// eps[band] per band (eps: nbands device doubles or null); orthogonal
// stencils only
int stencil_apply_full(const gh_stencil* h, const void* in, void* out,
                       int64_t nbands, const double* phase_cd,
                       const void* vt, const void* src, double beta,
                       double alpha, double gamma, const double* eps,
                       hipStream_t st);

// out = alpha * T(in)
int transformer_apply_scaled(const gh_transformer* h, const void* in,
                             void* out, int64_t nbands,
                             const double* phase_cd, double alpha,
                             hipStream_t st);

int64_t stencil_points(const gh_stencil* h);
int stencil_dtype(const gh_stencil* h);
int64_t transformer_in_points(const gh_transformer* h);
int64_t transformer_out_points(const gh_transformer* h);

}  // namespace gh
