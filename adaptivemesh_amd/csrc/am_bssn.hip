// BSSN right-hand side, fused RK4 stage update and constraint norms for
// gfx950.  The reference stops at metric derivatives and Christoffel symbols
// (BSSN.h:7-102); the equations here are the moving-puncture BSSN system
// (chi form, 1+log or harmonic slicing, Gamma-driver shift with B^i), the
// same system the FP64 oracle states term by term.
//
// Inputs per interior cell: 24 fp32 state values (stage input) and the
// compressed_float3 words written by the FD kernel (a gradient for every
// field, a full Hessian for chi, gt_ij, alpha, beta^i).  The lowered
// symbols Gt_ijk come from the metric gradient words and are raised in
// registers (Gt^i_jk = gt^il Gt_ljk): no Christoffel words are read, so the
// symbols carry one quantisation (that of the gradient words), not two.
//
// One thread per interior cell, 256 cells per workgroup, every array read
// as contiguous runs (words: 1 KB per workgroup per array; state: 16-float
// row segments).  This kernel is HBM-bound (~500-700 B per cell per stage
// against ~1.5 kflop): bulk arithmetic is fp32 with FMA contraction, the
// constraint sums are reduced in FP64.
#include "am_bssn_kernels.hpp"

namespace am {

template <int S, int P>
static int launch_stage(const am_bssn_fields *f, int stage, const am_bssn_params *p, const am_mesh *mesh,
                        const am_geometry *g, int push, void *stream)
{
    if (f->word_format == WORD_CF3) return launch_stage_fmt<S, P, WORD_CF3>(f, stage, p, mesh, g, push, stream);
    if (S == 16 && P == 3) return bssn_stage_words(f->word_format, f, stage, p, mesh, g, push, stream);
    return set_error(AM_ERR_INVALID, "am_bssn_rhs_stage: word format %d needs S=16, P=3", f->word_format);
}

template <int S, int P>
static int launch_cons(const am_bssn_fields *f, const am_bssn_params *p, const am_mesh *mesh, double *partial,
                       void *stream)
{
    if (f->word_format == WORD_CF3) return launch_cons_fmt<S, P, WORD_CF3>(f, p, mesh, partial, stream);
    if (S == 16 && P == 3) return bssn_cons_words(f->word_format, f, p, mesh, partial, stream);
    return set_error(AM_ERR_INVALID, "am_bssn_constraints: word format %d needs S=16, P=3", f->word_format);
}

}  // namespace am

extern "C" int am_bssn_rhs_stage(const am_bssn_fields *f, int32_t stage, const am_bssn_params *p, const am_mesh *mesh,
                                 const am_geometry *g, int32_t push_ghosts, void *stream)
{
    if (!f || !p || !mesh || !g || !mesh->tiles || !mesh->positions)
        return am::set_error(AM_ERR_INVALID, "am_bssn_rhs_stage: null argument");
    if (mesh->n_tiles <= 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_stage, f, stage, p, mesh, g, push_ghosts, stream);
}

extern "C" int am_bssn_constraints(const am_bssn_fields *f, const am_bssn_params *p, const am_mesh *mesh,
                                   const am_geometry *g, double *partial, void *stream)
{
    if (!f || !p || !mesh || !g || !mesh->tiles || !partial)
        return am::set_error(AM_ERR_INVALID, "am_bssn_constraints: null argument");
    if (mesh->n_tiles <= 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_cons, f, p, mesh, partial, stream);
}

extern "C" int64_t am_bssn_constraint_blocks(const am_geometry *g, int32_t n_tiles)
{
    if (!g) return -1;
    const int64_t S = g->size_domain;
    return (int64_t)n_tiles * ((S * S * S + 255) / 256);
}
