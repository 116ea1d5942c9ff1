// fp32 and fp16 derivative-word instantiations of the BSSN kernels (the
// precision study), for S = 16, P = 3 tiles only.
#include "am_bssn_kernels.hpp"

namespace am {

int bssn_stage_words(int fmt, const am_bssn_fields *f, int stage, const am_bssn_params *p, const am_mesh *mesh,
                     const am_geometry *g, int push, void *stream)
{
    if (fmt == WORD_F32) return launch_stage_fmt<16, 3, WORD_F32>(f, stage, p, mesh, g, push, stream);
    if (fmt == WORD_F16) return launch_stage_fmt<16, 3, WORD_F16>(f, stage, p, mesh, g, push, stream);
    return set_error(AM_ERR_INVALID, "am_bssn_rhs_stage: word format %d", fmt);
}

int bssn_cons_words(int fmt, const am_bssn_fields *f, const am_bssn_params *p, const am_mesh *mesh,
                    double *partial, void *stream)
{
    if (fmt == WORD_F32) return launch_cons_fmt<16, 3, WORD_F32>(f, p, mesh, partial, stream);
    if (fmt == WORD_F16) return launch_cons_fmt<16, 3, WORD_F16>(f, p, mesh, partial, stream);
    return set_error(AM_ERR_INVALID, "am_bssn_constraints: word format %d", fmt);
}

}  // namespace am
