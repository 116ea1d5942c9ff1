// The scalar wave equation u_tt = lap u as a System for rk4_AMR_data: fields
// u (gradient + diagonal Hessian words) and v = u_t (no derivatives).  The
// reference's own wave plugin (SimpleWave.h) is a Heun step of a Burgers-type
// equation and stays available as wave_AMR_data; this is the RK4 evolution
// of the linear wave on the same octree.
#pragma once
#include <math.h>

#include <vector>

#include "AMR.hpp"
#include "RK4.hpp"

namespace am {

struct swave_options {
    int32_t unused = 0;
};

struct swave_system {
    static constexpr int NV = 2;
    static constexpr int FD_KIND[NV] = {2, 0};
    static constexpr bool CHRISTOFFEL = false;
    static constexpr bool WORD_FORMATS = false;  // compressed_float3 words only
    static constexpr uint32_t CKPT_KIND = AM_CKPT_SWAVE;
    static constexpr int SHIFT_FIELD = -1;  // no shift: no upwinded advection
    using options_t = swave_options;
    static uint32_t advect_mask(const options_t &) { return 0u; }

    static am_swave_fields fields(const rk4_views &v)
    {
        am_swave_fields f{};
        f.tile_layout = v.tile_layout;
        f.push_layers = v.push_layers;
        f.push_axes = v.push_axes;
        for (int i = 0; i < NV; i++) {
            f.state[i] = v.state[i];
            f.old_[i] = v.old_[i];
            f.new_[i] = v.new_[i];
            f.acc[i] = v.acc[i];
        }
        f.d1_u = v.d1[0];
        f.d2diag_u = v.d2diag[0];
        return f;
    }
    static int rhs_stage(const rk4_views &v, int stage, const options_t &, const am_mesh &m, const am_geometry &g,
                         int push, void *stream)
    {
        const am_swave_fields f = fields(v);
        return am_swave_rhs_stage(&f, stage, &m, &g, push, stream);
    }
    // {energy, integral of u^2, max |u|, cells} over the leaves
    static void diagnostics(const rk4_views &v, const options_t &, const am_mesh &m, const am_geometry &g,
                            double *partial, double *out4, void *stream)
    {
        const am_swave_fields f = fields(v);
        check(am_swave_energy(&f, &m, &g, partial, stream), "energy");
        const int64_t nb = am_bssn_constraint_blocks(&g, m.n_tiles);
        std::vector<double> h((size_t)(4 * nb));
        check(am_memcpy_d2h(h.data(), partial, h.size() * sizeof(double), stream), "energy");
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        for (int64_t i = 0; i < nb; i++) {
            s0 += h[4 * i]; s1 += h[4 * i + 1]; s2 += h[4 * i + 2]; s3 = fmax(s3, h[4 * i + 3]);
        }
        out4[0] = s0;
        out4[1] = s1;
        out4[2] = s3;
        out4[3] = s2;
    }
};

using swave_rk4_data = rk4_AMR_data<swave_system>;

}  // namespace am
