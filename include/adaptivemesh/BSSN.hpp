// BSSN plugins of the drop-in API.
//
// BSSN_AMR_data mirrors the reference plugin (BSSN.h:104-171): the six
// conformal metric components with their derivative words, the metric trace
// of each component's Hessian and the Christoffel words, recomputed in
// predictor and corrector with the reference's 26 seed draws per substep.
// Like the reference it never advances the metric.
//
// BSSN_rk4_data is the evolution the reference leaves out: the 24 fields of
// the moving-puncture BSSN system advanced by the generic RK4 plugin
// (RK4.hpp) with sub-cycling.
#pragma once
#include <math.h>

#include <vector>

#include "AMR.hpp"
#include "RK4.hpp"

namespace am {

// --------------------------------------------------------------- reference
struct BSSN_AMR_data {
    fast_prng rng;
    AMR<BSSN_AMR_data> &parent;
    am_geometry geom;
    void *stream = nullptr;
    bool christoffel_compat = true;  // reproduce BSSN.h:82-83 (metric instead of inverse)

    // the reference's members and component names (BSSN.h:110-119)
    tensor2_sym_field cyij_old, cyij_new;  // conformal metric, padded tiles
    comp_tensor3_sym_field cyij_k;         // its gradient words
    comp_tensor3_sym_field cGijk;          // Christoffel words
    tensor2_sym_field cRij;                // gamma^ij d_i d_j gamma_kl, interior-only
    comp_tensor2_sym_field didjA, didjW;   // Hessian scratch (diagonal, off-diagonal words)

    static size_t tile_cells(const am_geometry &g)
    {
        const size_t T = (size_t)(g.size_domain + 2 * g.padding_domain);
        return T * T * T;
    }
    static size_t inner_cells(const am_geometry &g)
    {
        const size_t S = (size_t)g.size_domain;
        return S * S * S;
    }
    BSSN_AMR_data(uint32_t node_slots, AMR<BSSN_AMR_data> &p, const am_geometry &g)
        : rng(), parent(p), geom(g), cyij_old(node_slots, tile_cells(g)), cyij_new(node_slots, tile_cells(g)),
          cyij_k(node_slots, inner_cells(g)), cGijk(node_slots, inner_cells(g)), cRij(node_slots, inner_cells(g)),
          didjA(node_slots, inner_cells(g)), didjW(node_slots, inner_cells(g))
    {
        // zeroed, so a run does not depend on what the allocation held (the
        // reference leaves its buffers uninitialised)
        for (int c = 0; c < 6; c++) {
            cyij_old[c].zero(); cyij_new[c].zero(); cyij_k[c].zero(); cGijk[c].zero(); cRij[c].zero();
        }
        for (int c = 0; c < 2; c++) { didjA[c].zero(); didjW[c].zero(); }
        check(am_stream_sync(stream), "BSSN_AMR_data");
    }
    void destroy()
    {
        cyij_old.destroy(); cyij_new.destroy(); cyij_k.destroy(); cGijk.destroy(); cRij.destroy();
        didjA.destroy(); didjW.destroy();
    }
    void node_ctor(int node_idx)  // BSSN.h:131-137
    {
        const uint32_t off = (uint32_t)parent.positions.cpu_buffer_ptr[node_idx].final_offset();
        const int par = parent.parent_idx_b.cpu_buffer_ptr[node_idx];
        for (int c = 0; c < 6; c++) check(copy_to_child(cyij_old[c], off, node_idx, par, geom, stream), "node_ctor");
        for (int c = 0; c < 6; c++) check(copy_to_child(cyij_new[c], off, node_idx, par, geom, stream), "node_ctor");
    }
    void node_dtor(int) {}
    void modify_nodes() {}
    void copy_back()  // BSSN.h:126-130
    {
        for (int c = 0; c < 6; c++) check(copy_new_to_old(cyij_old[c], cyij_new[c], parent, geom, stream), "copy_back");
    }
    // metric_derivatives_noncovariant + compute_raised_christoffel_symbols
    // (BSSN.h:38-53, 93-102): per component differentiate (1 draw), mixed
    // partials (1 draw), metric trace; then the Christoffel words (1 draw).
    void derivatives(tensor2_sym_field &metric)
    {
        const am_mesh m = parent.mesh_all();
        const float *met[6];
        for (int c = 0; c < 6; c++) met[c] = metric[c].gpu_buffer_ptr;
        for (int c = 0; c < 6; c++) {
            auto &h = (c & 1) ? didjW : didjA;
            check(differentiate(metric[c], cyij_k[c], h[0], parent, geom, stream, rng), "derivatives");
            check(mixed_partials(metric[c], h[1], parent, geom, stream, rng), "derivatives");
            check(am_deriv_trace(words(h[0]), words(h[1]), met, cRij[c].gpu_buffer_ptr, &m, &geom, stream),
                  "derivatives");
        }
        const uint32_t *dg[6];
        uint32_t *out[6];
        for (int c = 0; c < 6; c++) {
            dg[c] = words(cyij_k[c]);
            out[c] = words(cGijk[c]);
        }
        check(am_christoffel(met, dg, out, christoffel_compat ? 1 : 0, &m, &geom, rng.generate_int(), stream),
              "christoffel");
    }
    void predictor() { derivatives(cyij_old); }
    void copy_bounds()
    {
        for (int c = 0; c < 6; c++) check(copy_boundaries(cyij_old[c], cyij_new[c], parent, geom, stream), "copy_bounds");
    }
    void corrector() { derivatives(cyij_new); }
    void timestep()  // BSSN.h:160-170
    {
        do {
            copy_back();
            modify_nodes();
            predictor();
            copy_bounds();
            corrector();
            parent.increment_timer();
        } while (parent.timer_helper != 0u);
        check(am_stream_sync(stream), "timestep");
    }
};

// -------------------------------------------------------------- evolution
// The 24-field moving-puncture system for rk4_AMR_data.
struct bssn_system {
    static constexpr int NV = AM_BSSN_NVARS;
    // chi, gt_ij, alpha, beta^i need full Hessians; everything else a gradient
    static constexpr int FD_KIND[NV] = {3, 3, 3, 3, 3, 3, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 1, 3, 3, 3, 3, 1, 1, 1};
    static constexpr bool CHRISTOFFEL = false;  // raised in registers by the RHS kernel
    static constexpr bool WORD_FORMATS = true;  // fp32 / fp16 words for S=16, P=3 (precision study)
    static constexpr uint32_t CKPT_KIND = AM_CKPT_BSSN;
    // upwinded advection: beta^i is fields 18-20; fields 0-17 are always
    // advected, beta and B with advect_shift
    static constexpr int SHIFT_FIELD = 18;
    using options_t = am_bssn_params;
    static uint32_t advect_mask(const options_t &p) { return p.advect_shift ? 0xFFFFFFu : 0x3FFFFu; }

    static am_bssn_fields fields(const rk4_views &v)
    {
        am_bssn_fields f{};
        f.tile_layout = v.tile_layout;
        if (v.adv[0])
            for (int i = 0; i < NV; i++) f.adv[i] = v.adv[i];
        f.word_format = v.word_format;
        f.push_layers = v.push_layers;
        f.push_axes = v.push_axes;
        for (int i = 0; i < NV; i++) {
            f.state[i] = v.state[i];
            f.old_[i] = v.old_[i];
            f.new_[i] = v.new_[i];
            f.acc[i] = v.acc[i];
            f.d1[i] = v.d1[i];
        }
        for (int h = 0; h < AM_BSSN_NHESS; h++) {
            f.d2diag[h] = v.d2diag[h];
            f.d2off[h] = v.d2off[h];
        }
        return f;
    }
    static int rhs_stage(const rk4_views &v, int stage, const options_t &p, const am_mesh &m, const am_geometry &g,
                         int push, void *stream)
    {
        const am_bssn_fields f = fields(v);
        return am_bssn_rhs_stage(&f, stage, &p, &m, &g, push, stream);
    }
    // {rms H, rms |M|, max |H|, cells} over the leaves
    static void diagnostics(const rk4_views &v, const options_t &p, const am_mesh &m, const am_geometry &g,
                            double *partial, double *out4, void *stream)
    {
        const am_bssn_fields f = fields(v);
        check(am_bssn_constraints(&f, &p, &m, &g, partial, stream), "constraints");
        const int64_t nb = am_bssn_constraint_blocks(&g, m.n_tiles);
        std::vector<double> h((size_t)(4 * nb));
        check(am_memcpy_d2h(h.data(), partial, h.size() * sizeof(double), stream), "constraints");
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        for (int64_t i = 0; i < nb; i++) {
            s0 += h[4 * i]; s1 += h[4 * i + 1]; s2 += h[4 * i + 2]; s3 = fmax(s3, h[4 * i + 3]);
        }
        out4[0] = s2 > 0 ? sqrt(s0 / s2) : 0;
        out4[1] = s2 > 0 ? sqrt(s1 / s2) : 0;
        out4[2] = s3;
        out4[3] = s2;
    }
};

using BSSN_rk4_data = rk4_AMR_data<bssn_system>;

}  // namespace am
