// BSSN plugins of the drop-in API.
//
// BSSN_AMR_data mirrors the reference plugin (BSSN.h:104-171): the six
// conformal metric components with their derivative words, the metric trace
// of each component's Hessian and the Christoffel words, recomputed in
// predictor and corrector with the reference's 26 seed draws per substep.
// Like the reference it never advances the metric.
//
// BSSN_rk4_data is the evolution the reference leaves out: 24 fields of the
// moving-puncture BSSN system, classical RK4 with Berger-Oliger sub-cycling
// (each depth steps 2^(max_depth - d) times less often, AMR.h:640-644), run
// depth by depth, coarse first, so that a fine tile's ghost cells always
// interpolate a coarse step that is complete.  Per stage and depth:
//   FD words for all 24 fields (one batched launch) -> Christoffel words ->
//   RHS + RK4 update -> ghost fill at the stage time.
#pragma once
#include <math.h>

#include <vector>

#include "AMR.hpp"

namespace am {

// --------------------------------------------------------------- reference
struct BSSN_AMR_data {
    fast_prng rng;
    AMR<BSSN_AMR_data> &parent;
    am_geometry geom;
    void *stream = nullptr;
    bool christoffel_compat = true;  // reproduce BSSN.h:82-83 (metric instead of inverse)

    smart_gpu_buffer<float> cyij_old[6], cyij_new[6];     // xx,xy,xz,yy,yz,zz (BSSN.h:111-112)
    smart_gpu_buffer<compressed_float3> cyij_k[6];        // gradient words (BSSN.h:115)
    smart_gpu_buffer<compressed_float3> cGijk[6];         // Christoffel words (BSSN.h:116)
    smart_gpu_buffer<float> cRij[6];                      // gamma^ij d_i d_j gamma_kl (BSSN.h:117)
    smart_gpu_buffer<compressed_float3> didjA[2], didjW[2];  // Hessian scratch (diag, off) (BSSN.h:118-119)

    BSSN_AMR_data(uint32_t node_slots, AMR<BSSN_AMR_data> &p, const am_geometry &g) : rng(), parent(p), geom(g)
    {
        const size_t T = (size_t)(g.size_domain + 2 * g.padding_domain), S = (size_t)g.size_domain;
        for (int c = 0; c < 6; c++) {
            cyij_old[c] = smart_gpu_buffer<float>(node_slots * T * T * T);
            cyij_new[c] = smart_gpu_buffer<float>(node_slots * T * T * T);
            cyij_k[c] = smart_gpu_buffer<compressed_float3>(node_slots * S * S * S);
            cGijk[c] = smart_gpu_buffer<compressed_float3>(node_slots * S * S * S);
            cRij[c] = smart_gpu_buffer<float>(node_slots * S * S * S);
        }
        for (int c = 0; c < 2; c++) {
            didjA[c] = smart_gpu_buffer<compressed_float3>(node_slots * S * S * S);
            didjW[c] = smart_gpu_buffer<compressed_float3>(node_slots * S * S * S);
        }
    }
    void destroy()
    {
        for (int c = 0; c < 6; c++) {
            cyij_old[c].destroy(); cyij_new[c].destroy(); cyij_k[c].destroy(); cGijk[c].destroy(); cRij[c].destroy();
        }
        for (int c = 0; c < 2; c++) { didjA[c].destroy(); didjW[c].destroy(); }
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
    void derivatives(smart_gpu_buffer<float> (&metric)[6])
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
struct bssn_rk4_options {
    am_bssn_params params{2.0f, 0.75f, 1e-4f, 1, 0};
    int32_t interp = AM_INTERP_LAGRANGE;  // ghost fill: reference quadratic, or tricubic + quadratic time
    int32_t z_block = 0;                  // FD launch shape (0 = whole tile per workgroup)
};

// Optional per-launch timing with HIP events on the launch stream.
struct kernel_timer {
    enum { FD = 0, CHRISTOFFEL, RHS, GHOST, COPY_BACK, N_CLASSES };
    bool on = false;
    std::vector<void *> pool;
    std::vector<std::pair<int, int>> marks;  // (class, index of the start event)
    size_t next = 0;
    void *stream = nullptr;

    void enable(bool e, void *st)
    {
        stream = st;
        on = e;
        if (e && pool.empty()) {
            pool.resize(16384);
            for (auto &ev : pool) check(am_event_create(&ev), "am_event_create");
        }
        marks.clear();
        next = 0;
    }
    int begin(int cls)
    {
        if (!on || next + 2 > pool.size()) return -1;
        const int i = (int)next;
        next += 2;
        check(am_event_record(pool[i], stream), "am_event_record");
        marks.emplace_back(cls, i);
        return i;
    }
    void end(int i)
    {
        if (i >= 0) check(am_event_record(pool[i + 1], stream), "am_event_record");
    }
    // total ms and launch count per class since enable(); resets the marks
    void totals(double *ms, int64_t *count)
    {
        for (int c = 0; c < N_CLASSES; c++) { ms[c] = 0; count[c] = 0; }
        for (auto &m : marks) {
            float t = 0.f;
            check(am_event_elapsed_ms(pool[m.second], pool[m.second + 1], &t), "am_event_elapsed_ms");
            ms[m.first] += t;
            count[m.first]++;
        }
        marks.clear();
        next = 0;
    }
    void destroy()
    {
        for (auto ev : pool) am_event_destroy(ev);
        pool.clear();
    }
};

struct BSSN_rk4_data {
    static constexpr int NV = AM_BSSN_NVARS, NH = AM_BSSN_NHESS;
    static constexpr int HESS_VAR[NH] = {0, 1, 2, 3, 4, 5, 6, 17, 18, 19, 20};

    fast_prng rng;
    AMR<BSSN_rk4_data> &parent;
    am_geometry geom;
    bssn_rk4_options opt;
    void *stream = nullptr;
    size_t cells = 0, inner = 0, slots = 0;

    // one allocation per kind; variable v at base + v * slots * cells
    smart_gpu_buffer<float> old_, new_, acc, older;
    smart_gpu_buffer<uint32_t> d1, d2d, d2o, chr;
    smart_gpu_buffer<double> partial;
    uint64_t zone_cycles = 0;  // interior cells advanced by one full step, all depths
    kernel_timer timer;

    BSSN_rk4_data(uint32_t node_slots, AMR<BSSN_rk4_data> &p, const am_geometry &g,
                  const bssn_rk4_options &o = bssn_rk4_options())
        : rng(), parent(p), geom(g), opt(o)
    {
        const size_t T = (size_t)(g.size_domain + 2 * g.padding_domain), S = (size_t)g.size_domain;
        cells = T * T * T;
        inner = S * S * S;
        slots = node_slots;
        old_ = smart_gpu_buffer<float>(NV * slots * cells);
        new_ = smart_gpu_buffer<float>(NV * slots * cells);
        acc = smart_gpu_buffer<float>(NV * slots * cells);
        if (opt.interp == AM_INTERP_TRICUBIC) older = smart_gpu_buffer<float>(NV * slots * cells);
        d1 = smart_gpu_buffer<uint32_t>(NV * slots * inner);
        d2d = smart_gpu_buffer<uint32_t>(NH * slots * inner);
        d2o = smart_gpu_buffer<uint32_t>(NH * slots * inner);
        chr = smart_gpu_buffer<uint32_t>(6 * slots * inner);
        partial = smart_gpu_buffer<double>(4 * (size_t)am_bssn_constraint_blocks(&geom, (int32_t)slots) + 4);
        old_.zero(); new_.zero(); acc.zero();
        if (older.created) older.zero();
    }
    void destroy()
    {
        old_.destroy(); new_.destroy(); acc.destroy(); older.destroy();
        d1.destroy(); d2d.destroy(); d2o.destroy(); chr.destroy(); partial.destroy();
        timer.destroy();
    }
    float *field(smart_gpu_buffer<float> &b, int v) { return b.gpu_buffer_ptr + (size_t)v * slots * cells; }
    uint32_t *word(smart_gpu_buffer<uint32_t> &b, int v) { return b.gpu_buffer_ptr + (size_t)v * slots * inner; }

    am_field_batch batch(bool with_older)
    {
        am_field_batch b{};
        b.n_vars = NV;
        for (int v = 0; v < NV; v++) {
            b.old_[v] = field(old_, v);
            b.new_[v] = field(new_, v);
            b.older[v] = with_older && older.created ? field(older, v) : nullptr;
        }
        return b;
    }
    am_bssn_fields fields(bool stage_input_is_old)
    {
        am_bssn_fields f{};
        for (int v = 0; v < NV; v++) {
            f.state[v] = stage_input_is_old ? field(old_, v) : field(new_, v);
            f.old_[v] = field(old_, v);
            f.new_[v] = field(new_, v);
            f.acc[v] = field(acc, v);
            f.d1[v] = word(d1, v);
        }
        for (int h = 0; h < NH; h++) {
            f.d2diag[h] = word(d2d, h);
            f.d2off[h] = word(d2o, h);
        }
        for (int c = 0; c < 6; c++) f.chr[c] = word(chr, c);
        return f;
    }

    void node_ctor(int node_idx)
    {
        const uint32_t off = (uint32_t)parent.positions.cpu_buffer_ptr[node_idx].final_offset();
        const int par = parent.parent_idx_b.cpu_buffer_ptr[node_idx];
        for (int v = 0; v < NV; v++) {
            check(am_copy_to_child(field(old_, v), off, node_idx, par, &geom, AM_INTERP_LAGRANGE, stream), "node_ctor");
            check(am_copy_to_child(field(new_, v), off, node_idx, par, &geom, AM_INTERP_LAGRANGE, stream), "node_ctor");
        }
    }
    void node_dtor(int) {}
    void modify_nodes() {}

    // derivative words of the state held in `src` for the tiles of `m`
    void derivatives(smart_gpu_buffer<float> &src, const am_mesh &m)
    {
        am_fd_batch b{};
        b.n_vars = NV;
        b.z_block = opt.z_block;
        int h = 0;
        for (int v = 0; v < NV; v++) {
            b.field[v] = field(src, v);
            b.d1[v] = word(d1, v);
            b.seed_d[v] = rng.generate_int();
            if (h < NH && HESS_VAR[h] == v) {
                b.d2diag[v] = word(d2d, h);
                b.d2off[v] = word(d2o, h);
                b.seed_m[v] = rng.generate_int();
                h++;
            }
        }
        int ev = timer.begin(kernel_timer::FD);
        check(am_fd_batched(&b, &m, &geom, stream), "fd");
        timer.end(ev);
        const float *met[6];
        const uint32_t *dg[6];
        uint32_t *out[6];
        for (int c = 0; c < 6; c++) {
            met[c] = field(src, 1 + c);
            dg[c] = word(d1, 1 + c);
            out[c] = word(chr, c);
        }
        ev = timer.begin(kernel_timer::CHRISTOFFEL);
        check(am_christoffel(met, dg, out, 0, &m, &geom, rng.generate_int(), stream), "christoffel");
        timer.end(ev);
    }

    // copy-back: old <- new (+ restriction), keeping the previous old when a
    // third time level is in use (AMR.h:998-1051)
    void copy_back()
    {
        am_field_batch b = batch(true);
        const am_mesh m = parent.mesh_all();
        const int ev = timer.begin(kernel_timer::COPY_BACK);
        check(am_copy_new_to_old_batched(&b, &m, &geom, stream), "copy_back");
        timer.end(ev);
    }

    // one RK4 stage of every tile at depth d: derivative words of the stage
    // input, RHS + update, ghost fill of new at the stage time
    void stage_depth(int d, int s)
    {
        const am_mesh m = parent.mesh_depth(d);
        if (m.n_tiles == 0) return;
        static const float frac[4] = {.5f, .5f, 1.f, 1.f};
        {
            derivatives(s == 0 ? old_ : new_, m);
            const am_bssn_fields f = fields(s == 0);
            // equal-depth face ghosts are pushed by the RHS kernel; faces with a
            // coarser target are interpolated afterwards
            int ev = timer.begin(kernel_timer::RHS);
            check(am_bssn_rhs_stage(&f, s, &opt.params, &m, &geom, 1, stream), "rhs");
            timer.end(ev);
            if (d > 0 && parent.coarse_faces_at_depth(d) > 0) {
                am_field_batch b = batch(true);
                ev = timer.begin(kernel_timer::GHOST);
                check(am_ghost_fill_stage_batched(&b, &m, &geom, frac[s], opt.interp, AM_EQUAL_SKIP, stream),
                      "ghost");
                timer.end(ev);
            }
        }
    }
    // one RK4 step of every tile at depth d
    void step_depth(int d)
    {
        for (int s = 0; s < 4; s++) stage_depth(d, s);
        zone_cycles += (uint64_t)parent.tiles_at_depth(d) * inner;
    }

    void substep()
    {
        copy_back();
        for (int d = 0; d <= parent.read_max_depth(); d++)
            if (parent.depth_active(d)) step_depth(d);
        parent.increment_timer();
    }
    // the reference's timestep(): sub-steps until the counter wraps
    void timestep()
    {
        do {
            substep();
        } while (parent.timer_helper != 0u);
    }
    // fill every face ghost of `new` from the current data (initial data)
    void fill_ghosts()
    {
        for (int d = 1; d <= parent.read_max_depth(); d++) {
            const am_mesh m = parent.mesh_depth(d);
            if (m.n_tiles == 0) continue;
            am_field_batch b = batch(false);
            check(am_ghost_fill_stage_batched(&b, &m, &geom, 1.f, opt.interp, AM_EQUAL_COPY, stream), "ghost");
        }
    }
    // FP64 L2 norms of the Hamiltonian and momentum constraints of `new`
    // over the leaves: {rms H, rms |M|, max |H|, cells}
    void constraints(double *out4)
    {
        const am_mesh m = parent.mesh_leaves();
        derivatives(new_, m);
        const am_bssn_fields f = fields(false);
        check(am_bssn_constraints(&f, &opt.params, &m, &geom, partial.gpu_buffer_ptr, stream), "constraints");
        const int64_t nb = am_bssn_constraint_blocks(&geom, m.n_tiles);
        std::vector<double> h((size_t)(4 * nb));
        check(am_memcpy_d2h(h.data(), partial.gpu_buffer_ptr, h.size() * sizeof(double), stream), "constraints");
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

}  // namespace am
