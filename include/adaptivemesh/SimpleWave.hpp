// SimpleWave plugin of the drop-in API (SimpleWave.h:191-279): viscous
// Burgers-type advection-diffusion u_t = 0.01 lap u - u u_x integrated with
// Heun's method, each substep running copy_back -> modify_nodes -> predictor
// -> copy_bounds -> corrector -> increment_timer until the counter wraps.
// modify_nodes re-grids from the refinement criterion as the reference does
// (SimpleWave.h:146-248): am_amr_criterion gives each tile's octant maxima,
// the host averages them into two "pressures" and removes tiles below, and
// adds children to octants above, a weight that rises as the slots fill.
// Differences: the derivative words start at zero (the reference reads them
// uninitialised at the first call); no pressure line is printed; add_node
// refuses a parent removed earlier in the same call (the reference would
// build a child from an empty position); `regrid` can turn it off.
#pragma once
#include <cmath>
#include <vector>

#include "AMR.hpp"
#include "checkpoint.hpp"

namespace am {

struct wave_AMR_data {
    fast_prng rng;
    AMR<wave_AMR_data> &parent;
    am_geometry geom;
    void *stream = nullptr;

    smart_gpu_buffer<float> old_field, new_field;
    smart_gpu_buffer<compressed_float3> first_derivs, hessian_diag;
    smart_gpu_buffer<float> criterion;  // 8 octant values per slot (device)
    std::vector<float> crit_host;       // octants of every used slot, then the per-slot maxima
    bool regrid = true;                 // run modify_nodes
    float last_pressure[2] = {0.f, 0.f};

    wave_AMR_data(uint32_t node_slots, AMR<wave_AMR_data> &p, const am_geometry &g)
        : rng(), parent(p), geom(g), old_field(cells(g) * node_slots), new_field(cells(g) * node_slots),
          first_derivs(inner(g) * node_slots), hessian_diag(inner(g) * node_slots), criterion(8 * (size_t)node_slots)
    {
        first_derivs.zero();
        hessian_diag.zero();
    }
    static size_t cells(const am_geometry &g)
    {
        const size_t T = (size_t)(g.size_domain + 2 * g.padding_domain);
        return T * T * T;
    }
    static size_t inner(const am_geometry &g)
    {
        const size_t S = (size_t)g.size_domain;
        return S * S * S;
    }
    void destroy()
    {
        old_field.destroy();
        new_field.destroy();
        first_derivs.destroy();
        hessian_diag.destroy();
        criterion.destroy();
    }

    void copy_back() { check(copy_new_to_old(old_field, new_field, parent, geom, stream), "copy_back"); }
    void node_ctor(int node_idx)  // SimpleWave.h:219-225
    {
        const uint32_t off = (uint32_t)parent.positions.cpu_buffer_ptr[node_idx].final_offset();
        const int par = parent.parent_idx_b.cpu_buffer_ptr[node_idx];
        check(copy_to_child(old_field, off, node_idx, par, geom, stream), "node_ctor");
        check(copy_to_child(new_field, off, node_idx, par, geom, stream), "node_ctor");
    }
    void node_dtor(int) {}

    // checkpoint sections (checkpoint.hpp): old and new of the used slots,
    // the derivative words of every slot
    static constexpr uint32_t CKPT_KIND = AM_CKPT_WAVE;
    uint32_t ckpt_vars() const { return 1u; }
    uint32_t ckpt_flags() const { return regrid ? 1u : 0u; }
    uint32_t ckpt_sections() const { return 4u; }
    void ckpt_save(checkpoint_out &o)
    {
        const size_t fb = parent.curr_used_slots() * cells(geom) * sizeof(float);
        const size_t wb = parent.max_slots * inner(geom) * sizeof(uint32_t);
        o.device("old", {{old_field.gpu_buffer_ptr, fb}}, stream);
        o.device("new", {{new_field.gpu_buffer_ptr, fb}}, stream);
        o.device("d1", {{first_derivs.gpu_buffer_ptr, wb}}, stream);
        o.device("d2diag", {{hessian_diag.gpu_buffer_ptr, wb}}, stream);
    }
    void ckpt_load(checkpoint_in &in, uint32_t used, bool apply)
    {
        const size_t fb = used * cells(geom) * sizeof(float);
        const size_t wb = parent.max_slots * inner(geom) * sizeof(uint32_t);
        const char *o = in.take("old", fb), *n = in.take("new", fb);
        const char *w1 = in.take("d1", wb), *w2 = in.take("d2diag", wb);
        if (!apply) return;
        upload(old_field.gpu_buffer_ptr, o, fb, stream);
        upload(new_field.gpu_buffer_ptr, n, fb, stream);
        upload(first_derivs.gpu_buffer_ptr, w1, wb, stream);
        upload(hessian_diag.gpu_buffer_ptr, w2, wb, stream);
    }

    static float clampf(float f, float a, float b) { return fmaxf(a, fminf(f, b)); }
    // amr_criterion (SimpleWave.h:146-187): octant values of every used slot
    // into crit_host[0, 8s), the per-slot maxima into crit_host[8s, 9s), and
    // the two average pressures (per-slot maxima; octants / 8), each over the
    // slots that have a value
    void criterion_pressures(float &avg1, float &avg2)
    {
        const am_mesh m = parent.mesh_all();
        const uint32_t s = parent.curr_used_slots();
        check(am_amr_criterion(words(first_derivs), words(hessian_diag), &m, &geom, criterion.gpu_buffer_ptr,
                               stream),
              "amr_criterion");
        crit_host.assign((size_t)s * 9u, 0.f);
        check(am_memcpy_d2h(crit_host.data(), criterion.gpu_buffer_ptr, (size_t)s * 8u * sizeof(float), stream),
              "amr_criterion");
        float p1 = 0.f, p2 = 0.f;
        uint32_t active = 0;
        for (uint32_t i = 0; i < s; i++)
            for (int j = 0; j < 8; j++)
                if (!std::isnan(crit_host[i * 8u + j])) p2 += clampf(crit_host[i * 8u + j], 0.f, 2.f);
        for (uint32_t i = 0; i < s; i++) {
            float t = crit_host[i * 8u];
            for (int j = 1; j < 8; j++) t = fmaxf(t, crit_host[i * 8u + j]);
            crit_host[s * 8u + i] = t;
            if (!std::isnan(t)) {
                p1 += clampf(t, 0.f, 2.f);
                active++;
            }
        }
        avg1 = p1 / (float)active;
        avg2 = p2 * .125f / (float)active;
    }
    void modify_nodes()  // SimpleWave.h:228-248
    {
        if (!regrid) return;
        float avg1, avg2;
        criterion_pressures(avg1, avg2);
        const uint32_t s = parent.curr_used_slots();
        if (s == 1u) avg1 = avg2 = .15f;
        last_pressure[0] = avg1;
        last_pressure[1] = avg2;
        const float urgency = (float)(s - 1u) / (float)parent.max_slots;
        for (int i = 1; i < (int)s; i++) {  // the root stays
            const float depth_ratio =
                (float)parent.positions.cpu_buffer_ptr[i].depth() / (float)parent.read_max_depth();
            if (depth_ratio <= 0.f) continue;
            const float weight = clampf((float)parent.lifetime.cpu_buffer_ptr[i] * .1f - .2f, 0.f, 1.f) * urgency *
                                 urgency * avg1 * sqrtf(depth_ratio);
            if (crit_host[s * 8u + (uint32_t)i] < weight) parent.remove_node(i);
        }
        const float weight = avg2 * 3.f / (1.f - urgency);
        for (uint32_t i = 0, n = parent.curr_used_slots() * 8u; i < n; i++)
            if (crit_host[i] > weight && parent.children_idx_b.cpu_buffer_ptr[i] == -1)
                parent.add_node((int)(i >> 3u), (int)(i & 7u));
    }
    void predictor()  // SimpleWave.h:249-255
    {
        check(differentiate(old_field, first_derivs, hessian_diag, parent, geom, stream, rng), "predictor");
        const am_mesh m = parent.mesh_all();
        check(am_wave_predictor(old_field.gpu_buffer_ptr, words(first_derivs), words(hessian_diag),
                                new_field.gpu_buffer_ptr, &m, &geom, stream),
              "predictor");
    }
    void copy_bounds() { check(copy_boundaries(old_field, new_field, parent, geom, stream), "copy_bounds"); }
    void corrector()  // SimpleWave.h:261-267
    {
        check(differentiate(new_field, first_derivs, hessian_diag, parent, geom, stream, rng), "corrector");
        const am_mesh m = parent.mesh_all();
        check(am_wave_corrector(old_field.gpu_buffer_ptr, words(first_derivs), words(hessian_diag),
                                new_field.gpu_buffer_ptr, &m, &geom, stream),
              "corrector");
    }
    void timestep()  // SimpleWave.h:268-278
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

// init_temp (kernel.cu:34-43): the wave test's two-bump initial data in
// both fields of the used tiles (all cells at depth 0, interior deeper)
template <class T>
void init_temp(smart_gpu_buffer<float> &old_d, smart_gpu_buffer<float> &new_d, AMR<T> &amr, float mul = 1.f)
{
    const am_mesh m = amr.mesh_all();
    check(am_init_temp(old_d.gpu_buffer_ptr, new_d.gpu_buffer_ptr, &m, &amr.sim_data.geom, mul, amr.sim_data.stream),
          "init_temp");
}

}  // namespace am
