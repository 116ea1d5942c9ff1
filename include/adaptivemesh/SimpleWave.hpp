// SimpleWave plugin of the drop-in API (SimpleWave.h:191-279): viscous
// Burgers-type advection-diffusion u_t = 0.01 lap u - u u_x integrated with
// Heun's method, each substep running copy_back -> modify_nodes -> predictor
// -> copy_bounds -> corrector -> increment_timer until the counter wraps.
// The reference's modify_nodes re-grids from a refinement criterion
// (SimpleWave.h:228-248); this plugin keeps a static mesh, so it is empty.
#pragma once
#include "AMR.hpp"

namespace am {

struct wave_AMR_data {
    fast_prng rng;
    AMR<wave_AMR_data> &parent;
    am_geometry geom;
    void *stream = nullptr;

    smart_gpu_buffer<float> old_field, new_field;
    smart_gpu_buffer<compressed_float3> first_derivs, hessian_diag;

    wave_AMR_data(uint32_t node_slots, AMR<wave_AMR_data> &p, const am_geometry &g)
        : rng(), parent(p), geom(g), old_field(cells(g) * node_slots), new_field(cells(g) * node_slots),
          first_derivs(inner(g) * node_slots), hessian_diag(inner(g) * node_slots)
    {
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
    void modify_nodes() {}
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

}  // namespace am
