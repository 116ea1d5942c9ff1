// A plugin written the way the reference writes its own (AMR.h:296-334,
// BSSN.h:104-171): component bundles (vector_field, tensor2_field,
// tensor2_sym_field, comp_tensor2_sym_field) driven through the wrappers'
// reference-shaped forms.  Two modes run the same program:
//
//   stream_pool_demo pool     the bundle forms over round_robin_threads
//                             (copy_to_child, copy_new_to_old,
//                             copy_boundaries, yield_first_second_derivs),
//                             with cuda_sync() between phases
//   stream_pool_demo serial   one buffer at a time on the default stream
//                             (copy_boundaries(old, new, amr), ...,
//                             differentiate + mixed_partials)
//
// Each prints one line per buffer, "<name> <fnv1a-64 of its used slots>
// <nonzero 32-bit words>", then "seed <s> used <n>"; both modes must print
// the same lines.
#include <stdio.h>
#include <string.h>

#include <string>
#include <vector>

#include <adaptivemesh/AMR.hpp>
#include <adaptivemesh/SimpleWave.hpp>

using namespace am;

static bool pool_mode = false;

struct pool_AMR_data {
    fast_prng rng;
    AMR<pool_AMR_data> &parent;
    am_geometry geom;
    void *stream = nullptr;
    round_robin_threads threads;

    vector_field shift_old, shift_new;
    tensor2_field k_old, k_new;
    tensor2_sym_field metric_old, metric_new;
    smart_gpu_buffer<compressed_float3> gradient;
    comp_tensor2_sym_field hessian;

    static size_t tile(const am_geometry &g)
    {
        const size_t T = (size_t)(g.size_domain + 2 * g.padding_domain);
        return T * T * T;
    }
    static size_t inner(const am_geometry &g)
    {
        const size_t S = (size_t)g.size_domain;
        return S * S * S;
    }
    pool_AMR_data(uint32_t slots, AMR<pool_AMR_data> &p, const am_geometry &g)
        : rng(), parent(p), geom(g), shift_old(slots, tile(g)), shift_new(slots, tile(g)), k_old(slots, tile(g)),
          k_new(slots, tile(g)), metric_old(slots, tile(g)), metric_new(slots, tile(g)), gradient(slots * inner(g)),
          hessian(slots, inner(g))
    {
        for (int c = 0; c < 3; c++) { shift_old[c].zero(); shift_new[c].zero(); }
        for (int c = 0; c < 9; c++) { k_old[c].zero(); k_new[c].zero(); }
        for (int c = 0; c < 6; c++) { metric_old[c].zero(); metric_new[c].zero(); }
        gradient.zero();
        hessian[0].zero();
        hessian[1].zero();
        check(cuda_sync(), "pool_AMR_data");
    }

    template <class B>
    void for_bundles(B &&f)
    {
        f(shift_old, shift_new);
        f(k_old, k_new);
        f(metric_old, metric_new);
    }

    void node_ctor(int node_idx)
    {
        const uint32_t off = (uint32_t)parent.positions.cpu_buffer_ptr[node_idx].final_offset();
        const int par = parent.parent_idx_b.cpu_buffer_ptr[node_idx];
        for_bundles([&](auto &o, auto &n) {
            if (pool_mode) {
                check(copy_to_child(o, off, node_idx, par, geom, threads), "copy_to_child");
                check(copy_to_child(n, off, node_idx, par, geom, threads), "copy_to_child");
            } else {
                for (int c = 0; c < o.components; c++) {
                    check(copy_to_child(o[c], off, node_idx, par, geom, stream), "copy_to_child");
                    check(copy_to_child(n[c], off, node_idx, par, geom, stream), "copy_to_child");
                }
            }
        });
        check(cuda_sync(), "node_ctor");
    }
    void node_dtor(int) {}
    void modify_nodes() {}
    void copy_back()
    {
        for_bundles([&](auto &o, auto &n) {
            if (pool_mode)
                check(copy_new_to_old(o, n, parent, threads), "copy_new_to_old");
            else
                for (int c = 0; c < o.components; c++) check(copy_new_to_old(o[c], n[c], parent), "copy_new_to_old");
        });
        check(cuda_sync(), "copy_back");
    }
    // the metric's xx and yz components through the derivative wrappers
    void derivatives(tensor2_sym_field &metric)
    {
        for (int c : {0, 4}) {
            if (pool_mode) {
                check(yield_first_second_derivs(metric[c], gradient, hessian, parent, threads, rng), "derivs");
            } else {
                check(differentiate(metric[c], gradient, hessian.diag, parent, stream, rng), "differentiate");
                check(mixed_partials(metric[c], hessian.off_diag, parent, stream, rng), "mixed_partials");
            }
            check(cuda_sync(), "derivatives");
        }
    }
    void predictor() { derivatives(metric_old); }
    void copy_bounds()
    {
        for_bundles([&](auto &o, auto &n) {
            if (pool_mode)
                check(copy_boundaries(o, n, parent, threads), "copy_boundaries");
            else
                for (int c = 0; c < o.components; c++) check(copy_boundaries(o[c], n[c], parent), "copy_boundaries");
        });
        check(cuda_sync(), "copy_bounds");
    }
    void corrector() { derivatives(metric_new); }
    void timestep()  // AMR.h:324-333
    {
        do {
            copy_back();
            modify_nodes();
            predictor();
            copy_bounds();
            corrector();
            parent.increment_timer();
        } while (parent.timer_helper != 0u);
    }
};

template <class E>
static void report(const char *name, const smart_gpu_buffer<E> &b, size_t n)
{
    std::vector<uint32_t> h(n * sizeof(E) / 4);
    check(am_memcpy_d2h(h.data(), b.gpu_buffer_ptr, n * sizeof(E), nullptr), "report");
    uint64_t x = 1469598103934665603ull;
    size_t nz = 0;
    for (uint32_t w : h) {
        nz += w != 0u;
        for (int k = 0; k < 4; k++) {
            x ^= (w >> (8 * k)) & 255u;
            x *= 1099511628211ull;
        }
    }
    printf("%s %016llx %zu\n", name, (unsigned long long)x, nz);
}

int main(int argc, char **argv)
{
    if (argc != 2 || (strcmp(argv[1], "pool") && strcmp(argv[1], "serial"))) {
        fprintf(stderr, "usage: %s pool|serial\n", argv[0]);
        return 2;
    }
    pool_mode = !strcmp(argv[1], "pool");
    am_geometry g;
    if (am_geometry_init(&g, 16, 3, 4, 50.f, .4f)) return 3;
    try {
        AMR<pool_AMR_data> amr(32u, g);
        pool_AMR_data &d = amr.sim_data;
        float mul = 1.f;
        d.for_bundles([&](auto &o, auto &n) {
            for (int c = 0; c < o.components; c++, mul += .25f) init_temp(o[c], n[c], amr, mul);
        });
        for (int c = 0; c < 8; c++) amr.add_node(0, c);
        const int a = amr.add_node(1, 7), b = amr.add_node(8, 0);
        for (int c = 0; c < 8; c += 3) amr.add_node(a, c);
        amr.add_node(b, 2);
        d.timestep();
        check(cuda_sync(), "timestep");

        const size_t used = amr.curr_used_slots();
        const size_t tile = pool_AMR_data::tile(g), inner = pool_AMR_data::inner(g);
        const char *sn[3] = {"shift.x", "shift.y", "shift.z"};
        const char *kn[9] = {"k.xx", "k.xy", "k.xz", "k.yx", "k.yy", "k.yz", "k.zx", "k.zy", "k.zz"};
        const char *mn[6] = {"metric.xx", "metric.xy", "metric.xz", "metric.yy", "metric.yz", "metric.zz"};
        for (int c = 0; c < 3; c++) {
            report((std::string(sn[c]) + ".old").c_str(), d.shift_old[c], used * tile);
            report((std::string(sn[c]) + ".new").c_str(), d.shift_new[c], used * tile);
        }
        for (int c = 0; c < 9; c++) {
            report((std::string(kn[c]) + ".old").c_str(), d.k_old[c], used * tile);
            report((std::string(kn[c]) + ".new").c_str(), d.k_new[c], used * tile);
        }
        for (int c = 0; c < 6; c++) {
            report((std::string(mn[c]) + ".old").c_str(), d.metric_old[c], used * tile);
            report((std::string(mn[c]) + ".new").c_str(), d.metric_new[c], used * tile);
        }
        report("gradient", d.gradient, used * inner);
        report("hessian.diag", d.hessian.diag, used * inner);
        report("hessian.off_diag", d.hessian.off_diag, used * inner);
        printf("seed %d used %zu\n", d.rng.seed, used);
    } catch (const std::exception &e) {
        fprintf(stderr, "error: %s\n", e.what());
        return 1;
    }
    return 0;
}
