// Native host runtime behind adaptivemesh_host.h: instantiates the drop-in
// C++ API (include/adaptivemesh/*.hpp) and exports it as opaque handles.
#include <new>
#include <type_traits>
#include <vector>
#include <string.h>
#include <stdexcept>

#include "../../include/adaptivemesh/AMR.hpp"
#include "../../include/adaptivemesh/BSSN.hpp"
#include "../../include/adaptivemesh/ScalarWave.hpp"
#include "../../include/adaptivemesh/SimpleWave.hpp"
#include "../../include/adaptivemesh/image.hpp"
#include "../../include/adaptivemesh_host.h"
#include "am_common.hpp"

using namespace am;

// the drop-in image templates compile against the driver types
template void am::rasterise_to_image<AMR<wave_AMR_data>, smart_gpu_buffer<float>, smart_gpu_cpu_buffer<rgba_f32>>(
    const smart_gpu_buffer<float> &, smart_gpu_cpu_buffer<rgba_f32> &, AMR<wave_AMR_data> &, uint32_t, uint32_t, float,
    bool, void *);
template void am::save_image<smart_gpu_cpu_buffer<uint32_t>, smart_gpu_cpu_buffer<rgba_f32>>(
    smart_gpu_cpu_buffer<uint32_t> &, const smart_gpu_cpu_buffer<rgba_f32> &, uint32_t, uint32_t, const char *, void *);

struct am_tree {
    AMR<blank_AMR_data> *amr;
};

struct am_wave_sim {
    AMR<wave_AMR_data> *amr;
};

// RK4 drivers: one handle type per System, one set of generic bodies
template <class Data>
struct rk4_sim {
    AMR<Data> *amr;
};
struct am_bssn_sim : rk4_sim<BSSN_rk4_data> {};
struct am_swave_sim : rk4_sim<swave_rk4_data> {};

namespace {

template <class F>
int guarded(F &&f)
{
    try {
        return f();
    } catch (const am::checkpoint_error &e) {
        return am::set_error(e.code, "%s", e.what());
    } catch (const std::exception &e) {
        return am::set_error(AM_ERR_INVALID, "%s", e.what());
    }
}

am_geometry null_geometry()
{
    am_geometry g{};
    am_geometry_init(&g, 16, 3, 6, 1.f, .25f);
    return g;
}

template <class Tree>
int export_tree(const Tree &a, int32_t *parent, int32_t *children, uint32_t *positions, uint64_t *boundaries,
                uint32_t *lifetime, int32_t *scalars)
{
    const size_t n = a.max_slots;
    if (parent) memcpy(parent, a.parent_idx_b.cpu_buffer_ptr, n * sizeof(int32_t));
    if (children) memcpy(children, a.children_idx_b.cpu_buffer_ptr, 8 * n * sizeof(int32_t));
    if (positions) memcpy(positions, a.positions.cpu_buffer_ptr, n * sizeof(uint32_t));
    if (boundaries)
        for (size_t i = 0; i < 6 * n; i++) boundaries[i] = a.boundaries.cpu_buffer_ptr[i].packed();
    if (lifetime) memcpy(lifetime, a.lifetime.cpu_buffer_ptr, n * sizeof(uint32_t));
    if (scalars) {
        scalars[0] = a.read_max_depth();
        scalars[1] = (int32_t)a.read_first_free_slot();
        scalars[2] = (int32_t)a.curr_used_slots() - 1;
        scalars[3] = (int32_t)a.timer_helper;
    }
    return AM_OK;
}

}  // namespace

// ------------------------------------------------- RK4 driver bodies
// Shared by every System's handle (am_bssn_sim, am_swave_sim).
namespace {

template <class H, class Data, class SysOpt>
H *rk4_create(const char *name, const am_geometry *g, uint32_t max_slots, int32_t interp, int32_t z_block,
              int32_t periodic, int32_t seed, int32_t word_format, const SysOpt &so)
{
    if (!g || !am_geometry_supported(g)) {
        am::set_error(AM_ERR_UNSUPPORTED_GEOMETRY, "%s: unsupported geometry", name);
        return nullptr;
    }
    try {
        am::rk4_options opt;
        opt.interp = interp;
        opt.z_block = z_block;
        opt.word_format = word_format;
        auto *s = new H;
        s->amr = new AMR<Data>(max_slots, periodic != 0, *g, opt, so);
        s->amr->sim_data.rng.seed = seed;
        return s;
    } catch (const std::exception &e) {
        am::set_error(AM_ERR_ALLOC, "%s: %s", name, e.what());
        return nullptr;
    }
}
template <class H>
void rk4_destroy(H *s)
{
    if (!s) return;
    s->amr->sim_data.destroy();
    s->amr->destroy();
    delete s->amr;
    delete s;
}
template <class H>
int rk4_add_node(H *s, int32_t parent, int32_t offset)
{
    if (!s) return am::set_error(AM_ERR_INVALID, "null sim");
    return guarded([&] {
        const int r = s->amr->add_node(parent, offset);
        if (r < 0) am::set_error(r, "add_node(%d, %d) refused", parent, offset);
        return r;
    });
}
template <class H>
smart_gpu_buffer<float> *rk4_state_buffer(H *s, int which)
{
    auto &d = s->amr->sim_data;
    switch (which) {
    case 0: return &d.old_;
    case 1: return &d.new_;
    case 2: return &d.acc;
    case 3: return d.older.created ? &d.older : nullptr;
    default: return nullptr;
    }
}
// interior cells of padded host tiles, and back (ghost cells zero)
std::vector<float> interior_of(const float *padded, size_t slots, const am_geometry &g)
{
    const size_t S = (size_t)g.size_domain, P = (size_t)g.padding_domain, T = S + 2 * P;
    std::vector<float> out(slots * S * S * S);
    for (size_t s = 0; s < slots; s++)
        for (size_t z = 0; z < S; z++)
            for (size_t y = 0; y < S; y++)
                memcpy(&out[((s * S + z) * S + y) * S], &padded[((s * T + z + P) * T + y + P) * T + P],
                       S * sizeof(float));
    return out;
}
void pad_from_interior(const std::vector<float> &in, float *padded, size_t slots, const am_geometry &g)
{
    const size_t S = (size_t)g.size_domain, P = (size_t)g.padding_domain, T = S + 2 * P;
    memset(padded, 0, slots * T * T * T * sizeof(float));
    for (size_t s = 0; s < slots; s++)
        for (size_t z = 0; z < S; z++)
            for (size_t y = 0; y < S; y++)
                memcpy(&padded[((s * T + z + P) * T + y + P) * T + P], &in[((s * S + z) * S + y) * S],
                       S * sizeof(float));
}

template <class H>
int rk4_set_field(H *s, int32_t which, int32_t var, const float *host)
{
    using D = typename std::remove_reference<decltype(s->amr->sim_data)>::type;
    if (!s || !host || var < 0 || var >= D::NV) return am::set_error(AM_ERR_INVALID, "set_field: bad argument");
    auto *b = rk4_state_buffer(s, which);
    if (!b) return am::set_error(AM_ERR_INVALID, "set_field: no buffer %d", which);
    auto &d = s->amr->sim_data;
    if (b == &d.acc) {  // interior-only on the device, padded on the host
        const std::vector<float> in = interior_of(host, d.slots, d.geom);
        if (int rc = am_memcpy_h2d(d.field(*b, var), in.data(), in.size() * sizeof(float), d.stream)) return rc;
        return am_stream_sync(d.stream);
    }
    return guarded([&] {  // padded on the host, the driver's layout on the device
        d.from_padded_host(d.field(*b, var), host, d.slots);
        return am_stream_sync(d.stream);
    });
}
template <class H>
int rk4_get_field(H *s, int32_t which, int32_t var, float *host)
{
    using D = typename std::remove_reference<decltype(s->amr->sim_data)>::type;
    if (!s || !host || var < 0 || var >= D::NV) return am::set_error(AM_ERR_INVALID, "get_field: bad argument");
    auto *b = rk4_state_buffer(s, which);
    if (!b) return am::set_error(AM_ERR_INVALID, "get_field: no buffer %d", which);
    auto &d = s->amr->sim_data;
    if (b == &d.acc) {
        std::vector<float> in(d.slots * d.inner);
        if (int rc = am_memcpy_d2h(in.data(), d.field(*b, var), in.size() * sizeof(float), d.stream)) return rc;
        pad_from_interior(in, host, d.slots, d.geom);
        return AM_OK;
    }
    return guarded([&] { return am_memcpy_d2h(host, d.to_padded(d.field(*b, var)), d.slots * d.padded * sizeof(float),
                                              d.stream); });
}
template <class H>
void *rk4_device_ptr(H *s, int32_t which)
{
    if (!s) return nullptr;
    auto &d = s->amr->sim_data;
    if (which >= 0 && which <= 3) {
        auto *b = rk4_state_buffer(s, which);
        return b ? (void *)b->gpu_buffer_ptr : nullptr;
    }
    switch (which) {
    case 4: return d.d1.gpu_buffer_ptr;
    case 5: return d.d2d.gpu_buffer_ptr;
    case 6: return d.d2o.gpu_buffer_ptr;
    case 7: return d.chr.gpu_buffer_ptr;
    case 8: return d.adv.gpu_buffer_ptr;
    default: return nullptr;
    }
}
template <class H, class F>
int rk4_call(H *s, F &&f)
{
    if (!s) return am::set_error(AM_ERR_INVALID, "null sim");
    return guarded([&] {
        f(s->amr->sim_data);
        return AM_OK;
    });
}
template <class H>
int rk4_counter(const H *s, uint32_t *timer, int32_t *max_depth)
{
    if (!s) return am::set_error(AM_ERR_INVALID, "null sim");
    if (timer) *timer = s->amr->timer_helper;
    if (max_depth) *max_depth = s->amr->read_max_depth();
    return AM_OK;
}

}  // namespace

extern "C" {

am_tree *am_tree_create(uint32_t max_slots, int32_t periodic)
{
    if (max_slots == 0 || max_slots > (1u << 23)) {
        am::set_error(AM_ERR_INVALID, "am_tree_create: max_slots %u", max_slots);
        return nullptr;
    }
    try {
        auto *t = new am_tree;
        t->amr = new AMR<blank_AMR_data>(max_slots, periodic != 0, null_geometry());
        return t;
    } catch (const std::exception &e) {
        am::set_error(AM_ERR_ALLOC, "am_tree_create: %s", e.what());
        return nullptr;
    }
}
void am_tree_destroy(am_tree *t)
{
    if (!t) return;
    t->amr->destroy();
    delete t->amr;
    delete t;
}
int am_tree_add_node(am_tree *t, int32_t parent, int32_t offset)
{
    if (!t) return am::set_error(AM_ERR_INVALID, "null tree");
    return guarded([&] {
        const int r = t->amr->add_node(parent, offset);
        if (r < 0) am::set_error(r, "add_node(%d, %d) refused", parent, offset);
        return r;
    });
}
int am_tree_remove_node(am_tree *t, int32_t node)
{
    if (!t) return am::set_error(AM_ERR_INVALID, "null tree");
    return guarded([&] { return t->amr->remove_node(node); });
}
int am_tree_increment_timer(am_tree *t)
{
    if (!t) return am::set_error(AM_ERR_INVALID, "null tree");
    t->amr->increment_timer();
    return AM_OK;
}
uint32_t am_tree_find_node_idx(const am_tree *t, uint32_t packed_pos, uint32_t max_depth)
{
    return t ? t->amr->find_node_idx(packed_pos, max_depth) : 0u;
}
int am_tree_export(const am_tree *t, int32_t *parent, int32_t *children, uint32_t *positions, uint64_t *boundaries,
                   uint32_t *lifetime, int32_t *scalars)
{
    if (!t) return am::set_error(AM_ERR_INVALID, "null tree");
    return export_tree(*t->amr, parent, children, positions, boundaries, lifetime, scalars);
}
int am_tree_mesh(am_tree *t, int32_t depth, am_mesh *out)
{
    if (!t || !out) return am::set_error(AM_ERR_INVALID, "null tree or output");
    if (depth < -1 || depth > (int)depth_limit) return am::set_error(AM_ERR_INVALID, "bad depth %d", depth);
    return guarded([&] {
        *out = depth < 0 ? t->amr->mesh_all() : t->amr->mesh_depth(depth);
        return AM_OK;
    });
}

am_wave_sim *am_wave_sim_create(const am_geometry *g, uint32_t max_slots, int32_t periodic)
{
    if (!g || !am_geometry_supported(g)) {
        am::set_error(AM_ERR_UNSUPPORTED_GEOMETRY, "am_wave_sim_create: unsupported geometry");
        return nullptr;
    }
    try {
        auto *s = new am_wave_sim;
        s->amr = new AMR<wave_AMR_data>(max_slots, periodic != 0, *g);
        s->amr->sim_data.old_field.zero();
        s->amr->sim_data.new_field.zero();
        return s;
    } catch (const std::exception &e) {
        am::set_error(AM_ERR_ALLOC, "am_wave_sim_create: %s", e.what());
        return nullptr;
    }
}
void am_wave_sim_destroy(am_wave_sim *s)
{
    if (!s) return;
    s->amr->sim_data.destroy();
    s->amr->destroy();
    delete s->amr;
    delete s;
}
int am_wave_sim_add_node(am_wave_sim *s, int32_t parent, int32_t offset)
{
    if (!s) return am::set_error(AM_ERR_INVALID, "null sim");
    return guarded([&] {
        const int r = s->amr->add_node(parent, offset);
        if (r < 0) am::set_error(r, "add_node(%d, %d) refused", parent, offset);
        return r;
    });
}
int am_wave_sim_export(const am_wave_sim *s, int32_t *parent, int32_t *children, uint32_t *positions,
                       uint64_t *boundaries, uint32_t *lifetime, int32_t *scalars)
{
    if (!s) return am::set_error(AM_ERR_INVALID, "null sim");
    return export_tree(*s->amr, parent, children, positions, boundaries, lifetime, scalars);
}
int am_wave_sim_set_field(am_wave_sim *s, int32_t which, const float *host)
{
    if (!s || !host || which < 0 || which > 1) return am::set_error(AM_ERR_INVALID, "am_wave_sim_set_field");
    auto &b = which == 0 ? s->amr->sim_data.old_field : s->amr->sim_data.new_field;
    if (int rc = am_memcpy_h2d(b.gpu_buffer_ptr, host, b.total_size(), s->amr->sim_data.stream)) return rc;
    return am_stream_sync(s->amr->sim_data.stream);
}
int am_wave_sim_get_field(am_wave_sim *s, int32_t which, float *host)
{
    if (!s || !host || which < 0 || which > 1) return am::set_error(AM_ERR_INVALID, "am_wave_sim_get_field");
    auto &b = which == 0 ? s->amr->sim_data.old_field : s->amr->sim_data.new_field;
    return am_memcpy_d2h(host, b.gpu_buffer_ptr, b.total_size(), s->amr->sim_data.stream);
}
int am_wave_sim_copy_bounds(am_wave_sim *s)
{
    if (!s) return am::set_error(AM_ERR_INVALID, "null sim");
    return guarded([&] {
        s->amr->sim_data.copy_bounds();
        return am_stream_sync(s->amr->sim_data.stream);
    });
}
int am_wave_sim_timestep(am_wave_sim *s)
{
    if (!s) return am::set_error(AM_ERR_INVALID, "null sim");
    return guarded([&] {
        s->amr->sim_data.timestep();
        return AM_OK;
    });
}
int32_t am_wave_sim_rng_state(const am_wave_sim *s) { return s ? s->amr->sim_data.rng.seed : 0; }
int am_wave_sim_set_regrid(am_wave_sim *s, int32_t on)
{
    if (!s) return am::set_error(AM_ERR_INVALID, "null sim");
    s->amr->sim_data.regrid = on != 0;
    return AM_OK;
}
int am_checkpoint_inspect(const char *path, am_checkpoint_header *out)
{
    if (!path) return am::set_error(AM_ERR_INVALID, "null path");
    return guarded([&] {
        const am_checkpoint_header h = inspect_checkpoint(path);
        if (out) *out = h;
        return AM_OK;
    });
}
int am_wave_sim_save(am_wave_sim *s, const char *path)
{
    if (!s || !path) return am::set_error(AM_ERR_INVALID, "null argument");
    return guarded([&] {
        save_checkpoint(*s->amr, path);
        return AM_OK;
    });
}
int am_wave_sim_load(am_wave_sim *s, const char *path)
{
    if (!s || !path) return am::set_error(AM_ERR_INVALID, "null argument");
    return guarded([&] {
        load_checkpoint(*s->amr, path);
        return AM_OK;
    });
}
int am_wave_sim_init_temp(am_wave_sim *s, float mul)
{
    if (!s) return am::set_error(AM_ERR_INVALID, "null sim");
    return guarded([&] {
        auto &d = s->amr->sim_data;
        init_temp(d.old_field, d.new_field, *s->amr, mul);
        return am_stream_sync(d.stream);
    });
}
int am_wave_sim_raster(am_wave_sim *s, int32_t which, uint32_t width, uint32_t height, float z_slice, int32_t boxes,
                       uint32_t *pixels)
{
    if (!s || !pixels || which < 0 || which > 1) return am::set_error(AM_ERR_INVALID, "am_wave_sim_raster");
    return guarded([&] {
        auto &d = s->amr->sim_data;
        raster_to_host(*s->amr, (which ? d.new_field : d.old_field).gpu_buffer_ptr, width, height, z_slice, boxes != 0,
                       pixels, d.stream);
        return AM_OK;
    });
}
int am_wave_sim_pressure(const am_wave_sim *s, float *out2)
{
    if (!s || !out2) return am::set_error(AM_ERR_INVALID, "null argument");
    out2[0] = s->amr->sim_data.last_pressure[0];
    out2[1] = s->amr->sim_data.last_pressure[1];
    return AM_OK;
}

// ----------------------------------------------------- RK4 driver exports
#define AM_RK4_SIM_EXPORTS(P, H)                                                                                     \
    void P##_destroy(H *s) { rk4_destroy(s); }                                                                     \
    int P##_add_node(H *s, int32_t parent, int32_t offset) { return rk4_add_node(s, parent, offset); }             \
    int P##_export(const H *s, int32_t *parent, int32_t *children, uint32_t *positions, uint64_t *boundaries,      \
                   uint32_t *lifetime, int32_t *scalars)                                                           \
    {                                                                                                              \
        if (!s) return am::set_error(AM_ERR_INVALID, "null sim");                                                  \
        return export_tree(*s->amr, parent, children, positions, boundaries, lifetime, scalars);                   \
    }                                                                                                              \
    int P##_set_field(H *s, int32_t which, int32_t var, const float *host)                                         \
    {                                                                                                              \
        return rk4_set_field(s, which, var, host);                                                                 \
    }                                                                                                              \
    int P##_get_field(H *s, int32_t which, int32_t var, float *host) { return rk4_get_field(s, which, var, host); } \
    void *P##_device_ptr(H *s, int32_t which) { return rk4_device_ptr(s, which); }                                 \
    int P##_fill_ghosts(H *s)                                                                                      \
    {                                                                                                              \
        return rk4_call(s, [](auto &d) { d.fill_ghosts(); });                                                      \
    }                                                                                                              \
    int P##_copy_back(H *s)                                                                                        \
    {                                                                                                              \
        return rk4_call(s, [](auto &d) { d.copy_back(); });                                                        \
    }                                                                                                              \
    int P##_step_depth(H *s, int32_t depth)                                                                        \
    {                                                                                                              \
        return rk4_call(s, [&](auto &d) { d.step_depth(depth); });                                                 \
    }                                                                                                              \
    int P##_stage_depth(H *s, int32_t depth, int32_t stage)                                                        \
    {                                                                                                              \
        if (stage < 0 || stage > 3) return am::set_error(AM_ERR_INVALID, "stage %d", stage);                       \
        return rk4_call(s, [&](auto &d) { d.stage_depth(depth, stage); });                                         \
    }                                                                                                              \
    int P##_substep(H *s)                                                                                          \
    {                                                                                                              \
        return rk4_call(s, [](auto &d) { d.substep(); });                                                          \
    }                                                                                                              \
    int P##_timestep(H *s)                                                                                         \
    {                                                                                                              \
        return rk4_call(s, [](auto &d) { d.timestep(); });                                                         \
    }                                                                                                              \
    int P##_sync(H *s)                                                                                             \
    {                                                                                                              \
        if (!s) return am::set_error(AM_ERR_INVALID, "null sim");                                                  \
        return am_stream_sync(s->amr->sim_data.stream);                                                            \
    }                                                                                                              \
    uint64_t P##_zone_cycles(const H *s) { return s ? s->amr->sim_data.zone_cycles : 0; }                          \
    int32_t P##_rng_state(const H *s) { return s ? s->amr->sim_data.rng.seed : 0; }                                \
    int P##_set_rng_state(H *s, int32_t seed)                                                                      \
    {                                                                                                              \
        return rk4_call(s, [&](auto &d) { d.rng.seed = seed; });                                                   \
    }                                                                                                              \
    int P##_counter(const H *s, uint32_t *timer, int32_t *max_depth) { return rk4_counter(s, timer, max_depth); }  \
    int P##_profile(H *s, int32_t enable)                                                                          \
    {                                                                                                              \
        return rk4_call(s, [&](auto &d) { d.timer.enable(enable != 0, d.stream); });                              \
    }                                                                                                              \
    int P##_kernel_times(H *s, double *ms5, int64_t *count5)                                                       \
    {                                                                                                              \
        if (!ms5 || !count5) return am::set_error(AM_ERR_INVALID, "null output");                                  \
        return rk4_call(s, [&](auto &d) { d.timer.totals(ms5, count5); });                                         \
    }                                                                                                              \
    int P##_set_launch(H *s, int32_t z_block, int32_t fd_grid_limit, int32_t rhs_grid_limit)                       \
    {                                                                                                              \
        if (fd_grid_limit < 0 || rhs_grid_limit < 0) return am::set_error(AM_ERR_INVALID, "negative grid limit");  \
        return rk4_call(s, [&](auto &d) {                                                                          \
            d.opt.z_block = z_block;                                                                               \
            d.opt.fd_grid_limit = fd_grid_limit;                                                                   \
            d.opt.rhs_grid_limit = rhs_grid_limit;                                                                 \
            d.drop_graph();                                                                                        \
        });                                                                                                        \
    }                                                                                                              \
    int P##_get_launch(const H *s, int32_t *out4)                                                                  \
    {                                                                                                              \
        if (!s || !out4) return am::set_error(AM_ERR_INVALID, "null argument");                                    \
        const auto &o = s->amr->sim_data.opt;                                                                      \
        out4[0] = o.z_block;                                                                                       \
        out4[1] = o.fd_grid_limit;                                                                                 \
        out4[2] = o.rhs_grid_limit;                                                                                \
        out4[3] = o.fd_prefetch;                                                                                   \
        return AM_OK;                                                                                              \
    }                                                                                                              \
    int P##_set_push_layers(H *s, int32_t layers)                                                                  \
    {                                                                                                              \
        return rk4_call(s, [&](auto &d) {                                                                          \
            if (layers < 0 || layers > d.geom.padding_domain || (layers > 0 && layers < d.geom.fd_order / 2))    \
                throw std::runtime_error("push layers " + std::to_string(layers) +                               \
                                         " (0, or fd_order/2..P: the FD stencils read fd_order/2 layers)");      \
            d.opt.push_layers = layers;                                                                            \
            d.drop_graph();                                                                                        \
        });                                                                                                        \
    }                                                                                                              \
    int P##_get_push_layers(const H *s) { return s ? s->amr->sim_data.push_layers() : -1; }                       \
    int P##_set_direct_faces(H *s, int32_t enable)                                                                 \
    {                                                                                                              \
        return rk4_call(s, [&](auto &d) {                                                                          \
            if (enable < 0 || enable > 7 || (enable & 1))                                                          \
                throw std::runtime_error("direct faces axis mask " + std::to_string(enable) + " (0, 2, 4 or 6)"); \
            d.opt.direct_faces = enable;                                                                               \
            d.drop_graph();                                                                                        \
        });                                                                                                        \
    }                                                                                                              \
    int P##_get_direct_faces(const H *s) { return s ? s->amr->sim_data.opt.direct_faces : -1; }                   \
    int P##_set_overlap_depths(H *s, int32_t enable)                                                               \
    {                                                                                                              \
        if (enable != 0 && enable != 1)                                                                            \
            return am::set_error(AM_ERR_INVALID, "overlap depths %d (0 or 1)", enable);                            \
        return rk4_call(s, [&](auto &d) {                                                                          \
            d.opt.overlap_depths = enable;                                                                         \
            d.drop_graph();                                                                                        \
        });                                                                                                        \
    }                                                                                                              \
    int P##_get_overlap_depths(const H *s) { return s ? s->amr->sim_data.opt.overlap_depths : -1; }               \
    int P##_set_fd_prefetch(H *s, int32_t planes)                                                                  \
    {                                                                                                              \
        if (planes != 1)                                                                                           \
            return am::set_error(AM_ERR_INVALID, "fd prefetch %d (the interior-first FD prefetches 1)", planes);  \
        return rk4_call(s, [&](auto &d) {                                                                          \
            d.opt.fd_prefetch = planes;                                                                            \
            d.drop_graph();                                                                                        \
        });                                                                                                        \
    }                                                                                                              \
    int P##_use_graph(H *s, int32_t enable)                                                                        \
    {                                                                                                              \
        return rk4_call(s, [&](auto &d) { d.set_use_graph(enable != 0); });                                        \
    }                                                                                                              \
    int P##_autotune(H *s, int32_t reps, am_tune_report *out)                                                      \
    {                                                                                                              \
        if (reps < 1) return am::set_error(AM_ERR_INVALID, "autotune needs reps >= 1");                            \
        return rk4_call(s, [&](auto &d) { d.autotune(reps, reinterpret_cast<am::rk4_tune_report *>(out)); });      \
    }                                                                                                              \
    int P##_set_prolongation(H *s, const float *params, float threshold)                                           \
    {                                                                                                              \
        return rk4_call(s, [&](auto &d) { d.set_prolongation(params, threshold); });                               \
    }                                                                                                              \
    int P##_save(H *s, const char *path)                                                                           \
    {                                                                                                              \
        if (!path) return am::set_error(AM_ERR_INVALID, "null path");                                              \
        return rk4_call(s, [&](auto &d) { save_checkpoint(d.parent, path); });                                     \
    }                                                                                                              \
    int P##_load(H *s, const char *path)                                                                           \
    {                                                                                                              \
        if (!path) return am::set_error(AM_ERR_INVALID, "null path");                                              \
        return rk4_call(s, [&](auto &d) { load_checkpoint(d.parent, path); });                                     \
    }                                                                                                              \
    int P##_raster(H *s, int32_t which, int32_t var, uint32_t width, uint32_t height, float z_slice, int32_t boxes, \
                   uint32_t *pixels)                                                                               \
    {                                                                                                              \
        if (!pixels || which < 0 || which > 1) return am::set_error(AM_ERR_INVALID, "raster: bad argument");       \
        return rk4_call(s, [&](auto &d) {                                                                          \
            if (var < 0 || var >= d.NV) throw std::runtime_error("raster: variable out of range");                \
            raster_to_host(d.parent, d.to_padded(d.field(which ? d.new_ : d.old_, var)), width, height, z_slice,   \
                           boxes != 0, pixels, d.stream);                                                          \
        });                                                                                                        \
    }                                                                                                              \
    int P##_prolongation_stats(H *s, uint64_t *out4)                                                               \
    {                                                                                                              \
        if (!out4) return am::set_error(AM_ERR_INVALID, "null output");                                            \
        return rk4_call(s, [&](auto &d) {                                                                          \
            check(am_memcpy_d2h(out4, d.mlp_stats.gpu_buffer_ptr, 4 * sizeof(uint64_t), d.stream), "stats");       \
        });                                                                                                        \
    }

static_assert(sizeof(am_tune_report) == sizeof(am::rk4_tune_report), "tune report layout");
AM_RK4_SIM_EXPORTS(am_bssn_sim, am_bssn_sim)
AM_RK4_SIM_EXPORTS(am_swave_sim, am_swave_sim)

am_bssn_sim *am_bssn_sim_create(const am_geometry *g, uint32_t max_slots, const am_bssn_sim_options *o)
{
    if (!o) {
        am::set_error(AM_ERR_INVALID, "am_bssn_sim_create: null options");
        return nullptr;
    }
    return rk4_create<am_bssn_sim, BSSN_rk4_data>("am_bssn_sim_create", g, max_slots, o->interp, o->z_block,
                                                  o->periodic, o->seed, o->word_format, o->params);
}
int am_bssn_sim_constraints(am_bssn_sim *s, double *out4)
{
    if (!out4) return am::set_error(AM_ERR_INVALID, "null output");
    return rk4_call(s, [&](auto &d) { d.diagnostics(out4); });
}
int am_bssn_sim_set_dissipation(am_bssn_sim *s, float ko_sigma, int32_t upwind)
{
    return rk4_call(s, [&](auto &d) { d.set_dissipation(ko_sigma, upwind != 0 ? 1 : 0); });
}
int am_bssn_sim_get_dissipation(const am_bssn_sim *s, float *ko_sigma, int32_t *upwind)
{
    if (!s) return am::set_error(AM_ERR_INVALID, "null sim");
    if (ko_sigma) *ko_sigma = s->amr->sim_data.opt.ko_sigma;
    if (upwind) *upwind = s->amr->sim_data.opt.upwind;
    return AM_OK;
}

am_swave_sim *am_swave_sim_create(const am_geometry *g, uint32_t max_slots, const am_swave_sim_options *o)
{
    if (!o) {
        am::set_error(AM_ERR_INVALID, "am_swave_sim_create: null options");
        return nullptr;
    }
    return rk4_create<am_swave_sim, swave_rk4_data>("am_swave_sim_create", g, max_slots, o->interp, o->z_block,
                                                    o->periodic, o->seed, AM_WORD_CF3, swave_options{});
}
int am_swave_sim_energy(am_swave_sim *s, double *out4)
{
    if (!out4) return am::set_error(AM_ERR_INVALID, "null output");
    return rk4_call(s, [&](auto &d) { d.diagnostics(out4); });
}

}  // extern "C"
