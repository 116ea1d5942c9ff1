// Native host runtime behind adaptivemesh_host.h: instantiates the drop-in
// C++ API (include/adaptivemesh/*.hpp) and exports it as opaque handles.
#include <new>
#include <stdexcept>

#include "../../include/adaptivemesh/AMR.hpp"
#include "../../include/adaptivemesh/SimpleWave.hpp"
#include "../../include/adaptivemesh_host.h"
#include "am_common.hpp"

using namespace am;

struct am_tree {
    AMR<blank_AMR_data> *amr;
};

struct am_wave_sim {
    AMR<wave_AMR_data> *amr;
};

namespace {

template <class F>
int guarded(F &&f)
{
    try {
        return f();
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

}  // extern "C"
