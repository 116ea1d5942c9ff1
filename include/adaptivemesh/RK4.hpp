// Generic classical-RK4 plugin for the octree driver, parameterised by a
// physics System.  The reference integrates with Heun's method and a single
// old/new pair (SimpleWave.h:249-278); this plugin keeps the reference's
// octree, counter and phase order (copy_back at the start of each sub-step,
// then the active depths, then increment_timer) and adds RK4 stages with
// Berger-Oliger sub-cycling: depth d advances every 2^(max_depth - d)
// sub-steps, depths run coarse first, so a fine tile's ghosts interpolate a
// coarse step that is complete.  Per depth and stage:
//   FD words of every field that needs derivatives (one batched launch)
//   -> Christoffel words (if the System uses them)
//   -> System RHS + RK4 update, which also pushes equal-depth face ghosts
//   -> interpolated ghosts for faces whose target is coarser.
//
// A System provides: NV fields; FD_KIND[v] (0 none, 1 gradient, 2 gradient
// + diagonal Hessian, 3 gradient + full Hessian); CHRISTOFFEL (metric in
// fields 1..6); options_t; rhs_stage(...) and diagnostics(...).
//
// Launch machinery (no counterpart in the reference, which launches eagerly
// between device-wide syncs): the driver owns one stream; every depth level
// is one launch per kernel over its tile list, optionally persistent (a
// capped grid whose workgroups loop over the level's work items); with
// use_graph, a whole timestep() -- every sub-step of the coarse step -- is
// recorded once as a hipGraph and replayed, the dither seeds coming from a
// device table that a one-thread kernel refills before each replay with
// exactly the draws the eager schedule would make; autotune() times the
// launch shapes on the live mesh and keeps the fastest.
#pragma once
#include <math.h>

#include <algorithm>
#include <stdexcept>
#include <vector>

#include "AMR.hpp"
#include "checkpoint.hpp"

namespace am {

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

struct rk4_options {
    int32_t interp = AM_INTERP_LAGRANGE;  // coarse-face ghosts: reference quadratic, tricubic + quadratic time,
                                          // or the learned prolongation (tricubic where it does not apply)
    int32_t z_block = 0;                  // FD planes per workgroup (0 = whole tile)
    int32_t fd_grid_limit = 0;            // persistent FD launch: workgroup cap (0 = one per item)
    int32_t rhs_grid_limit = 0;           // persistent RHS launch: workgroup cap (0 = one per item)
    int32_t word_format = AM_WORD_CF3;    // derivative-word storage (am_word_format)
    int32_t fd_prefetch = 1;              // FD planes prefetched ahead (1: the interior-first layout's FD)
    int32_t push_layers = 0;              // equal-depth ghost layers pushed (0 = the FD half-width)
    // axes (bits 1 x, 2 y, 4 z) whose equal-depth faces stages 0-2 do not
    // push; the next stage's FD reads them from the neighbours' interiors.
    // Stage 3 pushes every face, so the ghost layers are current after
    // every fine step whatever the mask.  In the interior-first layout a z
    // ghost plane is the neighbour's contiguous interior plane and a y ghost
    // row one of its interior rows, while x ghost columns would be strided
    // reads, so the mask is 0, 2, 4 or 6.
    int32_t direct_faces = 4;
    // upwinded (lopsided fourth-order) advection with Kreiss-Oliger
    // dissipation of strength ko_sigma (sixth order), computed by the FD
    // kernel into `adv` and used by the System's RHS (BSSN only; needs P = 3
    // and pushes all P ghost layers)
    int32_t upwind = 0;
    float ko_sigma = 0.f;
    // The copy-back and the coarser depths of a sub-step run on a second
    // stream beside the finest depth, which reads nothing they write until
    // its first coarse-face fill (their kernels touch only their own tiles
    // and equal-depth neighbours); that fill waits for them, and the stage-0
    // RHS waits for the copy-back (substep()).  The launches, their order
    // on each stream and the dither draws are those of the one-stream
    // schedule, so the results are the same bit for bit.  0: one stream.
    int32_t overlap_depths = 1;
};

// What autotune() measured: each trial's launch shape and mean ms per launch.
struct rk4_tune_report {
    int32_t z_block, fd_grid_limit, rhs_grid_limit, n_trials;
    struct {
        int32_t kind;  // 0 FD z_block, 1 FD grid limit, 2 RHS grid limit, 3 FD prefetch planes, 4 direct faces,
                       // 5 overlap depths
        int32_t value;
        float ms;      // best of the repetitions, one whole coarse step
    } trial[32];
    int32_t fd_prefetch;
    int32_t direct_faces;
    int32_t overlap_depths;
};

// Device views of the driver's buffers handed to a System's kernels.
struct rk4_views {
    int32_t tile_layout;  // am_tile_layout of state, old_ and new_
    const float *state[AM_MAX_VARS];
    const float *old_[AM_MAX_VARS];
    float *new_[AM_MAX_VARS];
    float *acc[AM_MAX_VARS];
    const uint32_t *d1[AM_MAX_VARS];      // gradient word of field v
    const uint32_t *d2diag[AM_MAX_VARS];  // by Hessian index (fields with FD_KIND >= 2, in order)
    const uint32_t *d2off[AM_MAX_VARS];   // by full-Hessian index (FD_KIND == 3, in order)
    const uint32_t *chr[6];
    const float *adv[AM_MAX_VARS];  // upwinded advection + dissipation per field, or null
    int32_t word_format;  // am_word_format of d1/d2diag/d2off
    int32_t push_layers;  // equal-depth ghost layers the stage kernel pushes
    int32_t push_axes;    // and the axes whose faces it pushes (bits 1 x, 2 y, 4 z)
};

template <class System>
struct rk4_AMR_data {
    static constexpr int NV = System::NV;
    using options_t = typename System::options_t;
    // The driver keeps old, new and older in the interior-first tile layout
    // (am_tile_layout): the stage kernels then stream whole interiors and the
    // equal-depth push writes whole slab rows.  Everything it exchanges with
    // its user (set/get_field, checkpoints, images) is in the reference's
    // padded layout, converted on the device through `conv`.
    static constexpr int32_t LAYOUT = AM_LAYOUT_INTERIOR_FIRST;

    fast_prng rng;
    AMR<rk4_AMR_data> &parent;
    am_geometry geom;
    rk4_options opt;
    options_t sys;
    void *stream = nullptr;
    void *aux_stream = nullptr, *fork_ev = nullptr, *join_ev = nullptr;  // the overlap of coarser depths
    void *copied_ev = nullptr;      // the copy-back done, on the second stream
    void *stage0_wait = nullptr;    // the next stage-0 RHS waits for this event (then cleared)
    void *fill_wait = nullptr;      // the next coarse-face fill waits for this event (then cleared)
    size_t cells = 0, inner = 0, slots = 0;
    size_t padded = 0;  // T^3: floats per slot of a padded field
    size_t word_stride = 0;  // bytes of one field's words: slots * inner * am_word_bytes(word_format)
    int n_d2 = 0, n_off = 0;

    // one allocation per kind; field v at base + v * slots * cells (padded
    // tiles) or v * slots * inner (acc: interior-only, so every accumulator
    // row is a whole aligned run)
    smart_gpu_buffer<float> old_, new_, acc, older;
    smart_gpu_buffer<float> conv;  // one padded field, for conversions (allocated on first use)
    smart_gpu_buffer<uint32_t> d1, d2d, d2o, chr;
    smart_gpu_buffer<float> adv;  // NV interior-only arrays (allocated with set_dissipation)
    smart_gpu_buffer<double> partial;
    uint64_t zone_cycles = 0;  // interior cells advanced by one full step, all depths
    kernel_timer timer;

    // recorded timestep
    static constexpr int32_t SEED_CAPACITY = 1 << 16;
    smart_gpu_buffer<int32_t> seed_table;
    bool use_graph = false, recording = false;
    int32_t n_draws = 0;             // draws of one recorded timestep
    void *graph = nullptr;           // instantiated hipGraphExec
    uint64_t graph_version = ~0ull;  // octree version it was recorded against
    uint64_t graph_zone_cycles = 0;  // zone-cycles one replay advances

    // learned prolongation (AM_INTERP_LEARNED): caller-owned parameters
    const float *mlp_params = nullptr;
    float mlp_threshold = 1e-5f;
    smart_gpu_buffer<uint64_t> mlp_stats;
    smart_gpu_buffer<float> mlp_tile_resid;  // per slot: the per-tile gate's residual

    rk4_AMR_data(uint32_t node_slots, AMR<rk4_AMR_data> &p, const am_geometry &g, const rk4_options &o,
                 const options_t &so)
        : rng(), parent(p), geom(g), opt(o), sys(so)
    {
        const size_t T = (size_t)(g.size_domain + 2 * g.padding_domain), S = (size_t)g.size_domain;
        padded = T * T * T;
        const int64_t stride = am_tile_stride(&g, LAYOUT);
        if (stride <= 0) throw std::runtime_error("no interior-first layout for this geometry");
        cells = (size_t)stride;
        inner = S * S * S;
        slots = node_slots;
        for (int v = 0; v < NV; v++) {
            n_d2 += System::FD_KIND[v] >= 2;
            n_off += System::FD_KIND[v] == 3;
        }
        if (opt.word_format != AM_WORD_CF3 &&
            (!System::WORD_FORMATS || g.size_domain != 16 || g.padding_domain != 3 || g.fd_order != 4 ||
             am_word_bytes(opt.word_format) <= 0))
            throw std::runtime_error("word format " + std::to_string(opt.word_format) +
                                     " needs the BSSN system on S=16, P=3, fd_order 4 tiles");
        word_stride = slots * inner * (size_t)am_word_bytes(opt.word_format);
        // the stage and FD kernels address one field or word array with 32-bit byte offsets
        if (slots * cells * sizeof(float) > 0xFFFFFFFFull || word_stride > 0xFFFFFFFFull)
            throw std::runtime_error(std::to_string(slots) + " slots: one field array would exceed 4 GiB");
        old_ = smart_gpu_buffer<float>(NV * slots * cells);
        new_ = smart_gpu_buffer<float>(NV * slots * cells);
        acc = smart_gpu_buffer<float>(NV * slots * inner);
        if (opt.interp == AM_INTERP_TRICUBIC || opt.interp == AM_INTERP_LEARNED)
            older = smart_gpu_buffer<float>(NV * slots * cells);
        mlp_stats = smart_gpu_buffer<uint64_t>(4);
        mlp_stats.zero();
        if (opt.interp == AM_INTERP_LEARNED) mlp_tile_resid = smart_gpu_buffer<float>(slots);
        d1 = smart_gpu_buffer<uint32_t>(NV * word_stride / 4);
        if (n_d2) d2d = smart_gpu_buffer<uint32_t>(n_d2 * word_stride / 4);
        if (n_off) d2o = smart_gpu_buffer<uint32_t>(n_off * word_stride / 4);
        if (System::CHRISTOFFEL) chr = smart_gpu_buffer<uint32_t>(6 * slots * inner);
        partial = smart_gpu_buffer<double>(4 * (size_t)am_bssn_constraint_blocks(&geom, (int32_t)slots) + 4);
        seed_table = smart_gpu_buffer<int32_t>(SEED_CAPACITY);
        old_.zero(); new_.zero(); acc.zero();
        if (older.created) older.zero();
        check(am_stream_create(&stream), "am_stream_create");
        check(am_stream_create(&aux_stream), "am_stream_create");
        check(am_event_create_ordering(&fork_ev), "am_event_create_ordering");
        check(am_event_create_ordering(&join_ev), "am_event_create_ordering");
        check(am_event_create_ordering(&copied_ev), "am_event_create_ordering");
    }
    void destroy()
    {
        drop_graph();
        if (stream) am_stream_sync(stream);
        old_.destroy(); new_.destroy(); acc.destroy(); older.destroy(); conv.destroy(); adv.destroy();
        d1.destroy(); d2d.destroy(); d2o.destroy(); chr.destroy(); partial.destroy(); seed_table.destroy();
        mlp_stats.destroy();
        mlp_tile_resid.destroy();
        timer.destroy();
        if (aux_stream) am_stream_sync(aux_stream);
        if (fork_ev) am_event_destroy(fork_ev);
        if (join_ev) am_event_destroy(join_ev);
        if (copied_ev) am_event_destroy(copied_ev);
        am_stream_destroy(aux_stream);
        am_stream_destroy(stream);
        stream = aux_stream = fork_ev = join_ev = copied_ev = nullptr;
    }
    void drop_graph()
    {
        if (graph) am_graph_destroy(graph);
        graph = nullptr;
        graph_version = ~0ull;
    }
    float *field(smart_gpu_buffer<float> &b, int v)
    {
        return b.gpu_buffer_ptr + (size_t)v * slots * (&b == &acc || &b == &adv ? inner : cells);
    }
    // upwinded advection with Kreiss-Oliger dissipation (upwind = 1) or the
    // centred advection of the words (upwind = 0, sigma = 0)
    void set_dissipation(float sigma, int32_t upwind)
    {
        if (upwind && System::SHIFT_FIELD < 0) throw std::runtime_error("this system has no shift to upwind along");
        if (!(sigma >= 0.f) || (sigma > 0.f && !upwind))
            throw std::runtime_error("dissipation needs upwind = 1 and sigma >= 0");
        if (upwind && (geom.padding_domain != 3 || geom.size_domain != 16 || geom.fd_order != 4 ||
                       opt.word_format != AM_WORD_CF3))
            throw std::runtime_error("upwinding needs S = 16, P = 3, fd_order 4 and compressed_float3 words");
        opt.upwind = upwind;
        opt.ko_sigma = sigma;
        if (upwind && !adv.created) adv = smart_gpu_buffer<float>((size_t)NV * slots * inner);
        drop_graph();
    }
    uint32_t *word(smart_gpu_buffer<uint32_t> &b, int v)
    {
        // Christoffel words are always compressed_float3
        const size_t stride = &b == &chr ? slots * inner * 4 : word_stride;
        return reinterpret_cast<uint32_t *>(reinterpret_cast<char *>(b.gpu_buffer_ptr) + (size_t)v * stride);
    }

    // Equal-depth ghost layers the stage kernels refresh: the FD stencils
    // read fd_order / 2 of the P layers, so by default the outer ones are
    // left as they are (they only matter to a reader of raw ghost cells)
    int32_t push_layers() const
    {
        const int32_t P = geom.padding_domain;
        if (opt.upwind) return P;  // the lopsided and sixth-difference stencils read every layer
        const int32_t l = opt.push_layers > 0 ? opt.push_layers : std::max(1, geom.fd_order / 2);
        return std::min(l, P);
    }

    // the padded form of field `f` (device, in driver layout) in `conv`
    const float *to_padded(const float *f)
    {
        if (!conv.created) conv = smart_gpu_buffer<float>(slots * padded);
        check(am_tile_layout_convert(f, conv.gpu_buffer_ptr, (int32_t)slots, &geom, LAYOUT, AM_LAYOUT_PADDED, stream),
              "layout conversion");
        return conv.gpu_buffer_ptr;
    }
    // field `f` (device, in driver layout) from the padded host array `h`
    // of n_slots slots
    void from_padded_host(float *f, const float *h, size_t n_slots)
    {
        if (!conv.created) conv = smart_gpu_buffer<float>(slots * padded);
        upload(conv.gpu_buffer_ptr, h, n_slots * padded * sizeof(float), stream);
        check(am_tile_layout_convert(conv.gpu_buffer_ptr, f, (int32_t)n_slots, &geom, AM_LAYOUT_PADDED, LAYOUT, stream),
              "layout conversion");
    }

    am_field_batch batch(bool with_older)
    {
        am_field_batch b{};
        b.n_vars = NV;
        b.tile_layout = LAYOUT;
        for (int v = 0; v < NV; v++) {
            b.old_[v] = field(old_, v);
            b.new_[v] = field(new_, v);
            b.older[v] = with_older && older.created ? field(older, v) : nullptr;
        }
        return b;
    }
    rk4_views views(bool stage_input_is_old)
    {
        rk4_views f{};
        f.tile_layout = LAYOUT;
        f.word_format = opt.word_format;
        f.push_layers = push_layers();
        f.push_axes = 7;
        for (int v = 0; v < NV; v++) {
            f.state[v] = stage_input_is_old ? field(old_, v) : field(new_, v);
            f.old_[v] = field(old_, v);
            f.new_[v] = field(new_, v);
            f.acc[v] = field(acc, v);
            f.d1[v] = word(d1, v);
        }
        for (int h = 0; h < n_d2; h++) f.d2diag[h] = word(d2d, h);
        for (int h = 0; h < n_off; h++) f.d2off[h] = word(d2o, h);
        if (System::CHRISTOFFEL)
            for (int c = 0; c < 6; c++) f.chr[c] = word(chr, c);
        if (opt.upwind)
            for (int v = 0; v < NV; v++) f.adv[v] = field(adv, v);
        return f;
    }

    void node_ctor(int node_idx)
    {
        const uint32_t off = (uint32_t)parent.positions.cpu_buffer_ptr[node_idx].final_offset();
        const int par = parent.parent_idx_b.cpu_buffer_ptr[node_idx];
        const am_field_batch b = batch(true);
        check(am_copy_to_child_batched(&b, off, node_idx, par, &geom, AM_INTERP_LAGRANGE, stream), "node_ctor");
    }
    void node_dtor(int) {}
    void modify_nodes() {}

    // checkpoint sections (checkpoint.hpp): old, new and older of the used
    // slots, variable-major
    static constexpr uint32_t CKPT_KIND = System::CKPT_KIND;
    uint32_t ckpt_vars() const { return (uint32_t)NV; }
    uint32_t ckpt_flags() const
    {
        return (uint32_t)opt.interp | ((uint32_t)opt.word_format << 4) | ((uint32_t)(opt.upwind != 0) << 8);
    }
    uint32_t ckpt_sections() const { return older.created ? 3u : 2u; }
    // (the file holds padded tiles, whatever the driver's layout)
    void ckpt_save(checkpoint_out &o)
    {
        const size_t fb = parent.curr_used_slots() * padded * sizeof(float);
        smart_gpu_buffer<float> *bufs[3] = {&old_, &new_, &older};
        const char *names[3] = {"old", "new", "older"};
        for (uint32_t k = 0; k < ckpt_sections(); k++) {
            o.begin(names[k], (uint64_t)NV * fb);
            for (int v = 0; v < NV; v++) o.device_part(to_padded(field(*bufs[k], v)), fb, stream);
        }
    }
    void ckpt_load(checkpoint_in &in, uint32_t used, bool apply)
    {
        const size_t fb = used * padded * sizeof(float);
        smart_gpu_buffer<float> *bufs[3] = {&old_, &new_, &older};
        const char *names[3] = {"old", "new", "older"};
        for (uint32_t k = 0; k < ckpt_sections(); k++) {
            const char *src = in.take(names[k], NV * fb);
            if (apply)
                for (int v = 0; v < NV; v++)
                    from_padded_host(field(*bufs[k], v), reinterpret_cast<const float *>(src + v * fb), used);
        }
        if (apply) drop_graph();
    }

    // the next dither seed: a host draw, or while recording the index of the
    // draw in the seed table
    int32_t draw()
    {
        if (!recording) return rng.generate_int();
        if (n_draws >= SEED_CAPACITY) throw std::runtime_error("recorded timestep needs more than 65536 seeds");
        return n_draws++;
    }

    // derivative words of the state in `src` for the tiles of `m`; seeds are
    // drawn per field in order (gradient+diagonal, then mixed), then one for
    // the Christoffel words
    void derivatives(smart_gpu_buffer<float> &src, am_mesh m, int direct_axes = 0)
    {
        am_fd_batch b{};
        b.tile_layout = LAYOUT;
        b.direct_faces = direct_axes;
        b.z_block = opt.z_block;
        b.word_format = opt.word_format;
        b.prefetch_planes = opt.fd_prefetch;
        b.seed_table = recording ? seed_table.gpu_buffer_ptr : nullptr;
        m.grid_limit = opt.fd_grid_limit;
        int h = 0, o = 0;
        const uint32_t amask = opt.upwind ? System::advect_mask(sys) : 0u;
        if (opt.upwind) {
            for (int k = 0; k < 3; k++) b.shift[k] = field(src, System::SHIFT_FIELD + k);
            b.ko_sigma = opt.ko_sigma;
        }
        for (int v = 0; v < NV; v++) {
            const int kind = System::FD_KIND[v];
            if (kind == 0) continue;
            const int i = b.n_vars++;
            b.field[i] = field(src, v);
            b.d1[i] = word(d1, v);
            if (opt.upwind) {
                b.adv[i] = field(adv, v);
                b.advect_mask |= ((amask >> v) & 1u) << i;
            }
            b.seed_d[i] = draw();
            if (kind >= 2) b.d2diag[i] = word(d2d, h++);
            if (kind == 3) {
                b.d2off[i] = word(d2o, o++);
                b.seed_m[i] = draw();
            }
        }
        int ev = timer.begin(kernel_timer::FD);
        check(am_fd_batched(&b, &m, &geom, stream), "fd");
        timer.end(ev);
        if (System::CHRISTOFFEL) {
            const float *met[6];
            const uint32_t *dg[6];
            uint32_t *out[6];
            for (int c = 0; c < 6; c++) {
                met[c] = field(src, 1 + c);
                dg[c] = word(d1, 1 + c);
                out[c] = word(chr, c);
            }
            ev = timer.begin(kernel_timer::CHRISTOFFEL);
            if (recording) throw std::runtime_error("Christoffel words cannot be recorded (per-launch seed)");
            check(am_christoffel(met, dg, out, 0, &m, &geom, rng.generate_int(), stream), "christoffel");
            timer.end(ev);
        }
    }

    // copy-back: old <- new (+ restriction), keeping the previous old when a
    // third time level is in use (AMR.h:998-1051)
    void copy_back()
    {
        am_field_batch b = batch(true);
        const am_mesh m = parent.mesh_all();
        const int ev = timer.begin(kernel_timer::COPY_BACK);
        // stage 0 reads new, so old's ghost cells matter only on coarse
        // sources of a finer face
        check(am_copy_new_to_old_interior(&b, &m, parent.coarse_source.gpu_buffer_ptr, &geom, stream), "copy_back");
        timer.end(ev);
    }

    // one RK4 stage of every tile at depth d: derivative words of the stage
    // input, RHS + update (+ equal-depth ghost push), coarse-face ghosts at
    // the stage time
    void stage_depth(int d, int s)
    {
        const am_mesh m = parent.mesh_depth(d);
        if (m.n_tiles == 0) return;
        static const float frac[4] = {.5f, .5f, 1.f, 1.f};
        // the stage input: new at every stage (at stage 0 new equals old on
        // the tiles of this depth -- copy-back made it so, and nothing has
        // written their new since -- ghost cells included, which old need
        // not hold)
        // With direct faces the stage input's equal-depth ghost layers on
        // those axes are current only at stage 0 (pushed by the previous
        // step's stage 3).
        const int direct = opt.direct_faces & 7;
        derivatives(new_, m, s > 0 ? direct : 0);
        rk4_views f = views(false);
        if (s < 3) f.push_axes = 7 & ~direct;
        am_mesh mr = m;
        mr.grid_limit = opt.rhs_grid_limit;
        if (stage0_wait && s == 0) {  // the stage-0 RHS writes new, which the copy-back reads
            check(am_stream_wait_event(stream, stage0_wait), "copy-back wait");
            stage0_wait = nullptr;
        }
        int ev = timer.begin(kernel_timer::RHS);
        check(System::rhs_stage(f, s, sys, mr, geom, f.push_axes ? 1 : 0, stream), "rhs");
        timer.end(ev);
        if (d > 0 && parent.coarse_faces_at_depth(d) > 0) {
            if (fill_wait) {  // the coarser depths of this sub-step have finished their step
                check(am_stream_wait_event(stream, fill_wait), "coarser depths wait");
                fill_wait = nullptr;
            }
            am_field_batch b = batch(true);
            ev = timer.begin(kernel_timer::GHOST);
            // only the faces whose target is coarser have cells to fill
            const int32_t *faces = parent.coarse_face_list(d);
            const int32_t n_faces = parent.coarse_faces_at_depth(d);
            if (opt.interp == AM_INTERP_LEARNED) {
                if (!mlp_params) throw std::runtime_error("learned prolongation selected but no parameters set");
                check(am_ghost_fill_learned_faces(&b, &m, faces, n_faces, &geom, frac[s], mlp_params, mlp_threshold,
                                                  mlp_tile_resid.gpu_buffer_ptr, mlp_stats.gpu_buffer_ptr,
                                                  AM_EQUAL_SKIP, stream),
                      "learned ghost fill");
            } else {
                check(am_ghost_fill_stage_faces(&b, &m, faces, n_faces, &geom, frac[s], opt.interp, AM_EQUAL_SKIP,
                                                stream),
                      "ghost");
            }
            timer.end(ev);
        }
    }
    // the coarse-face fill of the chosen interpolation
    void ghost_fill(const am_field_batch &b, const am_mesh &m, float frac, int equal_mode)
    {
        if (opt.interp == AM_INTERP_LEARNED) {
            if (!mlp_params) throw std::runtime_error("learned prolongation selected but no parameters set");
            check(am_ghost_fill_learned(&b, &m, &geom, frac, mlp_params, mlp_threshold, mlp_stats.gpu_buffer_ptr,
                                        equal_mode, stream),
                  "learned ghost fill");
        } else {
            check(am_ghost_fill_stage_batched(&b, &m, &geom, frac, opt.interp, equal_mode, stream), "ghost");
        }
    }
    void set_prolongation(const float *params, float threshold)
    {
        mlp_params = params;
        mlp_threshold = threshold;
        drop_graph();
        check(am_memset_async(mlp_stats.gpu_buffer_ptr, 0, 4 * sizeof(uint64_t), stream), "stats");
    }
    void step_depth(int d)
    {
        for (int s = 0; s < 4; s++) stage_depth(d, s);
        zone_cycles += (uint64_t)parent.tiles_at_depth(d) * inner;
    }
    void substep()
    {
        int finest = -1, active = 0;
        for (int d = 0; d <= parent.read_max_depth(); d++)
            if (parent.depth_active(d)) { finest = d; active++; }
        // Two streams (per-launch event timing stays on one).  The finest
        // active depth is the tree's deepest level, so its tiles are leaves.
        // The copy-back and the coarser depths go to the second stream.  The
        // finest depth's kernels read only its own tiles and equal-depth
        // neighbours, except its coarse-face fills, which read the coarser
        // depths' step: its stage-0 derivatives only read new, its stage-0
        // RHS (the first write of new) waits for the copy-back, and its
        // first coarse-face fill waits for the coarser depths.  Without
        // coarse faces the streams join at the end of the sub-step.
        const bool overlap = opt.overlap_depths && !timer.on && finest > 0 && finest == parent.read_max_depth();
        if (!overlap) {
            copy_back();
            for (int d = 0; d <= finest; d++)
                if (parent.depth_active(d)) step_depth(d);
        } else {
            check(am_event_record(fork_ev, stream), "fork");
            check(am_stream_wait_event(aux_stream, fork_ev), "fork");
            void *const main_stream = stream;
            stream = aux_stream;
            try {
                copy_back();
                check(am_event_record(copied_ev, stream), "copy-back done");
                for (int d = 0; d < finest; d++)
                    if (parent.depth_active(d)) step_depth(d);
                check(am_event_record(join_ev, stream), "join");
            } catch (...) {
                stream = main_stream;
                throw;
            }
            stream = main_stream;
            stage0_wait = copied_ev;
            fill_wait = join_ev;
            try {
                step_depth(finest);
            } catch (...) {
                stage0_wait = fill_wait = nullptr;
                throw;
            }
            stage0_wait = fill_wait = nullptr;
            check(am_stream_wait_event(stream, join_ev), "join");
        }
        parent.increment_timer();
    }
    // the reference's timestep(): sub-steps until the counter wraps; with
    // use_graph (and the counter at 0) one replay of the recorded coarse step
    void timestep()
    {
        if (use_graph && parent.timer_helper == 0u) {
            if (!graph || graph_version != parent.tree_version()) record();
            check(am_seed_table_fill(seed_table.gpu_buffer_ptr, n_draws, rng.seed, stream), "seed table");
            for (int32_t i = 0; i < n_draws; i++) rng.generate_int();  // the host stream of draws moves on alike
            check(am_graph_launch(graph, stream), "graph launch");
            zone_cycles += graph_zone_cycles;
            return;
        }
        do {
            substep();
        } while (parent.timer_helper != 0u);
    }
    // record one coarse step (counter 0 back to 0) into a graph
    void record()
    {
        drop_graph();
        parent.mesh_all();  // device octree arrays current before capture
        const uint64_t z0 = zone_cycles;
        const bool timing = timer.on;
        timer.on = false;
        n_draws = 0;
        recording = true;
        check(am_graph_begin(stream), "graph begin");
        try {
            do {
                substep();
            } while (parent.timer_helper != 0u);
        } catch (...) {
            void *g = nullptr;
            am_graph_end(stream, &g);
            if (g) am_graph_destroy(g);
            recording = false;
            timer.on = timing;
            zone_cycles = z0;
            throw;
        }
        recording = false;
        timer.on = timing;
        check(am_graph_end(stream, &graph), "graph end");
        graph_zone_cycles = zone_cycles - z0;
        zone_cycles = z0;
        graph_version = parent.tree_version();
    }
    void set_use_graph(bool on)
    {
        use_graph = on;
        if (!on) drop_graph();
    }

    // Time the launch shapes on the live run and keep the fastest: whole
    // coarse steps (timestep(), eager) from a snapshot of the state, one
    // parameter at a time -- FD planes per workgroup, the direct-face mask
    // (every mask gives the same fields), then the FD and RHS workgroup
    // caps.  Each trial starts from the same snapshot; the
    // snapshot, the host seed, the counter and the zone-cycle count are
    // restored at the end, so the run continues as if nothing happened.
    void autotune(int reps, rk4_tune_report *rep)
    {
        drop_graph();
        const bool graph_on = use_graph;
        use_graph = false;
        const size_t fn = (size_t)NV * slots * cells, an = (size_t)NV * slots * inner;
        smart_gpu_buffer<float> s_old(fn), s_new(fn), s_acc(an), s_older;
        if (older.created) s_older = smart_gpu_buffer<float>(fn);
        auto copy_all = [&](bool save) {
            auto cp = [&](smart_gpu_buffer<float> &live, smart_gpu_buffer<float> &snap, size_t n) {
                check(save ? am_memcpy_d2d(snap.gpu_buffer_ptr, live.gpu_buffer_ptr, n * 4, stream)
                           : am_memcpy_d2d(live.gpu_buffer_ptr, snap.gpu_buffer_ptr, n * 4, stream),
                      "autotune snapshot");
            };
            cp(old_, s_old, fn);
            cp(new_, s_new, fn);
            cp(acc, s_acc, an);
            if (older.created) cp(older, s_older, fn);
        };
        copy_all(true);
        const int32_t seed = rng.seed;
        const uint32_t counter = parent.timer_helper;
        const uint64_t zc = zone_cycles;
        const std::vector<uint32_t> life(parent.lifetime.cpu_buffer_ptr, parent.lifetime.cpu_buffer_ptr + slots);
        const bool timing = timer.on;
        timer.on = false;
        void *e0 = nullptr, *e1 = nullptr;
        check(am_event_create(&e0), "event");
        check(am_event_create(&e1), "event");
        rk4_tune_report r{};
        auto trial = [&]() {
            float best_ms = 1e30f;
            for (int i = 0; i < reps; i++) {
                copy_all(false);
                rng.seed = seed;
                parent.timer_helper = counter;
                check(am_event_record(e0, stream), "event");
                timestep();
                check(am_event_record(e1, stream), "event");
                float ms = 0.f;
                check(am_event_elapsed_ms(e0, e1, &ms), "event");
                best_ms = std::min(best_ms, ms);
            }
            return best_ms;
        };
        auto sweep = [&](int kind, int32_t &knob, std::initializer_list<int> values) {
            float best = 1e30f;
            int keep = knob;
            for (int v : values) {
                knob = v;
                const float ms = trial();
                if (r.n_trials < 32) r.trial[r.n_trials++] = {kind, v, ms};
                if (ms < best) { best = ms; keep = v; }
            }
            knob = keep;
        };
        const int S = geom.size_domain, cu = std::max(am_device_cu_count(), 1);
        opt.z_block = S;
        opt.fd_grid_limit = 0;
        opt.rhs_grid_limit = 0;
        opt.fd_prefetch = 1;
        trial();  // warm-up
        sweep(0, opt.z_block, {S, S / 2, S / 4});
        sweep(4, opt.direct_faces, {4, 6, 0});
        sweep(1, opt.fd_grid_limit, {0, 8 * cu});
        sweep(2, opt.rhs_grid_limit, {0, 2 * cu, 4 * cu});
        sweep(5, opt.overlap_depths, {1, 0});
        copy_all(false);
        check(am_stream_sync(stream), "autotune");
        s_old.destroy(); s_new.destroy(); s_acc.destroy(); s_older.destroy();
        am_event_destroy(e0);
        am_event_destroy(e1);
        rng.seed = seed;
        parent.timer_helper = counter;
        std::copy(life.begin(), life.end(), parent.lifetime.cpu_buffer_ptr);
        zone_cycles = zc;
        timer.on = timing;
        use_graph = graph_on;
        r.z_block = opt.z_block;
        r.fd_grid_limit = opt.fd_grid_limit;
        r.rhs_grid_limit = opt.rhs_grid_limit;
        r.fd_prefetch = opt.fd_prefetch;
        r.direct_faces = opt.direct_faces;
        r.overlap_depths = opt.overlap_depths;
        if (rep) *rep = r;
    }
    // fill every face ghost of `new` from the current data (initial data):
    // equal-depth faces copy, coarser faces interpolate
    void fill_ghosts()
    {
        for (int d = 1; d <= parent.read_max_depth(); d++) {
            const am_mesh m = parent.mesh_depth(d);
            if (m.n_tiles == 0) continue;
            am_field_batch b = batch(false);
            ghost_fill(b, m, 1.f, AM_EQUAL_COPY);
        }
    }
    // the System's diagnostics of `new` over the leaves (FP64), four values;
    // the dither draws they use are given back, so measuring does not change
    // the trajectory
    void diagnostics(double *out4)
    {
        const int32_t seed = rng.seed;
        const am_mesh m = parent.mesh_leaves();
        derivatives(new_, m);
        const rk4_views f = views(false);
        System::diagnostics(f, sys, m, geom, partial.gpu_buffer_ptr, out4, stream);
        rng.seed = seed;
    }
};

}  // namespace am
