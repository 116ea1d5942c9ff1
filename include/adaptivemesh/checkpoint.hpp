// Checkpoint files of the drivers: the octree, its sub-cycling counter, the
// host seed and every array the next timestep reads, so that a run resumed
// from a file continues bit for bit.
//
// The reference only dumps raw buffers (io_helper.h:9-40, write_to_file and
// read_from_file).  A file here starts with am_checkpoint_header (geometry,
// slot count, counter, seed) and then holds named sections, each a 16-byte
// NUL-padded name, a uint64 byte count and the bytes.  A load reads and
// checks the whole file against the running driver before it changes
// anything; a save writes "<path>.tmp" and renames it over <path>.
#pragma once
#include <stdio.h>
#include <string.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "../adaptivemesh_host.h"
#include "AMR.hpp"
#include "buffers.hpp"

namespace am {

static_assert(sizeof(am_checkpoint_header) == 76, "checkpoint header layout");

struct checkpoint_error : std::runtime_error {
    int code;
    checkpoint_error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

class checkpoint_out {
    FILE *f = nullptr;
    std::string path, tmp;
    std::vector<char> staging;

   public:
    explicit checkpoint_out(const char *p) : path(p), tmp(std::string(p) + ".tmp")
    {
        f = fopen(tmp.c_str(), "wb");
        if (!f) throw checkpoint_error(AM_ERR_IO, "cannot create " + tmp);
    }
    ~checkpoint_out()
    {
        if (f) {
            fclose(f);
            remove(tmp.c_str());
        }
    }
    void raw(const void *d, size_t n)
    {
        if (n && fwrite(d, 1, n, f) != n) throw checkpoint_error(AM_ERR_IO, "write failed: " + tmp);
    }
    void begin(const char *name, uint64_t bytes)
    {
        char nm[16] = {};
        strncpy(nm, name, 15);
        raw(nm, 16);
        raw(&bytes, 8);
    }
    void host(const char *name, const void *d, size_t bytes)
    {
        begin(name, bytes);
        raw(d, bytes);
    }
    // one section from device ranges written back to back
    void device(const char *name, const std::vector<std::pair<const void *, size_t>> &parts, void *stream)
    {
        uint64_t total = 0;
        for (auto &p : parts) total += p.second;
        begin(name, total);
        for (auto &p : parts) {
            staging.resize(p.second);
            check(am_memcpy_d2h(staging.data(), p.first, p.second, stream), "checkpoint read-back");
            raw(staging.data(), p.second);
        }
    }
    // part of a section opened with begin(), read back from the device
    void device_part(const void *d, size_t bytes, void *stream)
    {
        staging.resize(bytes);
        check(am_memcpy_d2h(staging.data(), d, bytes, stream), "checkpoint read-back");
        raw(staging.data(), bytes);
    }
    void commit()
    {
        const bool ok = fflush(f) == 0;
        fclose(f);
        f = nullptr;
        if (!ok || rename(tmp.c_str(), path.c_str()) != 0) {
            remove(tmp.c_str());
            throw checkpoint_error(AM_ERR_IO, "cannot write " + path);
        }
    }
};

class checkpoint_in {
    std::vector<char> bytes;
    struct section {
        std::string name;
        size_t offset;
        uint64_t size;
    };
    std::vector<section> sections;
    size_t next = 0;  // sections are read in file order

   public:
    am_checkpoint_header header{};

    explicit checkpoint_in(const char *path)
    {
        FILE *f = fopen(path, "rb");
        if (!f) throw checkpoint_error(AM_ERR_IO, std::string("cannot open ") + path);
        fseek(f, 0, SEEK_END);
        const long n = ftell(f);
        fseek(f, 0, SEEK_SET);
        bytes.resize(n > 0 ? (size_t)n : 0);
        const bool ok = n >= 0 && fread(bytes.data(), 1, bytes.size(), f) == bytes.size();
        fclose(f);
        if (!ok) throw checkpoint_error(AM_ERR_IO, std::string("cannot read ") + path);
        if (bytes.size() < sizeof header) throw checkpoint_error(AM_ERR_FORMAT, "file shorter than its header");
        memcpy(&header, bytes.data(), sizeof header);
        if (memcmp(header.magic, AM_CHECKPOINT_MAGIC, 8) != 0)
            throw checkpoint_error(AM_ERR_FORMAT, "not a checkpoint file (magic)");
        size_t at = sizeof header;
        for (uint32_t i = 0; i < header.n_sections; i++) {
            if (bytes.size() - at < 24) throw checkpoint_error(AM_ERR_FORMAT, "truncated section table");
            char nm[17] = {};
            memcpy(nm, bytes.data() + at, 16);
            uint64_t sz;
            memcpy(&sz, bytes.data() + at + 16, 8);
            at += 24;
            if (sz > bytes.size() - at) throw checkpoint_error(AM_ERR_FORMAT, std::string("truncated section ") + nm);
            sections.push_back({nm, at, sz});
            at += (size_t)sz;
        }
        if (at != bytes.size()) throw checkpoint_error(AM_ERR_FORMAT, "trailing bytes after the last section");
    }
    // the next section, which must carry `name` and exactly `size` bytes
    const char *take(const char *name, uint64_t size)
    {
        if (next >= sections.size()) throw checkpoint_error(AM_ERR_FORMAT, std::string("missing section ") + name);
        const section &s = sections[next++];
        if (s.name != name)
            throw checkpoint_error(AM_ERR_FORMAT, "section " + s.name + " where " + name + " was expected");
        if (s.size != size)
            throw checkpoint_error(AM_ERR_FORMAT, "section " + s.name + " holds " + std::to_string(s.size) +
                                                      " bytes, the driver needs " + std::to_string(size));
        return bytes.data() + s.offset;
    }
    void rewind() { next = 0; }
    void finish() const
    {
        if (next != sections.size()) throw checkpoint_error(AM_ERR_FORMAT, "unexpected section " + sections[next].name);
    }
};

inline void upload(void *dst, const void *src, size_t bytes, void *stream)
{
    check(am_memcpy_h2d(dst, src, bytes, stream), "checkpoint upload");
}

// Header of a driver over `tree` (an AMR<>), before the tree and state
// sections.
template <class Tree>
am_checkpoint_header checkpoint_header(const Tree &tree, uint32_t kind, const am_geometry &g, int32_t seed,
                                       uint32_t n_vars, uint32_t flags, uint32_t n_state_sections)
{
    am_checkpoint_header h{};
    memcpy(h.magic, AM_CHECKPOINT_MAGIC, 8);
    h.kind = kind;
    h.geometry = g;
    h.max_slots = tree.max_slots;
    h.periodic = tree.periodic ? 1u : 0u;
    h.max_depth = tree.read_max_depth();
    h.first_free_slot = tree.read_first_free_slot();
    h.final_used_slot = tree.curr_used_slots() - 1u;
    h.timer = tree.timer_helper;
    h.seed = seed;
    h.n_vars = n_vars;
    h.flags = flags;
    h.n_sections = 5u + n_state_sections;
    return h;
}

// The header of `in` against what the running driver expects.
inline void checkpoint_match(const am_checkpoint_header &f, const am_checkpoint_header &want)
{
    auto fail = [](const std::string &m) { throw checkpoint_error(AM_ERR_FORMAT, "checkpoint does not fit: " + m); };
    if (f.kind != want.kind) fail("driver kind " + std::to_string(f.kind) + ", expected " + std::to_string(want.kind));
    if (memcmp(&f.geometry, &want.geometry, sizeof(am_geometry)) != 0) fail("tile geometry differs");
    if (f.max_slots != want.max_slots) fail("max_slots " + std::to_string(f.max_slots));
    if (f.periodic != want.periodic) fail("periodic flag differs");
    if (f.n_vars != want.n_vars) fail("field count differs");
    if (f.flags != want.flags) fail("driver options differ (flags " + std::to_string(f.flags) + ")");
    if (f.n_sections != want.n_sections) fail("section count differs");
    if (f.final_used_slot >= f.max_slots || f.first_free_slot > f.max_slots) fail("slot counters out of range");
    if (f.max_depth < 0 || f.max_depth > (int32_t)depth_limit) fail("max_depth out of range");
    if (f.timer >= (1u << f.max_depth)) fail("counter out of range");
}


// The header of a file after checking it on its own: the section table, the
// octree's consistency, and the state sections' names and sizes for its
// driver kind.  Host memory only; no GPU needed.
inline am_checkpoint_header inspect_checkpoint(const char *path)
{
    checkpoint_in in(path);
    const am_checkpoint_header &h = in.header;
    auto fail = [](const std::string &m) { throw checkpoint_error(AM_ERR_FORMAT, m); };
    if (h.kind < AM_CKPT_WAVE || h.kind > AM_CKPT_SWAVE) fail("unknown driver kind " + std::to_string(h.kind));
    if (h.max_slots == 0 || h.max_slots > (1u << 23)) fail("max_slots out of range");
    if (h.final_used_slot >= h.max_slots || h.first_free_slot > h.max_slots) fail("slot counters out of range");
    if (h.max_depth < 0 || h.max_depth > (int32_t)depth_limit) fail("max_depth out of range");
    const int64_t S = h.geometry.size_domain, P = h.geometry.padding_domain;
    if (S < 2 || S > 64 || P < 1 || P > 8 || h.n_vars == 0 || h.n_vars > AM_MAX_VARS) fail("geometry out of range");
    const size_t n = h.max_slots, used = h.final_used_slot + 1u, T = (size_t)(S + 2 * P);
    octree t((uint32_t)n, h.periodic != 0);
    octree::snapshot s{};
    s.parent = reinterpret_cast<const int32_t *>(in.take("parent", n * sizeof(int32_t)));
    s.children = reinterpret_cast<const int32_t *>(in.take("children", 8 * n * sizeof(int32_t)));
    s.positions = reinterpret_cast<const uint32_t *>(in.take("positions", n * sizeof(uint32_t)));
    s.boundaries = reinterpret_cast<const uint64_t *>(in.take("boundaries", 6 * n * sizeof(uint64_t)));
    s.lifetime = reinterpret_cast<const uint32_t *>(in.take("lifetime", n * sizeof(uint32_t)));
    s.max_depth = h.max_depth;
    s.first_free = h.first_free_slot;
    s.final_used = h.final_used_slot;
    s.timer = h.timer;
    const char *why = t.check_snapshot(s);
    t.destroy();
    if (why) fail(std::string("octree: ") + why);
    const size_t fb = used * T * T * T * sizeof(float);
    if (h.kind == AM_CKPT_WAVE) {
        if (h.n_vars != 1 || h.n_sections != 9) fail("SimpleWave checkpoints hold 1 field in 9 sections");
        in.take("old", fb);
        in.take("new", fb);
        in.take("d1", n * (size_t)(S * S * S) * 4);
        in.take("d2diag", n * (size_t)(S * S * S) * 4);
    } else {
        if (h.n_sections != 7 && h.n_sections != 8) fail("RK4 checkpoints hold 7 or 8 sections");
        in.take("old", h.n_vars * fb);
        in.take("new", h.n_vars * fb);
        if (h.n_sections == 8) in.take("older", h.n_vars * fb);
    }
    in.finish();
    return h;
}

// Save the driver of `amr` (an AMR<Data>).  Data supplies CKPT_KIND,
// ckpt_vars(), ckpt_flags(), ckpt_sections() and ckpt_save(out), and owns
// `geom`, `rng` and `stream`.
template <class AMR_t>
void save_checkpoint(AMR_t &amr, const char *path)
{
    auto &d = amr.sim_data;
    check(am_stream_sync(d.stream), "checkpoint");
    const am_checkpoint_header h = checkpoint_header(amr, d.CKPT_KIND, d.geom, d.rng.seed, d.ckpt_vars(),
                                                     d.ckpt_flags(), d.ckpt_sections());
    const size_t n = amr.max_slots;
    checkpoint_out o(path);
    o.raw(&h, sizeof h);
    o.host("parent", amr.parent_idx_b.cpu_buffer_ptr, n * sizeof(int32_t));
    o.host("children", amr.children_idx_b.cpu_buffer_ptr, 8 * n * sizeof(int32_t));
    o.host("positions", amr.positions.cpu_buffer_ptr, n * sizeof(uint32_t));
    o.host("boundaries", amr.boundaries.cpu_buffer_ptr, 6 * n * sizeof(uint64_t));
    o.host("lifetime", amr.lifetime.cpu_buffer_ptr, n * sizeof(uint32_t));
    d.ckpt_save(o);
    o.commit();
}

// Load a file saved by save_checkpoint into a driver built alike.  Nothing
// changes unless the whole file checks out.
template <class AMR_t>
void load_checkpoint(AMR_t &amr, const char *path)
{
    auto &d = amr.sim_data;
    checkpoint_in in(path);
    checkpoint_match(in.header, checkpoint_header(amr, d.CKPT_KIND, d.geom, d.rng.seed, d.ckpt_vars(),
                                                  d.ckpt_flags(), d.ckpt_sections()));
    const size_t n = amr.max_slots;
    const uint32_t used = in.header.final_used_slot + 1u;
    auto take_tree = [&]() {
        octree::snapshot s{};
        s.parent = reinterpret_cast<const int32_t *>(in.take("parent", n * sizeof(int32_t)));
        s.children = reinterpret_cast<const int32_t *>(in.take("children", 8 * n * sizeof(int32_t)));
        s.positions = reinterpret_cast<const uint32_t *>(in.take("positions", n * sizeof(uint32_t)));
        s.boundaries = reinterpret_cast<const uint64_t *>(in.take("boundaries", 6 * n * sizeof(uint64_t)));
        s.lifetime = reinterpret_cast<const uint32_t *>(in.take("lifetime", n * sizeof(uint32_t)));
        s.max_depth = in.header.max_depth;
        s.first_free = in.header.first_free_slot;
        s.final_used = in.header.final_used_slot;
        s.timer = in.header.timer;
        return s;
    };
    const octree::snapshot s = take_tree();
    if (const char *why = amr.check_snapshot(s)) throw checkpoint_error(AM_ERR_FORMAT, std::string("octree: ") + why);
    d.ckpt_load(in, used, false);
    in.finish();
    check(am_stream_sync(d.stream), "checkpoint");
    amr.restore(s);
    in.rewind();
    take_tree();
    d.ckpt_load(in, used, true);
    d.rng.seed = in.header.seed;
    check(am_stream_sync(d.stream), "checkpoint upload");
}

}  // namespace am
