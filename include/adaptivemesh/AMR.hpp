// The drop-in host API of AdaptiveMesh for MI355X: the CPU-managed octree
// (AMR.h:22-294), its boundary descriptors and sub-cycling counter, and the
// free-function kernel wrappers (AMR.h:880-1310) re-expressed over the C ABI
// in adaptivemesh.h.  Names, argument order and octree semantics follow the
// reference; the differences are
//   * the tile geometry is a runtime am_geometry instead of constexpr globals;
//   * operations that the reference silently skips (full slot table, depth
//     limit, occupied child) leave the tree unchanged as there, and also
//     return an AM_ERR_* code;
//   * an optional periodic outer boundary (the reference always points an
//     outer face at the root tile);
//   * ghost fills run one depth at a time, coarse first, so a fine tile never
//     reads a coarse ghost cell while it is being written (the reference's
//     single launch leaves that order to the hardware).
#pragma once
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../adaptivemesh.h"
#include "buffers.hpp"
#include "tensors.hpp"

namespace am {

constexpr uint32_t depth_limit = 8u;  // AMR.h:9

// The reference's component bundles (AMR.h:350-633), named as there: one
// buffer per component of a vector field (x, y, z), of a tensor field (xx ..
// zz, row-major), of a symmetric tensor field (xx, xy, xz, yy, yz, zz), of
// symmetric compressed_float3 words, and of a Hessian's diagonal and
// off-diagonal words.  operator[] walks the components in that order.  The
// reference's default for domain_cells is its compile-time tile size; the
// tile size is a runtime value here, so the argument is required.
template <class Bundle, class E, int N>
struct field_bundle {
    using component_type = E;
    static constexpr int components = N;
    smart_gpu_buffer<E> &operator[](int c) { return static_cast<Bundle *>(this)->*Bundle::member(c); }
    const smart_gpu_buffer<E> &operator[](int c) const
    {
        return static_cast<const Bundle *>(this)->*Bundle::member(c);
    }
    void destroy()
    {
        for (int c = 0; c < N; c++) (*this)[c].destroy();
    }
};
struct vector_field : field_bundle<vector_field, float, 3> {
    smart_gpu_buffer<float> x, y, z;
    vector_field() = default;
    vector_field(size_t node_slots, size_t domain_cells)
        : x(node_slots * domain_cells), y(node_slots * domain_cells), z(node_slots * domain_cells)
    {
    }
    static smart_gpu_buffer<float> vector_field::*member(int c)
    {
        static smart_gpu_buffer<float> vector_field::*const m[3] = {&vector_field::x, &vector_field::y,
                                                                    &vector_field::z};
        return m[c];
    }
};
struct tensor2_field : field_bundle<tensor2_field, float, 9> {
    smart_gpu_buffer<float> xx, xy, xz, yx, yy, yz, zx, zy, zz;
    tensor2_field() = default;
    tensor2_field(size_t node_slots, size_t domain_cells)
        : xx(node_slots * domain_cells), xy(node_slots * domain_cells), xz(node_slots * domain_cells),
          yx(node_slots * domain_cells), yy(node_slots * domain_cells), yz(node_slots * domain_cells),
          zx(node_slots * domain_cells), zy(node_slots * domain_cells), zz(node_slots * domain_cells)
    {
    }
    static smart_gpu_buffer<float> tensor2_field::*member(int c)
    {
        static smart_gpu_buffer<float> tensor2_field::*const m[9] = {
            &tensor2_field::xx, &tensor2_field::xy, &tensor2_field::xz, &tensor2_field::yx, &tensor2_field::yy,
            &tensor2_field::yz, &tensor2_field::zx, &tensor2_field::zy, &tensor2_field::zz};
        return m[c];
    }
};
struct tensor2_sym_field : field_bundle<tensor2_sym_field, float, 6> {
    smart_gpu_buffer<float> xx, xy, xz, yy, yz, zz;
    tensor2_sym_field() = default;
    tensor2_sym_field(size_t node_slots, size_t domain_cells)
        : xx(node_slots * domain_cells), xy(node_slots * domain_cells), xz(node_slots * domain_cells),
          yy(node_slots * domain_cells), yz(node_slots * domain_cells), zz(node_slots * domain_cells)
    {
    }
    static smart_gpu_buffer<float> tensor2_sym_field::*member(int c)
    {
        static smart_gpu_buffer<float> tensor2_sym_field::*const m[6] = {
            &tensor2_sym_field::xx, &tensor2_sym_field::xy, &tensor2_sym_field::xz,
            &tensor2_sym_field::yy, &tensor2_sym_field::yz, &tensor2_sym_field::zz};
        return m[c];
    }
};
struct comp_tensor3_sym_field : field_bundle<comp_tensor3_sym_field, compressed_float3, 6> {
    smart_gpu_buffer<compressed_float3> xx, xy, xz, yy, yz, zz;
    comp_tensor3_sym_field() = default;
    comp_tensor3_sym_field(size_t node_slots, size_t domain_cells)
        : xx(node_slots * domain_cells), xy(node_slots * domain_cells), xz(node_slots * domain_cells),
          yy(node_slots * domain_cells), yz(node_slots * domain_cells), zz(node_slots * domain_cells)
    {
    }
    static smart_gpu_buffer<compressed_float3> comp_tensor3_sym_field::*member(int c)
    {
        static smart_gpu_buffer<compressed_float3> comp_tensor3_sym_field::*const m[6] = {
            &comp_tensor3_sym_field::xx, &comp_tensor3_sym_field::xy, &comp_tensor3_sym_field::xz,
            &comp_tensor3_sym_field::yy, &comp_tensor3_sym_field::yz, &comp_tensor3_sym_field::zz};
        return m[c];
    }
};
struct comp_tensor2_sym_field : field_bundle<comp_tensor2_sym_field, compressed_float3, 2> {
    smart_gpu_buffer<compressed_float3> diag, off_diag;
    comp_tensor2_sym_field() = default;
    comp_tensor2_sym_field(size_t node_slots, size_t domain_cells)
        : diag(node_slots * domain_cells), off_diag(node_slots * domain_cells)
    {
    }
    static smart_gpu_buffer<compressed_float3> comp_tensor2_sym_field::*member(int c)
    {
        return c ? &comp_tensor2_sym_field::off_diag : &comp_tensor2_sym_field::diag;
    }
};

// LSB ddddd|zyxzyx...: depth in bits 27-31, a 3-bit octant per level with
// level 0 lowest (AMR.h:22-55).
struct octree_abs_pos {
    uint32_t data = ~0u;
    int depth() const { return (~data) ? (int)(data >> 27u) : -1; }
    int final_offset() const { return (int)((data >> ((uint32_t)(depth() - 1) * 3u)) & 7u); }
    void set_child(const octree_abs_pos &parent, int offset)
    {
        data = parent.data + (1u << 27u);
        data |= ((uint32_t)offset) << ((uint32_t)parent.depth() * 3u);
    }
    // centre of the tile in units of outer_size (AMR.h:33-43, without the scale)
    void centre(double *c) const
    {
        c[0] = c[1] = c[2] = 0.0;
        uint32_t s = data & 134217727u;
        for (int i = 0, d = depth(); i < d; i++, s >>= 3) {
            c[0] += ((s & 1u) ? .25 : -.25) / (1 << i);
            c[1] += ((s & 2u) ? .25 : -.25) / (1 << i);
            c[2] += ((s & 4u) ? .25 : -.25) / (1 << i);
        }
    }
};

// target slot (24 bits, sign-extended) | depth difference (4) | own depth
// (4), then rel_pos = 2^d_target (c_this - c_target) / outer_size as an
// undithered compressed_float3 (AMR.h:56-100).
struct domain_boundary {
    uint32_t node_data = ~0u;
    compressed_float3 rel_pos;

    domain_boundary() = default;
    domain_boundary(const octree_abs_pos &this_d, const octree_abs_pos &other_d, int other_idx, bool periodic = false,
                    int face = -1)
        : node_data(((uint32_t)other_idx & 16777215u) | ((uint32_t)((this_d.depth() - other_d.depth()) & 15) << 24u) |
                    ((uint32_t)(this_d.depth() & 15) << 28u))
    {
        float r[3] = {0.f, 0.f, 0.f};
        const float mult = .25f * (float)(1u << other_d.depth());
        uint32_t s = this_d.data & 134217727u;
        for (int i = 0, d = this_d.depth(); i < d; i++, s >>= 3) {
            r[0] += ((s & 1u) ? mult : -mult) / (float)(1 << i);
            r[1] += ((s & 2u) ? mult : -mult) / (float)(1 << i);
            r[2] += ((s & 4u) ? mult : -mult) / (float)(1 << i);
        }
        s = other_d.data & 134217727u;
        for (int i = 0, d = other_d.depth(); i < d; i++, s >>= 3) {
            r[0] -= ((s & 1u) ? mult : -mult) / (float)(1 << i);
            r[1] -= ((s & 2u) ? mult : -mult) / (float)(1 << i);
            r[2] -= ((s & 4u) ? mult : -mult) / (float)(1 << i);
        }
        if (periodic && face >= 0) {
            // the periodic image of the target across `face`: along the face axis
            // this - target < 0 for a + face, > 0 for a - face; the cube is
            // 2^d_target target widths wide
            const float w = (float)(1u << other_d.depth());
            const int a = face % 3;
            if (face < 3 && r[a] > 0.f) r[a] -= w;
            if (face >= 3 && r[a] < 0.f) r[a] += w;
        }
        rel_pos = compressed_float3(r[0], r[1], r[2]);
    }
    int target_idx() const { return (int)((node_data & 16777215u) | ((node_data & (1u << 23u)) ? 4278190080u : 0u)); }
    uint32_t depth_diff() const { return (node_data >> 24u) & 15u; }
    int curr_depth() const { return ((node_data >> 28u) == 15u) ? -1 : (int)(node_data >> 28u); }
    uint64_t packed() const { return (uint64_t)node_data | ((uint64_t)rel_pos.data << 32); }
};
static_assert(sizeof(domain_boundary) == 8, "domain_boundary is 8 bytes on the device");

inline uint32_t reverse_bits(uint32_t a)  // helper_math.h:2256-2262
{
    a = (a >> 16u) | (a << 16u);
    a = ((a & 4278255360u) >> 8u) | ((a & 16711935u) << 8u);
    a = ((a & 4042322160u) >> 4u) | ((a & 252645135u) << 4u);
    a = ((a & 3435973836u) >> 2u) | ((a & 858993459u) << 2u);
    return ((a & 2863311530u) >> 1u) | ((a & 1431655765u) << 1u);
}

// Per-lane Morton add / subtract on the bit-reversed path; a carry out of
// the root level lands in bits 27-29 and reports ~0u when `overflow`
// (AMR.h:102-115).  The lane masks are the reference's: they define the
// encoding.
constexpr uint32_t kLaneZ = 19173961u, kFill = 115043766u, kLaneOv = 153391689u;
inline uint32_t yield_sum_coordinates(uint32_t a, uint32_t b, bool overflow = true)
{
    uint32_t z = ((a | kFill) + (b & kLaneZ)) & kLaneOv;
    uint32_t y = ((a | (kFill << 1u)) + (b & (kLaneZ << 1u))) & (kLaneOv << 1u);
    uint32_t x = ((a | (kFill << 2u)) + (b & (kLaneZ << 2u))) & (kLaneOv << 2u);
    x |= y | z;
    return ((x & (7u << 27u)) && overflow) ? ~0u : (x & ((1u << 27u) - 1u));
}
inline uint32_t yield_diff_coordinates(uint32_t a, uint32_t b, bool overflow = true)
{
    uint32_t z = ((a & kLaneZ) - (b & kLaneZ)) & kLaneOv;
    uint32_t y = ((a & (kLaneZ << 1u)) - (b & (kLaneZ << 1u))) & (kLaneOv << 1u);
    uint32_t x = ((a & (kLaneZ << 2u)) - (b & (kLaneZ << 2u))) & (kLaneOv << 2u);
    x |= y | z;
    return ((x & (7u << 27u)) && overflow) ? ~0u : (x & ((1u << 27u) - 1u));
}
constexpr uint32_t directions[3] = {4u, 2u, 1u};  // x y z in the reversed path

// The octree proper, without a physics plugin.  AMR<AMR_data> adds one.
struct octree {
    smart_gpu_cpu_buffer<int32_t> parent_idx_b;
    smart_gpu_cpu_buffer<int32_t> children_idx_b;
    smart_gpu_cpu_buffer<octree_abs_pos> positions;
    smart_gpu_cpu_buffer<domain_boundary> boundaries;
    smart_cpu_buffer<uint32_t> lifetime;
    const uint32_t max_slots;
    uint32_t timer_helper = 0;
    bool periodic = false;
    void *stream = nullptr;  // stream used for octree uploads

    // device slot lists, built on the device (am_tile_lists) after every
    // structural change: [0] every slot up to the last used one, [1 + d] the
    // slots at depth d, [depth_limit + 2] the leaves (no child)
    lazy_gpu_buffer<int32_t> tile_lists;
    int32_t list_offset[depth_limit + 3] = {};
    int32_t list_count[depth_limit + 3] = {};
    int32_t coarse_faces[depth_limit + 1] = {};  // faces per depth whose target is coarser
    // device lists of those faces (slot * 6 + face), depth by depth
    lazy_gpu_buffer<int32_t> face_lists;
    int32_t face_offset[depth_limit + 1] = {};
    // per slot: 1 when some finer tile's face reads this slot (its ghost cells
    // feed the interpolation), else 0
    lazy_gpu_buffer<uint8_t> coarse_source;
    lazy_gpu_buffer<int32_t> list_workspace, list_counts;  // the compaction's scratch and its counts

   protected:
    bool dirty = true;
    uint64_t version = 0;  // bumped by every structural change (recorded graphs check it)
    int max_depth = 0;
    uint32_t first_free_slot = 1;
    uint32_t final_used_slot = 0;

   public:
    explicit octree(uint32_t node_slots, bool periodic_ = false)
        : parent_idx_b(node_slots), children_idx_b(node_slots * 8u), positions(node_slots),
          boundaries(node_slots * 6u), lifetime(node_slots), max_slots(node_slots), periodic(periodic_)
    {
        memset(parent_idx_b.cpu_buffer_ptr, -1, parent_idx_b.total_size());
        memset(children_idx_b.cpu_buffer_ptr, -1, children_idx_b.total_size());
        memset(static_cast<void *>(positions.cpu_buffer_ptr), -1, positions.total_size());
        memset(lifetime.cpu_buffer_ptr, 0, lifetime.total_size());
        for (uint32_t i = 0; i < node_slots * 6u; i++) boundaries.cpu_buffer_ptr[i] = domain_boundary();
        positions.cpu_buffer_ptr[0].data = 0u;  // root
    }
    void destroy()
    {
        parent_idx_b.destroy();
        children_idx_b.destroy();
        positions.destroy();
        boundaries.destroy();
        lifetime.destroy();
        tile_lists.destroy();
        coarse_source.destroy();
        face_lists.destroy();
        list_workspace.destroy();
        list_counts.destroy();
    }

    uint32_t read_first_free_slot() const { return first_free_slot; }
    uint32_t curr_used_slots() const { return final_used_slot + 1u; }
    int read_max_depth() const { return max_depth; }
    bool is_dirty() const { return dirty; }
    uint64_t tree_version() const { return version; }

    // A whole tree state from outside (a checkpoint file): check_snapshot
    // returns nullptr when it is a tree this class could have built, else
    // what is wrong; restore installs a checked one.
    struct snapshot {
        const int32_t *parent;
        const int32_t *children;
        const uint32_t *positions;
        const uint64_t *boundaries;
        const uint32_t *lifetime;
        int max_depth;
        uint32_t first_free, final_used, timer;
    };
    const char *check_snapshot(const snapshot &s) const
    {
        const uint32_t n = max_slots;
        if (s.final_used >= n || s.first_free > n || s.first_free == 0) return "slot counters out of range";
        if (s.positions[0] != 0u) return "slot 0 is not the root";
        int deepest = 0;
        uint32_t first_empty = n;
        for (uint32_t i = 0; i < n; i++) {
            octree_abs_pos p;
            p.data = s.positions[i];
            const int d = p.depth();
            if (d == -1) {
                if (first_empty == n) first_empty = i;
                if (s.parent[i] != -1) return "an empty slot has a parent";
                for (int c = 0; c < 8; c++)
                    if (s.children[i * 8u + c] != -1) return "an empty slot has children";
                continue;
            }
            if (i > s.final_used) return "a used slot lies beyond final_used";
            if (d > (int)depth_limit) return "depth beyond the limit";
            deepest = std::max(deepest, d);
            for (int c = 0; c < 8; c++) {
                const int ch = s.children[i * 8u + c];
                if (ch == -1) continue;
                if (ch <= 0 || (uint32_t)ch >= n || s.parent[ch] != (int32_t)i) return "child and parent disagree";
                octree_abs_pos q;
                q.data = s.positions[ch];
                if (q.depth() != d + 1 || q.final_offset() != c) return "child position disagrees with its octant";
            }
            if (i > 0) {
                const int par = s.parent[i];
                if (par < 0 || (uint32_t)par >= n) return "parent out of range";
                octree_abs_pos q;
                q.data = s.positions[par];
                if (q.depth() != d - 1 || s.children[(uint32_t)par * 8u + (uint32_t)p.final_offset()] != (int32_t)i)
                    return "parent does not list the slot as its child";
                for (int f = 0; f < 6; f++) {
                    domain_boundary b;
                    memcpy(static_cast<void *>(&b), &s.boundaries[i * 6u + f], sizeof b);
                    if (b.target_idx() < 0 || (uint32_t)b.target_idx() >= n) return "boundary target out of range";
                }
            }
        }
        if (s.final_used > 0 && octree_abs_pos{s.positions[s.final_used]}.depth() == -1)
            return "final_used is an empty slot";
        if (s.first_free != first_empty) return "first_free is not the first empty slot";
        if (s.max_depth != deepest) return "max_depth differs from the deepest slot";
        if (s.timer >= (1u << (uint32_t)s.max_depth)) return "counter out of range";
        return nullptr;
    }
    void restore(const snapshot &s)
    {
        const size_t n = max_slots;
        memcpy(parent_idx_b.cpu_buffer_ptr, s.parent, n * sizeof(int32_t));
        memcpy(children_idx_b.cpu_buffer_ptr, s.children, 8 * n * sizeof(int32_t));
        memcpy(static_cast<void *>(positions.cpu_buffer_ptr), s.positions, n * sizeof(uint32_t));
        memcpy(static_cast<void *>(boundaries.cpu_buffer_ptr), s.boundaries, 6 * n * sizeof(uint64_t));
        memcpy(lifetime.cpu_buffer_ptr, s.lifetime, n * sizeof(uint32_t));
        max_depth = s.max_depth;
        first_free_slot = s.first_free;
        final_used_slot = s.final_used;
        timer_helper = s.timer;
        dirty = true;
        version++;
    }

    // AMR::increment_timer (AMR.h:180-188)
    void increment_timer()
    {
        for (uint32_t i = 0; i <= final_used_slot; i++) {
            const int depth = positions.cpu_buffer_ptr[i].depth();
            if (depth == -1) continue;
            lifetime.cpu_buffer_ptr[i] += ((timer_helper & ((1u << (max_depth - depth)) - 1u)) == 0u);
        }
        timer_helper = (timer_helper + 1u) & ((1u << max_depth) - 1u);
    }
    bool depth_active(int depth) const
    {
        return (timer_helper & ((1u << (uint32_t)(max_depth - depth)) - 1u)) == 0u;
    }

    // AMR::find_node_idx (AMR.h:255-263)
    uint32_t find_node_idx(uint32_t packed_pos, uint32_t max_d = ~0u) const
    {
        uint32_t curr = 0u, depth = 0u;
        while (true) {
            const int child = children_idx_b.cpu_buffer_ptr[curr * 8u + (packed_pos & 7u)];
            if (child == -1 || depth >= max_d) return curr;
            packed_pos >>= 3u;
            curr = (uint32_t)child;
            depth++;
        }
    }

    // AMR::yield_boundary (AMR.h:276-290)
    domain_boundary yield_boundary(int node_idx, uint32_t face) const
    {
        const int depth = positions.cpu_buffer_ptr[node_idx].depth();
        if (depth <= 0) return domain_boundary();
        uint32_t p = reverse_bits(positions.cpu_buffer_ptr[node_idx].data) >> 5u;
        const uint32_t step = (directions[face % 3u] << ((9u - (uint32_t)depth) * 3u)) & ((1u << 27u) - 1u);
        p = face < 3 ? yield_sum_coordinates(p, step, !periodic) : yield_diff_coordinates(p, step, !periodic);
        int new_idx = (int)find_node_idx(reverse_bits(p) >> 5u, (uint32_t)depth);
        if (p == ~0u) new_idx = 0;  // the root tile stands in for the outer boundary
        return domain_boundary(positions.cpu_buffer_ptr[node_idx], positions.cpu_buffer_ptr[new_idx], new_idx,
                               periodic, (int)face);
    }

   protected:
    // AMR::add_node (AMR.h:208-232) minus the plugin call, split so AMR<>
    // can run node_ctor at the reference's point in the sequence.
    int add_node_begin(int parent_idx, int offset, int &slot)
    {
        if (parent_idx < 0 || (uint32_t)parent_idx >= max_slots || offset < 0 || offset > 7)
            return AM_ERR_INVALID;
        if (first_free_slot >= max_slots) return AM_ERR_SLOTS_EXHAUSTED;
        if (children_idx_b.cpu_buffer_ptr[parent_idx * 8 + offset] != -1) return AM_ERR_OCCUPIED;
        if (positions.cpu_buffer_ptr[parent_idx].depth() < 0) return AM_ERR_INVALID;
        if ((uint32_t)positions.cpu_buffer_ptr[parent_idx].depth() >= depth_limit) return AM_ERR_DEPTH_LIMIT;
        slot = (int)first_free_slot;
        lifetime.cpu_buffer_ptr[slot] = 0u;
        parent_idx_b.cpu_buffer_ptr[slot] = parent_idx;
        children_idx_b.cpu_buffer_ptr[parent_idx * 8 + offset] = slot;
        positions.cpu_buffer_ptr[slot].set_child(positions.cpu_buffer_ptr[parent_idx], offset);
        const int new_depth = positions.cpu_buffer_ptr[slot].depth();
        const int change = std::max(max_depth, new_depth) - max_depth;
        timer_helper <<= change;
        max_depth += change;
        return AM_OK;
    }
    void add_node_end(int slot)
    {
        const int new_depth = positions.cpu_buffer_ptr[slot].depth();
        final_used_slot = std::max(final_used_slot, (uint32_t)slot);
        for (; first_free_slot < max_slots && positions.cpu_buffer_ptr[first_free_slot].depth() != -1;
             first_free_slot++) {
        }
        refresh_boundaries(new_depth);
        dirty = true;
        version++;
    }
    void refresh_boundaries(int min_depth)
    {
        for (uint32_t i = 0; i <= final_used_slot; i++)
            if (positions.cpu_buffer_ptr[i].depth() >= min_depth)
                for (uint32_t j = 0; j < 6; j++) boundaries.cpu_buffer_ptr[i * 6u + j] = yield_boundary((int)i, j);
    }
    template <class Dtor>
    void remove_node_recursive(int node_idx, Dtor &&dtor)  // AMR.h:134-146
    {
        if (node_idx == -1 || positions.cpu_buffer_ptr[node_idx].depth() == -1) return;
        dtor(node_idx);
        for (uint32_t i = 0; i < 8u; i++) remove_node_recursive(children_idx_b.cpu_buffer_ptr[node_idx * 8u + i], dtor);
        children_idx_b.cpu_buffer_ptr[parent_idx_b.cpu_buffer_ptr[node_idx] * 8 +
                                      positions.cpu_buffer_ptr[node_idx].final_offset()] = -1;
        parent_idx_b.cpu_buffer_ptr[node_idx] = -1;
        positions.cpu_buffer_ptr[node_idx].data = ~0u;
        first_free_slot = std::min(first_free_slot, (uint32_t)node_idx);
    }
    template <class Dtor>
    int remove_node_impl(int node_idx, Dtor &&dtor)  // AMR.h:233-254
    {
        if (node_idx <= 0 || (uint32_t)node_idx >= max_slots) return AM_ERR_INVALID;
        const int curr_depth = positions.cpu_buffer_ptr[node_idx].depth();
        if (curr_depth == -1) return AM_ERR_INVALID;
        remove_node_recursive(node_idx, dtor);
        const int old_max = max_depth;
        max_depth = 0;
        for (uint32_t i = 0; i < max_slots; i++)
            if ((int)i != node_idx) max_depth = std::max(max_depth, positions.cpu_buffer_ptr[i].depth());
        timer_helper >>= (old_max - max_depth);
        for (; positions.cpu_buffer_ptr[final_used_slot].depth() == -1 && final_used_slot > 0u; final_used_slot--) {
        }
        refresh_boundaries(curr_depth);
        dirty = true;
        version++;
        return AM_OK;
    }

   public:
    // AMR::copy_to_gpu (AMR.h:268-275), then the slot and face lists built
    // from the uploaded tree by the device compaction; only their lengths
    // come back to the host (launch sizes)
    void copy_to_gpu()
    {
        if (!dirty) return;
        parent_idx_b.copy_to_gpu(stream);
        children_idx_b.copy_to_gpu(stream);
        positions.copy_to_gpu(stream);
        boundaries.copy_to_gpu(stream);
        constexpr int NT = depth_limit + 3, NF = depth_limit + 1;
        int32_t counts[2 * NT + 2 * NF];
        const int32_t n = (int32_t)final_used_slot + 1;
        check(am_tile_lists(reinterpret_cast<const uint32_t *>(positions.gpu_buffer_ptr),
                            children_idx_b.gpu_buffer_ptr,
                            reinterpret_cast<const uint64_t *>(boundaries.gpu_buffer_ptr), n, (int32_t)max_slots,
                            (int32_t)depth_limit, tile_lists.ensure(3u * max_slots), face_lists.ensure(6u * max_slots),
                            coarse_source.ensure(max_slots),
                            list_workspace.ensure((size_t)am_tile_lists_workspace((int32_t)max_slots, depth_limit)),
                            list_counts.ensure(2 * NT + 2 * NF), stream),
              "am_tile_lists");
        check(am_memcpy_d2h(counts, list_counts.gpu_buffer_ptr, sizeof counts, stream), "tile list counts");
        check(am_stream_sync(stream), "tile list counts");
        for (int k = 0; k < NT; k++) {
            list_offset[k] = counts[k];
            list_count[k] = counts[NT + k];
        }
        for (int d = 0; d < NF; d++) {
            face_offset[d] = counts[2 * NT + d];
            coarse_faces[d] = counts[2 * NT + NF + d];
        }
        dirty = false;
    }
    int tiles_at_depth(int d) const { return list_count[d + 1]; }
    int coarse_faces_at_depth(int d) const { return coarse_faces[d]; }
    // the device list of those faces (slot * 6 + face)
    const int32_t *coarse_face_list(int d)
    {
        copy_to_gpu();
        return face_lists.gpu_buffer_ptr + face_offset[d];
    }
    int leaf_count() const { return list_count[depth_limit + 2]; }

    // Kernel view over every used slot (the reference's launch domain) ...
    am_mesh mesh_all()
    {
        copy_to_gpu();
        return make_mesh(tile_lists.gpu_buffer_ptr + list_offset[0], list_count[0]);
    }
    // ... or over the slots of one depth ...
    am_mesh mesh_depth(int d)
    {
        copy_to_gpu();
        return make_mesh(tile_lists.gpu_buffer_ptr + list_offset[d + 1], list_count[d + 1]);
    }
    // ... or over the leaves.
    am_mesh mesh_leaves()
    {
        copy_to_gpu();
        return make_mesh(tile_lists.gpu_buffer_ptr + list_offset[depth_limit + 2], list_count[depth_limit + 2]);
    }
    am_mesh make_mesh(const int32_t *tiles, int32_t n) const
    {
        am_mesh m;
        m.positions = reinterpret_cast<const uint32_t *>(positions.gpu_buffer_ptr);
        m.children = children_idx_b.gpu_buffer_ptr;
        m.boundaries = reinterpret_cast<const uint64_t *>(boundaries.gpu_buffer_ptr);
        m.tiles = tiles;
        m.n_tiles = n;
        m.substep = timer_helper;
        m.max_depth = (uint32_t)max_depth;
        m.grid_limit = 0;
        return m;
    }

    std::string to_string_debug() const { return print_node((uint32_t)0); }

   private:
    std::string print_node(uint32_t node) const  // AMR.h:147-174, abridged
    {
        const int depth = positions.cpu_buffer_ptr[node].depth();
        std::string pad((size_t)std::max(depth, 0) * 2, ' ');
        std::string r = pad + "Node index: " + std::to_string(node) + "\n" + pad +
                        "Parent index: " + std::to_string(parent_idx_b.cpu_buffer_ptr[node]) + "\n" + pad +
                        "Depth: " + std::to_string(depth) + "\n" + pad + "Children: ";
        int n = 0;
        for (uint32_t i = 0; i < 8; i++) {
            const int c = children_idx_b.cpu_buffer_ptr[node * 8u + i];
            if (c != -1) {
                r += "\n\n" + print_node((uint32_t)c);
                n++;
            }
        }
        if (!n) r += "None";
        return r;
    }
};

// The reference's AMR<AMR_data>: the octree plus one physics plugin built as
// AMR_data(node_slots, *this) (AMR.h:118-294).
template <class AMR_data>
struct AMR : octree {
    AMR_data sim_data;

    template <class... Args>
    AMR(uint32_t node_slots, const am_geometry &g, Args &&...args)
        : octree(node_slots, false), sim_data(node_slots, *this, g, std::forward<Args>(args)...)
    {
    }
    template <class... Args>
    AMR(uint32_t node_slots, bool periodic_, const am_geometry &g, Args &&...args)
        : octree(node_slots, periodic_), sim_data(node_slots, *this, g, std::forward<Args>(args)...)
    {
    }

    // AMR::add_node (AMR.h:208-232): returns the new slot, or an AM_ERR_*
    // code (< 0) with the tree unchanged.
    int add_node(int parent_idx, int offset)
    {
        int slot = -1;
        const int rc = add_node_begin(parent_idx, offset, slot);
        if (rc != AM_OK) return rc;
        sim_data.node_ctor(slot);
        add_node_end(slot);
        return slot;
    }
    // AMR::remove_node (AMR.h:233-254)
    int remove_node(int node_idx)
    {
        return remove_node_impl(node_idx, [this](int i) { sim_data.node_dtor(i); });
    }
};

// The canonical empty plugin (AMR.h:296-334).
struct blank_AMR_data {
    AMR<blank_AMR_data> &parent;
    blank_AMR_data(uint32_t, AMR<blank_AMR_data> &p, const am_geometry &) : parent(p) {}
    void node_ctor(int) {}
    void node_dtor(int) {}
    void modify_nodes() {}
    void predictor() {}
    void copy_bounds() {}
    void corrector() {}
    void copy_back() {}
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

// ----------------------------------------------------------- wrappers
// The reference's free-function wrappers (AMR.h:880-1310): each uploads the
// octree if it changed, draws one seed from the caller's fast_prng where the
// kernel dithers, and launches over every used slot.

inline uint32_t *words(smart_gpu_buffer<compressed_float3> &b)
{
    return reinterpret_cast<uint32_t *>(b.gpu_buffer_ptr);
}

template <class Tree>
int differentiate(const smart_gpu_buffer<float> &field, smart_gpu_buffer<compressed_float3> &dif, Tree &amr,
                  const am_geometry &g, void *stream, fast_prng &rng)  // AMR.h:1280-1290
{
    const am_mesh m = amr.mesh_all();
    return am_fd_d1(field.gpu_buffer_ptr, words(dif), &m, &g, rng.generate_int(), stream);
}
template <class Tree>
int differentiate(const smart_gpu_buffer<float> &field, smart_gpu_buffer<compressed_float3> &dif,
                  smart_gpu_buffer<compressed_float3> &d2if, Tree &amr, const am_geometry &g, void *stream,
                  fast_prng &rng)  // AMR.h:1292-1302
{
    const am_mesh m = amr.mesh_all();
    return am_fd_d1_d2diag(field.gpu_buffer_ptr, words(dif), words(d2if), &m, &g, rng.generate_int(), stream);
}
template <class Tree>
int mixed_partials(const smart_gpu_buffer<float> &field, smart_gpu_buffer<compressed_float3> &off, Tree &amr,
                   const am_geometry &g, void *stream, fast_prng &rng)  // AMR.h:1268-1278
{
    const am_mesh m = amr.mesh_all();
    return am_fd_mixed(field.gpu_buffer_ptr, words(off), &m, &g, rng.generate_int(), stream);
}
// yield_first_second_derivs (AMR.h:1304-1310) as one fused launch; the two
// seeds are drawn in the reference's order (differentiate, then mixed).
template <class Tree>
int yield_first_second_derivs(const smart_gpu_buffer<float> &field, smart_gpu_buffer<compressed_float3> &gradient,
                              smart_gpu_buffer<compressed_float3> &hdiag, smart_gpu_buffer<compressed_float3> &hoff,
                              Tree &amr, const am_geometry &g, void *stream, fast_prng &rng)
{
    const am_mesh m = amr.mesh_all();
    const int32_t sd = rng.generate_int();
    const int32_t sm = rng.generate_int();
    return am_fd_all(field.gpu_buffer_ptr, words(gradient), words(hdiag), words(hoff), &m, &g, sd, sm, stream);
}
// copy_boundaries (AMR.h:880-900), one launch per depth, coarse first
template <class Tree>
int copy_boundaries(smart_gpu_buffer<float> &old_bfr, smart_gpu_buffer<float> &new_bfr, Tree &amr,
                    const am_geometry &g, void *stream, bool copy_only_new = true,
                    int interp = AM_INTERP_LAGRANGE)
{
    amr.copy_to_gpu();
    for (int d = 1; d <= amr.read_max_depth(); d++) {
        if (!amr.depth_active(d) || amr.tiles_at_depth(d) == 0) continue;
        const am_mesh m = amr.mesh_depth(d);
        if (int rc = am_ghost_fill(old_bfr.gpu_buffer_ptr, new_bfr.gpu_buffer_ptr, nullptr, &m, &g,
                                   copy_only_new ? 1 : 0, interp, stream))
            return rc;
    }
    return AM_OK;
}
// copy_new_to_old (AMR.h:998-1018)
template <class Tree>
int copy_new_to_old(smart_gpu_buffer<float> &old_bfr, const smart_gpu_buffer<float> &new_bfr, Tree &amr,
                    const am_geometry &g, void *stream)
{
    const am_mesh m = amr.mesh_all();
    return am_copy_new_to_old(old_bfr.gpu_buffer_ptr, new_bfr.gpu_buffer_ptr, &m, &g, stream);
}
// copy_to_child (AMR.h:1075-1082)
inline int copy_to_child(smart_gpu_buffer<float> &buffer, uint32_t offset, int child_idx, int parent_idx,
                         const am_geometry &g, void *stream, int interp = AM_INTERP_LAGRANGE)
{
    return am_copy_to_child(buffer.gpu_buffer_ptr, offset, child_idx, parent_idx, &g, interp, stream);
}

// ------------------------------------------------- reference-shaped forms
// The wrappers as the reference's plugins call them: with the plugin's AMR<>
// and a stream (or none: the default stream), AMR.h:881-899, 998-1018,
// 1268-1302, the tile geometry taken from the plugin's `geom`.
template <class D>
int differentiate(const smart_gpu_buffer<float> &field, smart_gpu_buffer<compressed_float3> &dif, AMR<D> &amr,
                  void *stream, fast_prng &rng)
{
    return differentiate(field, dif, amr, amr.sim_data.geom, stream, rng);
}
template <class D>
int differentiate(const smart_gpu_buffer<float> &field, smart_gpu_buffer<compressed_float3> &dif,
                  smart_gpu_buffer<compressed_float3> &d2if, AMR<D> &amr, void *stream, fast_prng &rng)
{
    return differentiate(field, dif, d2if, amr, amr.sim_data.geom, stream, rng);
}
template <class D>
int mixed_partials(const smart_gpu_buffer<float> &field, smart_gpu_buffer<compressed_float3> &off, AMR<D> &amr,
                   void *stream, fast_prng &rng)
{
    return mixed_partials(field, off, amr, amr.sim_data.geom, stream, rng);
}
template <class D>
int copy_boundaries(smart_gpu_buffer<float> &old_bfr, smart_gpu_buffer<float> &new_bfr, AMR<D> &amr, void *stream,
                    bool copy_only_new = true)
{
    return copy_boundaries(old_bfr, new_bfr, amr, amr.sim_data.geom, stream, copy_only_new);
}
template <class D>
int copy_boundaries(smart_gpu_buffer<float> &old_bfr, smart_gpu_buffer<float> &new_bfr, AMR<D> &amr,
                    bool copy_only_new = true)
{
    return copy_boundaries(old_bfr, new_bfr, amr, amr.sim_data.geom, nullptr, copy_only_new);
}
template <class D>
int copy_new_to_old(smart_gpu_buffer<float> &old_bfr, const smart_gpu_buffer<float> &new_bfr, AMR<D> &amr,
                    void *stream = nullptr)
{
    return copy_new_to_old(old_bfr, new_bfr, amr, amr.sim_data.geom, stream);
}

// The component-bundle forms over a stream pool (AMR.h:902-933, 1020-1051,
// 1083-1115, 1304-1310): one launch per component, each on the pool's next
// stream.  As in the reference, launches on different streams are ordered
// neither with each other nor with later work: cuda_sync() before the
// results are read.  copy_to_child takes the tile geometry, which the
// reference's form reads from its compile-time constants.
template <class B>
using float_bundle = std::enable_if_t<std::is_same<typename B::component_type, float>::value, int>;

template <class B, class D, float_bundle<B> = 0>
int copy_boundaries(B &old_field, B &new_field, AMR<D> &amr, round_robin_threads &threads)
{
    for (int c = 0; c < B::components; c++)
        if (int rc = copy_boundaries(old_field[c], new_field[c], amr, amr.sim_data.geom, threads.yield_stream()))
            return rc;
    return AM_OK;
}
template <class B, class D, float_bundle<B> = 0>
int copy_new_to_old(B &old_field, const B &new_field, AMR<D> &amr, round_robin_threads &threads)
{
    for (int c = 0; c < B::components; c++)
        if (int rc = copy_new_to_old(old_field[c], new_field[c], amr, amr.sim_data.geom, threads.yield_stream()))
            return rc;
    return AM_OK;
}
template <class B, float_bundle<B> = 0>
int copy_to_child(B &field, uint32_t offset, int child_idx, int parent_idx, const am_geometry &g,
                  round_robin_threads &threads)
{
    for (int c = 0; c < B::components; c++)
        if (int rc = copy_to_child(field[c], offset, child_idx, parent_idx, g, threads.yield_stream())) return rc;
    return AM_OK;
}
// AMR.h:1304-1310 launches differentiate and mixed_partials on two streams of
// the pool; here they are one fused launch on the first of the two, with the
// seeds drawn in the same order, so the words are the same.
template <class D>
int yield_first_second_derivs(const smart_gpu_buffer<float> &field, smart_gpu_buffer<compressed_float3> &gradient,
                              comp_tensor2_sym_field &hessian, AMR<D> &amr, round_robin_threads &threads,
                              fast_prng &rng)
{
    void *stream = threads.yield_stream();
    threads.yield_stream();
    return yield_first_second_derivs(field, gradient, hessian.diag, hessian.off_diag, amr, amr.sim_data.geom, stream,
                                     rng);
}

}  // namespace am
