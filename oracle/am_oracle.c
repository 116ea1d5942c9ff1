/*
 * am_oracle.c -- CPU restatement of AdaptiveMesh's per-tile path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py load it.  The product path
 * (adaptivemesh_amd/csrc, the HIP kernels behind include/adaptivemesh.h) never
 * links, loads or calls anything here.
 *
 * Every function restates one reference routine with the same IEEE float32
 * operation order (compiled with -ffp-contract=off -fwrapv), so its output is
 * bit-exact with what the reference computes on an IEEE host:
 *
 *   fast_prng                  AdaptiveMesh/tensors.h:276-304
 *   compressed_float3 pack     AdaptiveMesh/tensors.h:350-393
 *   compressed_float3 unpack   AdaptiveMesh/tensors.h:335-348
 *   octree_abs_pos             AdaptiveMesh/AMR.h:22-55
 *   domain_boundary            AdaptiveMesh/AMR.h:56-100
 *   yield_sum/diff_coordinates AdaptiveMesh/AMR.h:102-115
 *   AMR::add_node/remove_node  AdaptiveMesh/AMR.h:134-146, 208-254
 *   AMR::find_node_idx         AdaptiveMesh/AMR.h:255-263
 *   AMR::yield_boundary        AdaptiveMesh/AMR.h:276-290
 *   AMR::increment_timer       AdaptiveMesh/AMR.h:180-188
 *   relative_time/active_depth AdaptiveMesh/AMR.h:634-644
 *   __interpolate_dat (both)   AdaptiveMesh/AMR.h:673-699, 767-785
 *   __copy_boundaries (both)   AdaptiveMesh/AMR.h:702-762, 789-848
 *   __copy_new_to_old          AdaptiveMesh/AMR.h:961-996
 *   __copy_to_child            AdaptiveMesh/AMR.h:1053-1073
 *   __differentiate (both)     AdaptiveMesh/AMR.h:1144-1219
 *   __mixed_partials           AdaptiveMesh/AMR.h:1221-1266
 *   __predictor/__corrector    AdaptiveMesh/SimpleWave.h:7-46
 *   __deriv_trace              AdaptiveMesh/BSSN.h:7-22
 *   __compute_raised_christoffel_symbols  AdaptiveMesh/BSSN.h:55-91
 *   __rasterise_to_image       AdaptiveMesh/AMR.h:1313-1353
 *   __init_temp                AdaptiveMesh/kernel.cu:9-43
 *   ___rgba(float4/float/float3) AdaptiveMesh/image_process.h:14-38
 *
 * Extensions that the reference does not have, restated here so the product's
 * versions can be checked: FD order decoupled from the padding (the
 * reference always uses order 2P), a periodic octree neighbour search, the
 * correct (inverse-metric) Christoffel mode, and tricubic/quadratic-time
 * ghost interpolation.
 *
 * Parity pins (tests/test_oracle_*.py): known answers for the codec and PRNG,
 * closed forms for the octree encodings and timer, exactness on polynomials
 * for every stencil and interpolant, exact-copy at equal depth, brute force
 * neighbour search on small trees.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

EXPORT int orc_num_threads(void) { return omp_get_max_threads(); }
EXPORT void orc_set_threads(int n) { omp_set_num_threads(n); }

/* ------------------------------------------------------------------ PRNG */

/* fast_prng::generate_int (tensors.h:281-285): wrapping add, arithmetic
 * shift, wrapping multiply on a signed 32-bit state. */
static inline int32_t prng_next(int32_t *s)
{
    int32_t v = (int32_t)((uint32_t)*s + 1124781675u);
    v ^= v >> 7;
    v = (int32_t)((uint32_t)v * 1246477263u);
    *s = v;
    return v;
}
/* fast_prng::generate_float (tensors.h:292-297) */
static inline float prng_float(int32_t *s)
{
    return (float)prng_next(s) * 2.3283064E-10f;
}
/* Per-thread seeding used by every dithering kernel (AMR.h:1160-1162):
 * seed = gs + x*2178614; seed ^= next()*y; seed ^= next()*z.  With C++17
 * sequencing the right-hand side runs first, so the xor sees the advanced
 * state. */
static inline int32_t thread_seed(int32_t gs, uint32_t x, uint32_t y, uint32_t z)
{
    int32_t s = (int32_t)((uint32_t)gs + x * 2178614u);
    uint32_t t = (uint32_t)prng_next(&s) * y;
    s = (int32_t)((uint32_t)s ^ t);
    t = (uint32_t)prng_next(&s) * z;
    s = (int32_t)((uint32_t)s ^ t);
    return s;
}

EXPORT int32_t orc_prng_int(int32_t *seed) { return prng_next(seed); }
EXPORT float orc_prng_float(int32_t *seed) { return prng_float(seed); }
EXPORT int32_t orc_thread_seed(int32_t gs, uint32_t x, uint32_t y, uint32_t z)
{
    return thread_seed(gs, x, y, z);
}

/* ------------------------------------------------- compressed_float3 codec */

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t umax(uint32_t a, uint32_t b) { return a > b ? a : b; }
static inline uint32_t umin(uint32_t a, uint32_t b) { return a < b ? a : b; }
static inline float fclampf(float f, float a, float b) { return fmaxf(a, fminf(f, b)); }

/* tensors.h:350-373 (rng != NULL) and :374-393 (rng == NULL) */
static uint32_t cf3_pack(float x, float y, float z, int32_t *rng)
{
    uint32_t e = umin(umax(umax((f2u(x) >> 23) & 255u,
                                umax((f2u(y) >> 23) & 255u, (f2u(z) >> 23) & 255u)), 112u) - 112u,
                      31u);
    const float sc = (32768u / (1.f / 128u)) / (float)(1u << e);
    x *= sc; y *= sc; z *= sc;
    if (rng) {
        x += prng_float(rng) * 0.95f;
        y += prng_float(rng) * 0.95f;
        z += prng_float(rng) * 0.95f;
    }
    x = fclampf(x, -255.f, 255.f);
    y = fclampf(y, -255.f, 255.f);
    z = fclampf(z, -255.f, 255.f);
    uint32_t w = e << 27;
    w |= (uint32_t)(fabsf(x) + .5f) | ((uint32_t)(x < 0.f) << 8);
    w |= ((uint32_t)(fabsf(y) + .5f) << 9) | ((uint32_t)(y < 0.f) << 17);
    w |= ((uint32_t)(fabsf(z) + .5f) << 18) | ((uint32_t)(z < 0.f) << 26);
    return w;
}
/* tensors.h:335-348 */
static inline void cf3_unpack(uint32_t w, float *o)
{
    const float base = u2f((127u - 15u + (w >> 27)) << 23);
    o[0] = base * ((float)(w & 255u) * (1.f - 2.f * (float)((w >> 8) & 1u)) * (1.f / 128u));
    o[1] = base * ((float)((w >> 9) & 255u) * (1.f - 2.f * (float)((w >> 17) & 1u)) * (1.f / 128u));
    o[2] = base * ((float)((w >> 18) & 255u) * (1.f - 2.f * (float)((w >> 26) & 1u)) * (1.f / 128u));
}

EXPORT uint32_t orc_cf3_pack(float x, float y, float z) { return cf3_pack(x, y, z, NULL); }
EXPORT uint32_t orc_cf3_pack_dither(float x, float y, float z, int32_t *seed)
{
    return cf3_pack(x, y, z, seed);
}
EXPORT void orc_cf3_unpack(uint32_t w, float *out3) { cf3_unpack(w, out3); }
EXPORT void orc_cf3_pack_n(const float *xyz, uint32_t *out, int64_t n)
{
    for (int64_t i = 0; i < n; i++) out[i] = cf3_pack(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], NULL);
}
EXPORT void orc_cf3_unpack_n(const uint32_t *w, float *out, int64_t n)
{
    for (int64_t i = 0; i < n; i++) cf3_unpack(w[i], out + 3 * i);
}

/* ----------------------------------------------------------- octree codes */

static inline int pos_depth(uint32_t d) { return (~d) ? (int)(d >> 27) : -1; }
static inline uint32_t pos_final_offset(uint32_t d)
{
    return (d >> ((uint32_t)(pos_depth(d) - 1) * 3u)) & 7u;
}
/* helper_math.h:2256-2262 */
static inline uint32_t reverse_bits(uint32_t a)
{
    a = (a >> 16) | (a << 16);
    a = ((a & 0xFF00FF00u) >> 8) | ((a & 0x00FF00FFu) << 8);
    a = ((a & 0xF0F0F0F0u) >> 4) | ((a & 0x0F0F0F0Fu) << 4);
    a = ((a & 0xCCCCCCCCu) >> 2) | ((a & 0x33333333u) << 2);
    return ((a & 0xAAAAAAAAu) >> 1) | ((a & 0x55555555u) << 1);
}
/* Per-lane Morton add/sub on the bit-reversed path (AMR.h:102-115): the
 * other two lanes are saturated with ones so a carry ripples along its own
 * lane; bits 27..29 catch a carry out of the root. */
#define LANE 19173961u   /* bits 0,3,...,24 */
#define FILL 115043766u  /* the two other lanes within 27 bits */
#define LANE_OV 153391689u
static uint32_t morton_sum(uint32_t a, uint32_t b, int overflow)
{
    uint32_t z = ((a | FILL) + (b & LANE)) & LANE_OV;
    uint32_t y = ((a | (FILL << 1)) + (b & (LANE << 1))) & (LANE_OV << 1);
    uint32_t x = ((a | (FILL << 2)) + (b & (LANE << 2))) & (LANE_OV << 2);
    x |= y | z;
    return ((x & (7u << 27)) && overflow) ? ~0u : (x & ((1u << 27) - 1u));
}
static uint32_t morton_diff(uint32_t a, uint32_t b, int overflow)
{
    uint32_t z = ((a & LANE) - (b & LANE)) & LANE_OV;
    uint32_t y = ((a & (LANE << 1)) - (b & (LANE << 1))) & (LANE_OV << 1);
    uint32_t x = ((a & (LANE << 2)) - (b & (LANE << 2))) & (LANE_OV << 2);
    x |= y | z;
    return ((x & (7u << 27)) && overflow) ? ~0u : (x & ((1u << 27) - 1u));
}
EXPORT uint32_t orc_morton_sum(uint32_t a, uint32_t b, int ov) { return morton_sum(a, b, ov); }
EXPORT uint32_t orc_morton_diff(uint32_t a, uint32_t b, int ov) { return morton_diff(a, b, ov); }
EXPORT uint32_t orc_reverse_bits(uint32_t a) { return reverse_bits(a); }

/* domain_boundary ctor (AMR.h:64-86).  Stored as node_data | rel_pos << 32. */
static uint64_t make_boundary(uint32_t this_d, uint32_t other_d, int other_idx, int periodic, int face)
{
    const int dt = pos_depth(this_d), dother = pos_depth(other_d);
    uint32_t nd = ((uint32_t)other_idx & 16777215u) | ((uint32_t)((dt - dother) & 15) << 24) |
                  ((uint32_t)(dt & 15) << 28);
    float r[3] = {0.f, 0.f, 0.f};
    const float mult = .25f * (float)(1u << dother);
    uint32_t s = this_d & 134217727u;
    for (int i = 0; i < dt; i++, s >>= 3) {
        r[0] += ((s & 1u) ? mult : -mult) / (float)(1 << i);
        r[1] += ((s & 2u) ? mult : -mult) / (float)(1 << i);
        r[2] += ((s & 4u) ? mult : -mult) / (float)(1 << i);
    }
    s = other_d & 134217727u;
    for (int i = 0; i < dother; i++, s >>= 3) {
        r[0] -= ((s & 1u) ? mult : -mult) / (float)(1 << i);
        r[1] -= ((s & 2u) ? mult : -mult) / (float)(1 << i);
        r[2] -= ((s & 4u) ? mult : -mult) / (float)(1 << i);
    }
    if (periodic && face >= 0) {
        /* the image of the target that lies across this face: along the face
         * axis this - target is negative for a + face and positive for a -
         * face; the unit cube is 2^dother target widths across */
        const float w = (float)(1u << dother);
        const int a = face % 3;
        if (face < 3 && r[a] > 0.f) r[a] -= w;
        if (face >= 3 && r[a] < 0.f) r[a] += w;
    }
    return (uint64_t)nd | ((uint64_t)cf3_pack(r[0], r[1], r[2], NULL) << 32);
}
#define BOUNDARY_NONE 0x00000000FFFFFFFFull

/* --------------------------------------------------------------- octree */

typedef struct {
    int max_slots;
    int periodic;
    int32_t *parent;     /* [max_slots] */
    int32_t *children;   /* [8*max_slots] */
    uint32_t *pos;       /* [max_slots] */
    uint64_t *bnd;       /* [6*max_slots] */
    uint32_t *lifetime;  /* [max_slots] */
    int max_depth;
    uint32_t first_free, final_used, timer;
} orc_tree;

EXPORT orc_tree *orc_tree_create(int max_slots, int periodic)
{
    orc_tree *t = (orc_tree *)calloc(1, sizeof(orc_tree));
    t->max_slots = max_slots;
    t->periodic = periodic;
    t->parent = (int32_t *)malloc(sizeof(int32_t) * max_slots);
    t->children = (int32_t *)malloc(sizeof(int32_t) * 8 * max_slots);
    t->pos = (uint32_t *)malloc(sizeof(uint32_t) * max_slots);
    t->bnd = (uint64_t *)malloc(sizeof(uint64_t) * 6 * max_slots);
    t->lifetime = (uint32_t *)calloc(max_slots, sizeof(uint32_t));
    memset(t->parent, 0xFF, sizeof(int32_t) * max_slots);
    memset(t->children, 0xFF, sizeof(int32_t) * 8 * max_slots);
    memset(t->pos, 0xFF, sizeof(uint32_t) * max_slots);
    /* the reference leaves unused boundary slots uninitialised; the oracle
     * gives them the "no boundary" value so comparisons are deterministic */
    for (int i = 0; i < 6 * max_slots; i++) t->bnd[i] = BOUNDARY_NONE;
    t->pos[0] = 0u;
    t->first_free = 1;
    t->final_used = 0;
    t->max_depth = 0;
    t->timer = 0;
    return t;
}
EXPORT void orc_tree_destroy(orc_tree *t)
{
    if (!t) return;
    free(t->parent); free(t->children); free(t->pos); free(t->bnd); free(t->lifetime);
    free(t);
}

/* AMR::find_node_idx (AMR.h:255-263) */
static uint32_t find_node_idx(const orc_tree *t, uint32_t packed, uint32_t maxd)
{
    uint32_t cur = 0, depth = 0;
    for (;;) {
        int c = t->children[cur * 8u + (packed & 7u)];
        if (c == -1 || depth >= maxd) return cur;
        packed >>= 3;
        cur = (uint32_t)c;
        depth++;
    }
}
EXPORT uint32_t orc_find_node_idx(const orc_tree *t, uint32_t packed, uint32_t maxd)
{
    return find_node_idx(t, packed, maxd);
}

/* AMR::yield_boundary (AMR.h:276-290); directions = {4,2,1} for x,y,z */
static uint64_t yield_boundary(const orc_tree *t, int node, uint32_t face)
{
    static const uint32_t dirs[3] = {4u, 2u, 1u};
    const int depth = pos_depth(t->pos[node]);
    if (depth <= 0) return BOUNDARY_NONE;
    uint32_t p = reverse_bits(t->pos[node]) >> 5;
    const int ov = !t->periodic;
    if (face < 3)
        p = morton_sum(p, (dirs[face] << ((9u - (uint32_t)depth) * 3u)) & ((1u << 27) - 1u), ov);
    else
        p = morton_diff(p, (dirs[face - 3] << ((9u - (uint32_t)depth) * 3u)) & ((1u << 27) - 1u), ov);
    int idx = (int)find_node_idx(t, reverse_bits(p) >> 5, (uint32_t)depth);
    if (p == ~0u) idx = 0;
    return make_boundary(t->pos[node], t->pos[idx], idx, t->periodic, (int)face);
}
EXPORT uint64_t orc_yield_boundary(const orc_tree *t, int node, uint32_t face)
{
    return yield_boundary(t, node, face);
}

/* AMR::add_node (AMR.h:208-232).  Returns the new slot, or -1 where the
 * reference silently returns. */
EXPORT int orc_add_node(orc_tree *t, int parent, int offset)
{
    if (t->first_free >= (uint32_t)t->max_slots || t->children[parent * 8 + offset] != -1) return -1;
    if (pos_depth(t->pos[parent]) >= 8) return -1;
    const uint32_t slot = t->first_free;
    t->lifetime[slot] = 0;
    t->parent[slot] = parent;
    t->children[parent * 8 + offset] = (int32_t)slot;
    /* octree_abs_pos::set_child (AMR.h:50-54) */
    t->pos[slot] = t->pos[parent] + (1u << 27);
    t->pos[slot] |= ((uint32_t)offset) << ((uint32_t)pos_depth(t->pos[parent]) * 3u);
    const int nd = pos_depth(t->pos[slot]);
    const int change = (t->max_depth > nd ? t->max_depth : nd) - t->max_depth;
    t->timer <<= change;
    t->max_depth += change;
    t->final_used = umax(t->final_used, slot);
    for (; t->first_free < (uint32_t)t->max_slots && pos_depth(t->pos[t->first_free]) != -1; t->first_free++) {}
    for (uint32_t i = 0; i <= t->final_used; i++)
        if (pos_depth(t->pos[i]) >= nd)
            for (uint32_t j = 0; j < 6; j++) t->bnd[i * 6 + j] = yield_boundary(t, (int)i, j);
    return (int)slot;
}

static void remove_rec(orc_tree *t, int node)
{
    if (node == -1 || pos_depth(t->pos[node]) == -1) return;
    for (int i = 0; i < 8; i++) remove_rec(t, t->children[node * 8 + i]);
    t->children[t->parent[node] * 8 + (int)pos_final_offset(t->pos[node])] = -1;
    t->parent[node] = -1;
    t->pos[node] = ~0u;
    t->first_free = umin(t->first_free, (uint32_t)node);
}
/* AMR::remove_node (AMR.h:233-254) */
EXPORT int orc_remove_node(orc_tree *t, int node)
{
    if (node == -1) return -1;
    const int cd = pos_depth(t->pos[node]);
    if (cd == -1) return -1;
    remove_rec(t, node);
    const int old = t->max_depth;
    t->max_depth = 0;
    for (int i = 0; i < t->max_slots; i++)
        if (i != node && pos_depth(t->pos[i]) > t->max_depth) t->max_depth = pos_depth(t->pos[i]);
    t->timer >>= (old - t->max_depth);
    for (; pos_depth(t->pos[t->final_used]) == -1 && t->final_used > 0; t->final_used--) {}
    for (uint32_t i = 0; i <= t->final_used; i++)
        if (pos_depth(t->pos[i]) >= cd)
            for (uint32_t j = 0; j < 6; j++) t->bnd[i * 6 + j] = yield_boundary(t, (int)i, j);
    return 0;
}

/* AMR::increment_timer (AMR.h:180-188) */
EXPORT void orc_increment_timer(orc_tree *t)
{
    for (uint32_t i = 0; i <= t->final_used; i++) {
        const int d = pos_depth(t->pos[i]);
        if (d == -1) continue;
        t->lifetime[i] += ((t->timer & ((1u << (t->max_depth - d)) - 1u)) == 0u);
    }
    t->timer = (t->timer + 1u) & ((1u << t->max_depth) - 1u);
}

EXPORT void orc_tree_export(const orc_tree *t, int32_t *parent, int32_t *children, uint32_t *pos,
                            uint64_t *bnd, uint32_t *lifetime, int32_t *scalars)
{
    const int n = t->max_slots;
    if (parent) memcpy(parent, t->parent, sizeof(int32_t) * n);
    if (children) memcpy(children, t->children, sizeof(int32_t) * 8 * n);
    if (pos) memcpy(pos, t->pos, sizeof(uint32_t) * n);
    if (bnd) memcpy(bnd, t->bnd, sizeof(uint64_t) * 6 * n);
    if (lifetime) memcpy(lifetime, t->lifetime, sizeof(uint32_t) * n);
    if (scalars) {
        scalars[0] = t->max_depth;
        scalars[1] = (int32_t)t->first_free;
        scalars[2] = (int32_t)t->final_used;
        scalars[3] = (int32_t)t->timer;
    }
}
EXPORT void orc_tree_set_timer(orc_tree *t, uint32_t timer) { t->timer = timer; }

/* relative_time / active_depth (AMR.h:634-644) */
static inline int active_depth(uint32_t sub, uint32_t depth, uint32_t maxd)
{
    return (sub & ((1u << (maxd - depth)) - 1u)) == 0u;
}
static inline int bnd_target(uint32_t nd)
{
    return (int)((nd & 16777215u) | ((nd & (1u << 23)) ? 4278190080u : 0u));
}
static inline uint32_t bnd_ddiff(uint32_t nd) { return (nd >> 24) & 15u; }
static inline int bnd_depth(uint32_t nd) { return ((nd >> 28) == 15u) ? -1 : (int)(nd >> 28); }

/* ------------------------------------------------------------ geometry */

typedef struct {
    int S, P, T, q;       /* interior, padding, padded side, FD half-width */
    float outer_dx, outer_dt;
} geom;

static geom mkgeom(int S, int P, int q, float outer_dx, float outer_dt)
{
    geom g = {S, P, S + 2 * P, q, outer_dx, outer_dt};
    return g;
}

/* FD tables (AMR.h:1144-1150); order-2q stencil starts at index q*q+q-1 */
static const float CFD1[15] = {-.5f, 0.f, .5f, 1.f / 12, -2.f / 3, 0.f, 2.f / 3, -1.f / 12,
                               -1.f / 60, 3.f / 20, -3.f / 4, 0.f, 3.f / 4, -3.f / 20, 1.f / 60};
static const float CFD2[15] = {1.f, -2.f, 1.f, -1.f / 12, 4.f / 3, -5.f / 2, 4.f / 3, -1.f / 12,
                               1.f / 90, -3.f / 20, 3.f / 2, -49.f / 18, 3.f / 2, -3.f / 20, 1.f / 90};

/* -------------------------------------------------------- FD stencils */

/* __differentiate, both overloads (AMR.h:1152-1219).  d2 == NULL selects the
 * first-derivative-only kernel.  Output words are interior-only,
 * slot*S^3 + (z*S+y)*S+x. */
/* One derivative triple into a word array: fmt 0 the dithered
 * compressed_float3 word, fmt 1 the three fp32 values themselves (3 floats
 * per cell, the precision study's fp32 words; its fp16 words are these
 * values rounded to nearest even). */
static inline void store3(void *out, int64_t o, int fmt, float x, float y, float z, int32_t *rng)
{
    if (fmt == 0) {
        ((uint32_t *)out)[o] = cf3_pack(x, y, z, rng);
    } else {
        float *p = (float *)out + 3 * o;
        p[0] = x; p[1] = y; p[2] = z;
    }
}

static void differentiate(const float *field, void *d1, void *d2, const uint32_t *pos, const int32_t *tiles,
                          int n_tiles, uint32_t substep, uint32_t maxd, int32_t gs, int S, int P, int q,
                          float outer_dx, int fmt)
{
    const geom g = mkgeom(S, P, q, outer_dx, 0.f);
    const int T = g.T, base = q * q + q - 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int z = 0; z < S; z++) {
            const int slot = tiles[ti];
            const int depth = pos_depth(pos[slot]);
            if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) continue;
            const float dx1 = (float)(1u << depth) / outer_dx;
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    int32_t rng = thread_seed(gs, (uint32_t)x, (uint32_t)y, (uint32_t)(slot * S + z));
                    const int64_t rw = (((int64_t)slot * T + z) * T + y) * T + x + P * (1 + T + T * T);
                    float a, vx = 0.f, vy = 0.f, vz = 0.f, xx = 0.f, yy = 0.f, zz = 0.f;
                    for (int i = -q; i <= q; i++) {
                        a = field[rw + i];
                        vx += a * CFD1[base + i];
                        xx += a * CFD2[base + i];
                        a = field[rw + (int64_t)i * T];
                        vy += a * CFD1[base + i];
                        yy += a * CFD2[base + i];
                        a = field[rw + (int64_t)i * T * T];
                        vz += a * CFD1[base + i];
                        zz += a * CFD2[base + i];
                    }
                    vx *= dx1; vy *= dx1; vz *= dx1;
                    const int64_t o = (((int64_t)slot * S + z) * S + y) * S + x;
                    if (d2) {
                        xx *= dx1 * dx1; yy *= dx1 * dx1; zz *= dx1 * dx1;
                        store3(d2, o, fmt, xx, yy, zz, &rng);
                    }
                    store3(d1, o, fmt, vx, vy, vz, &rng);
                }
        }
}

EXPORT void orc_differentiate(const float *field, uint32_t *d1, uint32_t *d2, const uint32_t *pos,
                              const int32_t *tiles, int n_tiles, uint32_t substep, uint32_t maxd,
                              int32_t gs, int S, int P, int q, float outer_dx)
{
    differentiate(field, d1, d2, pos, tiles, n_tiles, substep, maxd, gs, S, P, q, outer_dx, 0);
}
/* the same derivatives as fp32 triples (d1, d2: 3 floats per cell) */
EXPORT void orc_differentiate_f32(const float *field, float *d1, float *d2, const uint32_t *pos,
                                  const int32_t *tiles, int n_tiles, uint32_t substep, uint32_t maxd, int S, int P,
                                  int q, float outer_dx)
{
    differentiate(field, d1, d2, pos, tiles, n_tiles, substep, maxd, 0, S, P, q, outer_dx, 1);
}

/* __mixed_partial_c / _bd (AMR.h:1221-1230) */
static inline float mixed_c(const float *f, int64_t i, int64_t dx, int64_t dy)
{
    return .25f * (f[i + dx + dy] + f[i - dx - dy] - f[i + dx - dy] - f[i - dx + dy]);
}
static inline float mixed_bd(const float *f, int64_t i, int64_t dx, int64_t dy)
{
    return f[i] + .5f * (f[i + dx + dy] + f[i - dx - dy] - f[i - dx] - f[i + dx] - f[i - dy] - f[i + dy]);
}
static inline int iclamp(int v, int a, int b) { return v < a ? a : (v > b ? b : v); }

/* __mixed_partials (AMR.h:1232-1266) */
static void mixed_partials(const float *field, void *off, const uint32_t *pos, const int32_t *tiles, int n_tiles,
                           uint32_t substep, uint32_t maxd, int32_t gs, int S, int P, float outer_dx, int fmt)
{
    const int T = S + 2 * P;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int z = 0; z < S; z++) {
            const int slot = tiles[ti];
            const int depth = pos_depth(pos[slot]);
            if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) continue;
            float dx1 = (float)(1u << depth) / outer_dx;
            dx1 *= dx1;
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    int32_t rng = thread_seed(gs, (uint32_t)x, (uint32_t)y, (uint32_t)(slot * S + z));
                    const int64_t rw = (((int64_t)slot * T + z) * T + y) * T + x + P * (1 + T + T * T);
                    const int ex = x - iclamp(x, 1, S - 2);
                    const int ey = y - iclamp(y, 1, S - 2);
                    const int ez = z - iclamp(z, 1, S - 2);
                    const int cxy = ex * ey, cxz = ex * ez, cyz = ey * ez;
                    const int64_t T2 = (int64_t)T * T;
                    const float vxy = cxy == 0 ? mixed_c(field, rw, 1, T)
                                               : ((float)(-cxy) * mixed_bd(field, rw, -cxy, T));
                    const float vxz = cxz == 0 ? mixed_c(field, rw, 1, T2)
                                               : ((float)(-cxz) * mixed_bd(field, rw, -cxz, T2));
                    const float vyz = cyz == 0 ? mixed_c(field, rw, T, T2)
                                               : ((float)(-cyz) * mixed_bd(field, rw, (int64_t)(-cyz) * T, T2));
                    const int64_t o = (((int64_t)slot * S + z) * S + y) * S + x;
                    store3(off, o, fmt, vxy * dx1, vxz * dx1, vyz * dx1, &rng);
                }
        }
}

EXPORT void orc_mixed_partials(const float *field, uint32_t *off, const uint32_t *pos,
                               const int32_t *tiles, int n_tiles, uint32_t substep, uint32_t maxd,
                               int32_t gs, int S, int P, float outer_dx)
{
    mixed_partials(field, off, pos, tiles, n_tiles, substep, maxd, gs, S, P, outer_dx, 0);
}
/* The FD's advection + dissipation output (an extension for the upwinded,
 * dissipative BSSN runs; no reference counterpart), in float with no
 * contraction, as the product computes it:
 *   L_k = (sum over the five offsets m of the lopsided stencil of f_m * c_m,
 *          in increasing m) * inv_dx, on the side beta^k points to
 *          (c = (-3, -10, 18, -6, 1) / 12 at m = -1..3 for beta^k > 0, else
 *          (-1, 6, -18, 10, 3) / 12 at m = -3..1, each rounded to float)
 *   K_k = sum over m = -3..3 of f_m * (1, -6, 15, -20, 15, -6, 1)_m
 *   out = (advect ? (beta^x L_x + beta^y L_y) + beta^z L_z : 0)
 *         + ((K_x + K_y) + K_z) * ((sigma * (1/64)) * inv_dx)
 * with beta^k the shift fields at the cell and inv_dx = 2^depth /
 * outer_dx as float (AMR.h:1214).  out: interior-only layout, as the words. */
static const float LOPF_P[5] = {-3.f / 12, -10.f / 12, 18.f / 12, -6.f / 12, 1.f / 12};
static const float LOPF_M[5] = {-1.f / 12, 6.f / 12, -18.f / 12, 10.f / 12, 3.f / 12};
static const float KO6F[7] = {1.f, -6.f, 15.f, -20.f, 15.f, -6.f, 1.f};
EXPORT void orc_fd_adv(const float *field, const float *bx, const float *by, const float *bz, float *out,
                       const uint32_t *pos, const int32_t *tiles, int n_tiles, uint32_t substep, uint32_t maxd,
                       int S, int P, float outer_dx, float sigma, int advect)
{
    const int T = S + 2 * P;
    const int64_t st[3] = {1, T, (int64_t)T * T};
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int z = 0; z < S; z++) {
            const int slot = tiles[ti];
            const int depth = pos_depth(pos[slot]);
            if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) continue;
            const float dx1 = (float)(1u << depth) / outer_dx;
            const float kos = (sigma * (1.f / 64.f)) * dx1;
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    const int64_t rw = (((int64_t)slot * T + z) * T + y) * T + x + P * (1 + T + T * T);
                    const float beta[3] = {bx[rw], by[rw], bz[rw]};
                    float adv = 0.f, ko = 0.f;
                    for (int k = 0; k < 3; k++) {
                        float l = 0.f, kk = 0.f;
                        if (beta[k] > 0.f)
                            for (int m = -1; m <= 3; m++) l = l + field[rw + m * st[k]] * LOPF_P[m + 1];
                        else
                            for (int m = -3; m <= 1; m++) l = l + field[rw + m * st[k]] * LOPF_M[m + 3];
                        for (int m = -3; m <= 3; m++) kk = kk + field[rw + m * st[k]] * KO6F[m + 3];
                        adv = adv + beta[k] * (l * dx1);
                        ko = ko + kk;
                    }
                    out[(((int64_t)slot * S + z) * S + y) * S + x] = (advect ? adv : 0.f) + ko * kos;
                }
        }
}

EXPORT void orc_mixed_partials_f32(const float *field, float *off, const uint32_t *pos, const int32_t *tiles,
                                   int n_tiles, uint32_t substep, uint32_t maxd, int S, int P, float outer_dx)
{
    mixed_partials(field, off, pos, tiles, n_tiles, substep, maxd, 0, S, P, outer_dx, 1);
}

/* ------------------------------------------------ interpolation kernels */

/* __interpolate_dat, quadratic 10-point form (AMR.h:673-699) */
static inline float interp_lagrange(const float *d, const float *rp, int64_t n, int T)
{
    const int64_t T2 = (int64_t)T * T;
    float f = d[n], fx = d[n + 1], fy = d[n + T], fz = d[n + T2];
    float fxx = d[n - 1], fyy = d[n - T], fzz = d[n - T2];
    float fxy = d[n + T + 1], fxz = d[n + T2 + 1], fyz = d[n + T2 + T];
    fxy += f - fx - fy;
    fxz += f - fx - fz;
    fyz += f - fy - fz;
    fx = (fx - fxx) * .5f;
    fy = (fy - fyy) * .5f;
    fz = (fz - fzz) * .5f;
    fxx = fx + fxx - f;
    fyy = fy + fyy - f;
    fzz = fz + fzz - f;
    return f + (fx + fxx * rp[0]) * rp[0] + (fy + fyy * rp[1]) * rp[1] + (fz + fzz * rp[2]) * rp[2] +
           fxy * rp[0] * rp[1] + fxz * rp[0] * rp[2] + fyz * rp[1] * rp[2];
}
/* __interpolate_dat, trilinear form (AMR.h:767-785) */
static inline float interp_trilinear(const float *d, const float *rp, int64_t n, int T)
{
    const int64_t T2 = (int64_t)T * T;
    float r00 = d[n];
    r00 = r00 * (1.f - rp[0]) + d[n + 1] * rp[0];
    float r01 = d[n + T];
    r01 = r01 * (1.f - rp[0]) + d[n + T + 1] * rp[0];
    r00 = r00 * (1.f - rp[1]) + r01 * rp[1];
    float r10 = d[n + T2];
    r10 = r10 * (1.f - rp[0]) + d[n + T2 + 1] * rp[0];
    float r11 = d[n + T2 + T];
    r11 = r11 * (1.f - rp[0]) + d[n + T2 + T + 1] * rp[0];
    r10 = r10 * (1.f - rp[1]) + r11 * rp[1];
    return r00 * (1.f - rp[2]) + r10 * rp[2];
}

/* 4-point Lagrange weights on nodes -1,0,1,2 at offset t (tricubic mode, an
 * extension: exact on cubics in each axis). */
static inline void lag4(float t, float *w)
{
    w[0] = -t * (t - 1.f) * (t - 2.f) / 6.f;
    w[1] = (t + 1.f) * (t - 1.f) * (t - 2.f) / 2.f;
    w[2] = -(t + 1.f) * t * (t - 2.f) / 2.f;
    w[3] = (t + 1.f) * t * (t - 1.f) / 6.f;
}
static inline float interp_tricubic(const float *d, const float *rp, int64_t n, int T)
{
    const int64_t T2 = (int64_t)T * T;
    float wx[4], wy[4], wz[4];
    lag4(rp[0], wx); lag4(rp[1], wy); lag4(rp[2], wz);
    float acc = 0.f;
    for (int k = 0; k < 4; k++) {
        float ay = 0.f;
        for (int j = 0; j < 4; j++) {
            const float *r = d + n + (int64_t)(k - 1) * T2 + (int64_t)(j - 1) * T - 1;
            const float ax = r[0] * wx[0] + r[1] * wx[1] + r[2] * wx[2] + r[3] * wx[3];
            ay += ax * wy[j];
        }
        acc += ay * wz[k];
    }
    return acc;
}

enum { INTERP_LAGRANGE = 0, INTERP_TRILINEAR = 1, INTERP_TRICUBIC = 2 };

/* __copy_boundaries (AMR.h:702-762 quadratic, 789-848 trilinear), plus the
 * tricubic-space / quadratic-time extension.  For the tricubic mode the time
 * interpolation is quadratic through (older, old, new) at times (-1, 0, 1)
 * of the coarse step; older may be NULL, which selects linear time. */
EXPORT void orc_copy_boundaries(float *old_, float *new_, const float *older, const uint64_t *bnd,
                                const int32_t *tiles, int n_tiles, uint32_t substep, uint32_t maxd,
                                int copy_only_new, int mode, int S, int P, float frac, int equal_copy)
{
    const int T = S + 2 * P;
    const int64_t cells = (int64_t)T * T * T;
    const float invS = 1.f / (float)S;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int dir = 0; dir < 6; dir++) {
            const int node = tiles[ti];
            const uint64_t bb = bnd[(int64_t)node * 6 + dir];
            const uint32_t nd = (uint32_t)bb;
            const int tgt = bnd_target(nd), cd = bnd_depth(nd);
            if (tgt == -1 || cd <= 0) continue;
            if (!active_depth(substep, (uint32_t)cd, maxd)) continue;
            float relp[3];
            cf3_unpack((uint32_t)(bb >> 32), relp);
            const uint32_t ddiff = bnd_ddiff(nd);
            const float rel_scale = 1.f / (float)(1u << ddiff);
            /* relative_time */
            const float tmp = (float)((substep >> (maxd - (uint32_t)cd)) & ((1u << ddiff) - 1u));
            /* frac = 1 is the reference (k + 1); an intermediate Runge-Kutta time
             * uses k + frac, and an equal-depth target holds the same stage */
            const float tx = tmp / (float)(1u << ddiff),
                        ty = ddiff == 0 ? 1.f : (tmp + frac) / (float)(1u << ddiff);
            for (int l = 0; l < P; l++)
                for (int b = 0; b < S; b++)
                    for (int a = 0; a < S; a++) {
                        uint32_t w[3];
                        const uint32_t ix = (uint32_t)a, iy = (uint32_t)b, iz = (uint32_t)l;
                        const uint32_t uP = (uint32_t)P, uS = (uint32_t)S;
                        switch (dir) {
                        case 0: w[0] = iz + uP + uS; w[1] = ix + uP; w[2] = iy + uP; break;
                        case 1: w[0] = ix + uP; w[1] = iz + uP + uS; w[2] = iy + uP; break;
                        case 2: w[0] = ix + uP; w[1] = iy + uP; w[2] = iz + uP + uS; break;
                        case 3: w[0] = iz; w[1] = ix + uP; w[2] = iy + uP; break;
                        case 4: w[0] = ix + uP; w[1] = iz; w[2] = iy + uP; break;
                        default: w[0] = ix + uP; w[1] = iy + uP; w[2] = iz; break;
                        }
                        float rp[3];
                        uint32_t ri[3];
                        for (int k = 0; k < 3; k++) {
                            rp[k] = ((float)w[k] - (float)uP + .5f) * invS - .5f;
                            rp[k] = rp[k] * rel_scale + relp[k];
                            rp[k] = (rp[k] + .5f) * (float)uS + (float)uP - 0.5f;
                        }
                        if (equal_copy && ddiff == 0) {
                            /* the RK4 driver's equal-depth faces: an exact copy of the
                             * neighbour cell the ghost cell maps onto (an extension; the
                             * reference clamps and extrapolates on the last row).  rp is
                             * a whole index within the rounding of 1/S, so it is rounded */
                            const int64_t src = (int64_t)tgt * cells + (int64_t)(rp[0] + .5f) +
                                                ((int64_t)(rp[1] + .5f) + (int64_t)(rp[2] + .5f) * T) * T;
                            const int64_t wn = (int64_t)node * cells + w[0] + ((int64_t)w[1] + (int64_t)w[2] * T) * T;
                            if (!copy_only_new) old_[wn] = old_[src];
                            new_[wn] = new_[src];
                            continue;
                        }
                        if (mode == INTERP_TRILINEAR) {
                            for (int k = 0; k < 3; k++)
                                ri[k] = (tgt == 0) ? umin((uint32_t)fmaxf(rp[k], .5f), (uint32_t)T - 2u)
                                                   : umin((uint32_t)fmaxf(rp[k], .5f + (float)uP),
                                                          (uint32_t)T - uP - 2u);
                        } else if (mode == INTERP_TRICUBIC) {
                            /* floor, so the 4 nodes ri-1..ri+2 bracket the point; a
                             * non-root target is read on its interior only, so no ghost
                             * cell that another tile of this launch writes is read */
                            for (int k = 0; k < 3; k++)
                                ri[k] = (tgt == 0) ? umin((uint32_t)fmaxf(rp[k], 1.f), (uint32_t)T - 3u)
                                                   : umin((uint32_t)fmaxf(rp[k], (float)(uP + 1u)), (uint32_t)T - uP - 3u);
                        } else {
                            for (int k = 0; k < 3; k++)
                                ri[k] = (tgt == 0) ? umin((uint32_t)(fmaxf(rp[k], 0) + .5f), (uint32_t)T - 2u)
                                                   : umin((uint32_t)(fmaxf(rp[k], (float)uP) + .5f),
                                                          (uint32_t)T - uP - 2u);
                        }
                        const int64_t rn = (int64_t)tgt * cells + ri[0] + ((int64_t)ri[1] + (int64_t)ri[2] * T) * T;
                        for (int k = 0; k < 3; k++) rp[k] -= (float)ri[k];
                        const int64_t wn = (int64_t)node * cells + w[0] + ((int64_t)w[1] + (int64_t)w[2] * T) * T;
                        if (mode == INTERP_TRICUBIC) {
                            const float o = interp_tricubic(old_, rp, rn, T);
                            const float nw = interp_tricubic(new_, rp, rn, T);
                            if (older && ddiff > 0) {
                                const float od = interp_tricubic(older, rp, rn, T);
                                /* quadratic through t=-1,0,1 evaluated at s */
                                float s = ty;
                                float c0 = .5f * s * (s - 1.f), c1 = (1.f - s) * (1.f + s), c2 = .5f * s * (s + 1.f);
                                new_[wn] = od * c0 + o * c1 + nw * c2;
                                if (!copy_only_new) {
                                    s = tx;
                                    c0 = .5f * s * (s - 1.f); c1 = (1.f - s) * (1.f + s); c2 = .5f * s * (s + 1.f);
                                    old_[wn] = od * c0 + o * c1 + nw * c2;
                                }
                            } else {
                                if (!copy_only_new) old_[wn] = o * (1.f - tx) + nw * tx;
                                new_[wn] = o * (1.f - ty) + nw * ty;
                            }
                        } else {
                            float o, nw;
                            if (mode == INTERP_TRILINEAR) {
                                o = interp_trilinear(old_, rp, rn, T);
                                nw = interp_trilinear(new_, rp, rn, T);
                            } else {
                                o = interp_lagrange(old_, rp, rn, T);
                                nw = interp_lagrange(new_, rp, rn, T);
                            }
                            if (!copy_only_new) old_[wn] = o * (1.f - tx) + nw * tx;
                            new_[wn] = o * (1.f - ty) + nw * ty;
                        }
                    }
        }
}

/* __copy_new_to_old (AMR.h:961-996): old <- new on every cell of an active
 * tile; interior cells covered by a child take the mean of its 2x2x2 cells. */
EXPORT void orc_copy_new_to_old(float *old_, const float *new_, const uint32_t *pos, const int32_t *children,
                                const int32_t *tiles, int n_tiles, uint32_t substep, uint32_t maxd, int S, int P)
{
    const int T = S + 2 * P;
    const int64_t cells = (int64_t)T * T * T, T2 = (int64_t)T * T;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int z = 0; z < T; z++) {
            const int node = tiles[ti];
            const int depth = pos_depth(pos[node]);
            if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) continue;
            for (int y = 0; y < T; y++)
                for (int x = 0; x < T; x++) {
                    const uint32_t h = (uint32_t)S / 2u;
                    const uint32_t tx = ((uint32_t)x - (uint32_t)P) / h, ty = ((uint32_t)y - (uint32_t)P) / h,
                                   tz = ((uint32_t)z - (uint32_t)P) / h;
                    const int child = (tx < 2u && ty < 2u && tz < 2u)
                                          ? children[tx + ty * 2u + tz * 4u + (uint32_t)node * 8u] : -1;
                    const int64_t o = (int64_t)node * cells + ((int64_t)z * T + y) * T + x;
                    if (child == -1) { old_[o] = new_[o]; continue; }
                    const int64_t cx = (((uint32_t)x - (uint32_t)P) * 2u) % (uint32_t)S + (uint32_t)P;
                    const int64_t cy = (((uint32_t)y - (uint32_t)P) * 2u) % (uint32_t)S + (uint32_t)P;
                    const int64_t cz = (((uint32_t)z - (uint32_t)P) * 2u) % (uint32_t)S + (uint32_t)P;
                    const int64_t c = (int64_t)child * cells + (cz * T + cy) * T + cx;
                    old_[o] = (new_[c] + new_[c + 1] + new_[c + T] + new_[c + T + 1] + new_[c + T2] +
                               new_[c + T2 + 1] + new_[c + T2 + T] + new_[c + T2 + T + 1]) * .125f;
                }
        }
}

/* __copy_to_child (AMR.h:1053-1073): prolong the parent into every cell of
 * a new child, quadratic (mode 0) or trilinear (mode 1). */
EXPORT void orc_copy_to_child(float *buf, uint32_t offset, int child, int parent, int mode, int S, int P)
{
    const int T = S + 2 * P;
    const int64_t cells = (int64_t)T * T * T;
    const float invS = 1.f / (float)S;
    const float off[3] = {(float)(offset & 1u), (float)((offset & 2u) >> 1), (float)((offset & 4u) >> 2)};
    for (int z = 0; z < T; z++)
        for (int y = 0; y < T; y++)
            for (int x = 0; x < T; x++) {
                const uint32_t idx[3] = {(uint32_t)x, (uint32_t)y, (uint32_t)z};
                float rp[3];
                uint32_t ri[3];
                for (int k = 0; k < 3; k++) {
                    rp[k] = ((float)idx[k] - (float)(uint32_t)P + .5f) * invS - .5f;
                    rp[k] = rp[k] * .5f + (off[k] - .5f) * .5f;
                    rp[k] = (rp[k] + .5f) * (float)(uint32_t)S + (float)(uint32_t)P - .5f;
                    ri[k] = (mode == INTERP_TRILINEAR) ? (uint32_t)rp[k] : (uint32_t)(rp[k] + .5f);
                }
                const int64_t rn = (int64_t)parent * cells + ri[0] + ((int64_t)ri[1] + (int64_t)ri[2] * T) * T;
                for (int k = 0; k < 3; k++) rp[k] -= (float)ri[k];
                const float v = (mode == INTERP_TRILINEAR) ? interp_trilinear(buf, rp, rn, T)
                                                           : interp_lagrange(buf, rp, rn, T);
                buf[(int64_t)child * cells + x + ((int64_t)y + (int64_t)z * T) * T] = v;
            }
}

/* --------------------------------------------------------- SimpleWave */

/* __predictor_step / __corrector_step (SimpleWave.h:7-46), viscosity 0.01 */
EXPORT void orc_wave_stage(const float *old_, const uint32_t *d1, const uint32_t *d2, float *new_,
                           const uint32_t *pos, const int32_t *tiles, int n_tiles, uint32_t substep,
                           uint32_t maxd, int corrector, int S, int P, float outer_dt)
{
    const int T = S + 2 * P;
    const float nu = 0.01f;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int z = 0; z < S; z++) {
            const int node = tiles[ti];
            const int depth = pos_depth(pos[node]);
            if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) continue;
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    const int64_t i = (((int64_t)node * S + z) * S + y) * S + x;
                    float g[3], h[3];
                    cf3_unpack(d1[i], g);
                    cf3_unpack(d2[i], h);
                    const float lap = h[0] * nu + h[1] * nu + h[2] * nu;
                    const int64_t r = (((int64_t)node * T + z) * T + y) * T + x + P * (1 + T + T * T);
                    float u = old_[r];
                    if (!corrector) {
                        new_[r] = u + (lap - u * g[0]) * outer_dt / (float)(1u << depth);
                    } else {
                        u += (lap - u * g[0]) * outer_dt / (float)(1u << depth);
                        new_[r] = (new_[r] + u) * .5f;
                    }
                }
        }
}

/* -------------------------------------------------------------- BSSN.h */

/* float3x3_sym::inverse (tensors.h:208-225) */
static inline void sym_inverse(const float *m, float *r)
{
    const float xx = m[0], xy = m[1], xz = m[2], yy = m[3], yz = m[4], zz = m[5];
    const float det = -xz * yy * xz + xy * yz * xz + xz * xy * yz - xx * yz * yz - xy * xy * zz + xx * yy * zz;
    r[0] = -yz * yz + yy * zz;
    r[1] = yz * xz - xy * zz;
    r[2] = -yy * xz + xy * yz;
    r[3] = -xz * xz + xx * zz;
    r[4] = xy * xz - xx * yz;
    r[5] = -xy * xy + xx * yy;
    for (int k = 0; k < 6; k++) r[k] /= det;
}

/* __deriv_trace (BSSN.h:7-22): g^{ij} H_ij with H from (diag, off) words.
 * metric[6] are the six component arrays xx,xy,xz,yy,yz,zz (padded tiles). */
EXPORT void orc_deriv_trace(const uint32_t *hdiag, const uint32_t *hoff, const float *const *metric,
                            float *out, const uint32_t *pos, const int32_t *tiles, int n_tiles,
                            uint32_t substep, uint32_t maxd, int S, int P)
{
    const int T = S + 2 * P;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int z = 0; z < S; z++) {
            const int node = tiles[ti];
            const int depth = pos_depth(pos[node]);
            if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) continue;
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    const int64_t r = (((int64_t)node * T + z) * T + y) * T + x + P * (1 + T + T * T);
                    float m[6], inv[6], d[3], o[3];
                    for (int k = 0; k < 6; k++) m[k] = metric[k][r];
                    sym_inverse(m, inv);
                    const int64_t i = (((int64_t)node * S + z) * S + y) * S + x;
                    cf3_unpack(hdiag[i], d);
                    cf3_unpack(hoff[i], o);
                    /* trace_contra(sym, sym) (tensors.h:241-244) */
                    out[i] = d[0] * inv[0] + d[1] * inv[3] + d[2] * inv[5] +
                             (o[0] * inv[1] + o[1] * inv[2] + o[2] * inv[4]) * 2.f;
                }
        }
}

/* __compute_raised_christoffel_symbols (BSSN.h:55-91).  dg[6] and out[6] are
 * word arrays ordered xx,xy,xz,yy,yz,zz.  compat=1 reproduces the reference,
 * which contracts with the metric itself (BSSN.h:82-83); compat=0 uses the
 * inverse metric, giving the true Gamma^l_jk. */
EXPORT void orc_christoffel(const float *const *metric, const uint32_t *const *dg, uint32_t *const *out,
                            int compat, const uint32_t *pos, const int32_t *tiles, int n_tiles,
                            uint32_t substep, uint32_t maxd, int32_t gs, int S, int P)
{
    const int T = S + 2 * P;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int z = 0; z < S; z++) {
            const int node = tiles[ti];
            const int depth = pos_depth(pos[node]);
            if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) continue;
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    int32_t rng = thread_seed(gs, (uint32_t)x, (uint32_t)y, (uint32_t)(node * S + z));
                    const int64_t i = (((int64_t)node * S + z) * S + y) * S + x;
                    float v[6][3];
                    for (int k = 0; k < 6; k++) cf3_unpack(dg[k][i], v[k]);
                    /* float3x3_sym(float3 x, y, z) (tensors.h:138) for the rows
                     * (dg_xa), (dg_ya), (dg_za) */
                    const float *rx[3] = {v[0], v[1], v[2]};
                    const float *ry[3] = {v[1], v[3], v[4]};
                    const float *rz[3] = {v[2], v[4], v[5]};
                    float X[3][6];
                    const float *const *rows[3] = {rx, ry, rz};
                    for (int l = 0; l < 3; l++) {
                        const float *a = rows[l][0], *b = rows[l][1], *c = rows[l][2];
                        X[l][0] = a[0];
                        X[l][1] = (a[1] + b[0]) * .5f;
                        X[l][2] = (a[2] + c[0]) * .5f;
                        X[l][3] = b[1];
                        X[l][4] = (c[1] + b[2]) * .5f;
                        X[l][5] = c[2];
                    }
                    const int64_t r = (((int64_t)node * T + z) * T + y) * T + x + P * (1 + T + T * T);
                    float m[6], im[6];
                    for (int k = 0; k < 6; k++) m[k] = metric[k][r];
                    if (compat) memcpy(im, m, sizeof m);
                    else sym_inverse(m, im);
                    for (int jk = 0; jk < 6; jk++) {
                        float a0 = X[0][jk] - v[jk][0] * .5f;
                        float a1 = X[1][jk] - v[jk][1] * .5f;
                        float a2 = X[2][jk] - v[jk][2] * .5f;
                        /* float3x3_sym * float3 (tensors.h:201-206) */
                        const float o0 = a0 * im[0] + a1 * im[1] + a2 * im[2];
                        const float o1 = a0 * im[1] + a1 * im[3] + a2 * im[4];
                        const float o2 = a0 * im[2] + a1 * im[4] + a2 * im[5];
                        out[jk][i] = cf3_pack(o0, o1, o2, &rng);
                    }
                }
        }
}

/* __rasterise_to_image (AMR.h:1313-1353) with the quadratic __interpolate_dat
 * that AMR.h:4 selects and, when boxes, the tile-edge boxes of
 * DISPLAY_BOUNDING_BOXES.  out: width*height float4, row-major.  Index
 * arithmetic in uint32 as the reference writes it. */
EXPORT void orc_rasterise(const float *data, const int32_t *children, uint32_t width, uint32_t height, int S, int P,
                          float z_slice, int boxes, float *out)
{
    const uint32_t T = (uint32_t)(S + 2 * P);
    for (uint32_t py = 0; py < height; py++)
        for (uint32_t px = 0; px < width; px++) {
            float t[3] = {(float)px / (float)width - .5f, (float)py / (float)height - .5f, z_slice};
            uint32_t cur = 0, off = (t[0] > 0.f) | ((t[1] > 0.f) << 1) | ((t[2] > 0.f) << 2);
            int child = children[cur * 8u + off];
            uint32_t depth = 0;
            while (child != -1) {
                cur = (uint32_t)child;
                for (int k = 0; k < 3; k++) {
                    const float v = t[k] * 2.f;
                    t[k] = (v - floorf(v)) - .5f; /* fracf (helper_math.h:1475) */
                }
                off = (t[0] > 0.f) | ((t[1] > 0.f) << 1) | ((t[2] > 0.f) << 2);
                child = children[cur * 8u + off];
                depth++;
            }
            float *o = out + 4 * ((size_t)py * width + px);
            const float threshold = .5f - 2.f * (float)(1u << depth) / (float)width;
            if (boxes && fmaxf(fabsf(t[0]), fabsf(t[1])) >= threshold) {
                o[0] = 1.f, o[1] = 0.f, o[2] = 0.f, o[3] = 1.f;
                continue;
            }
            uint32_t ri[3];
            for (int k = 0; k < 3; k++) {
                t[k] = (t[k] + .5f) * (float)S + (float)P;
                ri[k] = (uint32_t)(t[k] + .5f);
                t[k] -= (float)ri[k];
            }
            uint32_t n = cur * T + ri[2];
            n = n * T + ri[1];
            n = n * T + ri[0];
            const float v = interp_lagrange(data, t, (int64_t)n, (int)T) + .5f;
            o[0] = v, o[1] = v, o[2] = v, o[3] = 1.f;
        }
}

/* ___rgba(float4) (image_process.h:14-21) per pixel, as __encode_img applies it */
EXPORT void orc_encode_rgba8(const float *raster, uint32_t *pixels, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        const float *v = raster + 4 * i;
        if (isnan(v[0]) || isnan(v[1]) || isnan(v[2]) || isnan(v[3])) {
            pixels[i] = 4294902015u;
            continue;
        }
        uint32_t c[4];
        for (int k = 0; k < 4; k++) c[k] = (uint32_t)(fclampf(v[k], 0.f, 1.f) * 255.f);
        pixels[i] = (c[3] << 24) | (c[2] << 16) | (c[1] << 8) | c[0];
    }
}

/* __init_temp (kernel.cu:9-43) with absolute_central_position (AMR.h:33-43):
 * the two-bump field in float with the reference's operation order, into
 * both arrays; all cells of depth-0 slots, interior cells of deeper ones. */
EXPORT void orc_init_temp(float *old_d, float *new_d, const uint32_t *pos, const int32_t *tiles, int n_tiles, int S,
                          int P, float outer_size, float outer_dx, float mul)
{
    const int T = S + 2 * P;
    for (int t = 0; t < n_tiles; t++) {
        const int node = tiles[t];
        const int depth = pos_depth(pos[node]);
        if (depth == -1) continue;
        float c[3] = {0.f, 0.f, 0.f};
        uint32_t s = pos[node] & 134217727u;
        for (int i = 0; i < depth; i++, s >>= 3) {
            c[0] += ((s & 1u) ? .25f : -.25f) / (float)(1 << i);
            c[1] += ((s & 2u) ? .25f : -.25f) / (float)(1 << i);
            c[2] += ((s & 4u) ? .25f : -.25f) / (float)(1 << i);
        }
        for (int z = 0; z < T; z++)
            for (int y = 0; y < T; y++)
                for (int x = 0; x < T; x++) {
                    if (depth > 0 && (x < P || x >= S + P || y < P || y >= S + P || z < P || z >= S + P)) continue;
                    const int idx[3] = {x, y, z};
                    float p[3];
                    for (int k = 0; k < 3; k++)
                        p[k] = c[k] * outer_size + ((((float)idx[k] + .5f) - (float)T * .5f) * outer_dx) /
                                                       (float)(1u << depth);
                    const float m0 = p[0] + 25.f, m1 = p[1] - .5f, m2 = p[2];
                    const float q0 = p[0] - 25.f, q1 = p[1] + .5f, q2 = p[2];
                    const float v = mul * (10.f / ((m0 * m0 + m1 * m1) + m2 * m2 + 10.f) -
                                           10.f / ((q0 * q0 + q1 * q1) + q2 * q2 + 10.f));
                    const int64_t o = (int64_t)node * T * T * T + ((int64_t)z * T + y) * T + x;
                    old_d[o] = v;
                    new_d[o] = v;
                }
    }
}

/* ___rgba(float) and ___rgba(float3) of a decoded word (image_process.h:22-38) */
EXPORT void orc_encode_rgba8_f32(const float *img, uint32_t *pixels, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        if (isnan(img[i])) {
            pixels[i] = 4294902015u;
            continue;
        }
        const uint32_t c = (uint32_t)(fclampf(img[i], 0.f, 1.f) * 255.f);
        pixels[i] = (255u << 24) | (c << 16) | (c << 8) | c;
    }
}
EXPORT void orc_encode_rgba8_cf3(const uint32_t *words, uint32_t *pixels, int64_t n)
{
    for (int64_t i = 0; i < n; i++) {
        float v[3];
        cf3_unpack(words[i], v);
        if (isnan(v[0]) || isnan(v[1]) || isnan(v[2])) {
            pixels[i] = 4294902015u;
            continue;
        }
        uint32_t c[3];
        for (int k = 0; k < 3; k++) c[k] = (uint32_t)(fclampf(v[k], 0.f, 1.f) * 255.f);
        pixels[i] = (255u << 24) | (c[2] << 16) | (c[1] << 8) | c[0];
    }
}
