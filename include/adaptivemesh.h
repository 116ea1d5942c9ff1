/*
 * adaptivemesh.h -- C ABI of the MI355X (gfx950) AdaptiveMesh per-tile path.
 *
 * Plain C: pointers, sizes and small POD structs only.  Every device pointer
 * is owned by the caller and is never freed here (the reference's
 * smart_gpu_buffer is likewise a raw cudaMalloc pointer with shallow copies,
 * CUDA_memory.h:89-134).  Every call enqueues on the HIP stream passed as
 * `stream` (NULL = the default stream) and returns 0 on success, a positive
 * hipError_t value, or a negative AM_ERR_* code; am_last_error_string()
 * describes the last failure on the calling thread.  Nothing throws across
 * this boundary.
 *
 * Data layout (identical to the reference):
 *   field      float  [slot][z][y][x] over the padded tile, T = S + 2P per
 *                     side (AMR.h:11-15, 752, 758)
 *   words      uint32 [slot][z][y][x] over the S^3 interior, one
 *                     compressed_float3 per cell (AMR.h:1117, 1177)
 *   positions  uint32 octree_abs_pos per slot (AMR.h:22-55)
 *   children   int32  8 per slot, -1 = none (AMR.h:123)
 *   boundaries uint64 6 per slot, node_data | rel_pos << 32 (AMR.h:56-100)
 *
 * Field arrays may instead use the interior-first tile layout
 * (AM_LAYOUT_INTERIOR_FIRST, below), selected per call by the tile_layout
 * member of the field structs; the RK4 drivers keep their state in it.
 *
 * Each function comment names the reference kernel and host wrapper it
 * replaces.  The reference launches over every used slot and lets inactive
 * depths exit early; here the caller passes a device list of slots
 * (am_mesh.tiles), and a list of every used slot reproduces the reference.
 */
#ifndef ADAPTIVEMESH_H
#define ADAPTIVEMESH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AM_MAX_VARS 32

enum am_status {
    AM_OK = 0,
    AM_ERR_INVALID = -1,              /* bad argument (null pointer, bad size) */
    AM_ERR_UNSUPPORTED_GEOMETRY = -2, /* no kernel instantiated for (S, P, order) */
    AM_ERR_SLOTS_EXHAUSTED = -3,      /* octree has no free slot (AMR.h:210) */
    AM_ERR_DEPTH_LIMIT = -4,          /* parent already at depth_limit (AMR.h:211) */
    AM_ERR_OCCUPIED = -5,             /* child offset already in use (AMR.h:210) */
    AM_ERR_MISALIGNED = -6,           /* field pointer not 16-byte aligned */
    AM_ERR_TOO_MANY_VARS = -7,        /* more than AM_MAX_VARS in one batch */
    AM_ERR_ALLOC = -8,                /* device allocation failed */
    AM_ERR_LAUNCH = -9,               /* kernel launch failed */
    AM_ERR_IO = -10,                  /* a file could not be opened, read or written */
    AM_ERR_FORMAT = -11               /* a checkpoint file is damaged or does not fit the driver */
};

enum am_interp {
    AM_INTERP_LAGRANGE = 0,  /* 10-point quadratic, the reference default (AMR.h:4, 673-699) */
    AM_INTERP_TRILINEAR = 1, /* the reference's alternative (AMR.h:767-785) */
    AM_INTERP_TRICUBIC = 2,  /* 4x4x4 Lagrange in space, quadratic in time (extension) */
    AM_INTERP_LEARNED = 3    /* learned per-octant prolongation on one-level-coarser faces, tricubic elsewhere */
};

/* Tile layouts of fp32 field arrays (the tile_layout member of the field
 * structs).  No counterpart in the reference, which uses the padded one.
 *   AM_LAYOUT_PADDED: cell (x, y, z) of the padded tile (0 <= x, y, z < T)
 *     of slot s at s * T^3 + (z * T + y) * T + x (AMR.h:752, 758).
 *   AM_LAYOUT_INTERIOR_FIRST: slot s starts at s * am_tile_stride() floats
 *     (T^3 rounded up to a multiple of 32); within it, first the S^3
 *     interior cells, ((z - P) * S + (y - P)) * S + (x - P); then the six
 *     face ghost slabs, face f (0 +x, 1 +y, 2 +z, 3 -x, 4 -y, 5 -z) at
 *     S^3 + f * P * S^2, each P layers of S^2 cells, layer l counted outward
 *     from the interior, cell (a, b) at (l * S + b) * S + a with (a, b) the
 *     interior coordinates of the two other axes in x, y, z order; then the
 *     edge and corner cells (am_tile_layout_convert gives the exact map).
 * Interior rows and slab rows are contiguous runs of S floats, so the
 * pointwise stage kernels stream whole tiles and the equal-depth ghost push
 * writes whole rows. */
enum am_tile_layout { AM_LAYOUT_PADDED = 0, AM_LAYOUT_INTERIOR_FIRST = 1 };

const char *am_last_error_string(void);
int am_version(void);

/* Tile geometry.  The reference fixes these as constexpr globals
 * (AMR.h:9-17); here they are a runtime value and the kernels dispatch to an
 * instantiation for (S, P, fd_order).  Instantiated: (16,3,{2,4,6}),
 * (16,2,{2,4}), (28,2,{2,4}). */
typedef struct am_geometry {
    int32_t size_domain;    /* S, interior cells per side (AMR.h:11) */
    int32_t padding_domain; /* P, ghost layers (AMR.h:12) */
    int32_t fd_order;       /* 2, 4 or 6 and <= 2P; the reference uses 2P (AMR.h:1171) */
    float outer_size;       /* root cube side (AMR.h:10) */
    float outer_dx;         /* outer_size / S (AMR.h:16) */
    float outer_dt;         /* outer_dx * cfl (AMR.h:17) */
} am_geometry;

int am_geometry_init(am_geometry *g, int32_t size_domain, int32_t padding_domain, int32_t fd_order,
                     float outer_size, float cfl);
int am_geometry_supported(const am_geometry *g);

/* Floats per slot of a field array in `layout` (enum am_tile_layout): T^3,
 * or T^3 rounded up to a multiple of 32 for AM_LAYOUT_INTERIOR_FIRST; -1 for
 * a bad argument.  No counterpart in the reference. */
int64_t am_tile_stride(const am_geometry *g, int32_t layout);

/* Host only: out[(z * T + y) * T + x] = offset within its slot of padded
 * cell (x, y, z) in `layout`, for every cell of one tile (T^3 entries).  No
 * counterpart in the reference. */
int am_tile_cell_offsets(const am_geometry *g, int32_t layout, uint32_t *out);

/* Copy n_slots tiles of one field from layout `from` to layout `to` (device
 * arrays, no overlap), every cell of every tile.  No counterpart in the
 * reference: it moves data between the padded layout of the reference's
 * interfaces and the interior-first layout of the RK4 drivers. */
int am_tile_layout_convert(const float *src, float *dst, int32_t n_slots, const am_geometry *g, int32_t from,
                           int32_t to, void *stream);

/* What a kernel sees of the octree: device copies of AMR's arrays (the
 * reference's AMR::copy_to_gpu, AMR.h:268-275), the slot list to visit, and
 * the sub-cycling counter timer_helper with max_depth (AMR.h:178-188). */
typedef struct am_mesh {
    const uint32_t *positions;
    const int32_t *children;
    const uint64_t *boundaries;
    const int32_t *tiles;
    int32_t n_tiles;
    uint32_t substep;
    uint32_t max_depth;
    /* launch shape: 0 = one workgroup per work item (tile, or tile x block);
     * > 0 = at most this many workgroups, each looping over work items
     * (a persistent launch over the level; results do not depend on it).
     * Honoured by am_fd_batched and the RHS stage kernels. */
    int32_t grid_limit;
} am_mesh;

/* ------------------------------------------------------- device memory */
int am_device_alloc(void **ptr, size_t bytes);
int am_device_free(void *ptr);
int am_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int am_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int am_memset_async(void *dst, int value, size_t bytes, void *stream);
int am_memcpy_d2d(void *dst, const void *src, size_t bytes, void *stream);
int am_stream_sync(void *stream);
/* a blocking stream at the default priority (it orders with the NULL stream) */
int am_stream_create(void **stream);
int am_stream_destroy(void *stream);
int am_device_sync(void);
/* compute units of the current device (for persistent launch shapes) */
int32_t am_device_cu_count(void);
/* HIP events, for timing kernels on the stream they run on */
int am_event_create(void **ev);
int am_event_record(void *ev, void *stream);
int am_event_elapsed_ms(void *start, void *stop, float *ms); /* synchronises on stop */
int am_event_destroy(void *ev);
/* Ordering between streams (no counterpart in the reference, which runs on
 * one stream): an event without timing, and a wait that makes later work on
 * `stream` start after the work recorded into `ev`.  Both work under graph
 * capture, where they become edges between the two streams' nodes. */
int am_event_create_ordering(void **ev);
int am_stream_wait_event(void *stream, void *ev);

/* ------------------------------------------------------------- graphs
 * No counterpart in the reference, which launches eagerly between device
 * syncs.  am_graph_begin starts capturing `stream`; am_graph_end stops and
 * instantiates what was captured into *exec; am_graph_launch replays it on
 * `stream`.  Captured work must not allocate, copy synchronously or memset
 * (kernels only). */
int am_graph_begin(void *stream);
int am_graph_end(void *stream, void **exec);
int am_graph_launch(void *exec, void *stream);
int am_graph_destroy(void *exec);

/* table[i] = the (i+1)-th fast_prng draw after state `seed`, i < n: the
 * dither seeds a recorded launch reads through am_fd_batch.seed_table
 * (one thread; n is a few thousand per coarse step). */
int am_seed_table_fill(int32_t *table, int32_t n, int32_t seed, void *stream);

/* ------------------------------------------------------- FD stencils */

/* First derivatives -> one word per interior cell.
 * Replaces __differentiate(field, dif, ...) AMR.h:1152-1179 and its wrapper
 * differentiate(field, dif, amr, stream, rng) AMR.h:1280-1290; `seed` is the
 * wrapper's rng.generate_int(). */
int am_fd_d1(const float *field, uint32_t *d1, const am_mesh *mesh, const am_geometry *g, int32_t seed,
             void *stream);

/* First derivatives + diagonal second derivatives (d2 packed first).
 * Replaces __differentiate(field, dif, d2if, ...) AMR.h:1181-1219 and
 * differentiate(field, dif, d2if, amr, stream, rng) AMR.h:1292-1302. */
int am_fd_d1_d2diag(const float *field, uint32_t *d1, uint32_t *d2diag, const am_mesh *mesh,
                    const am_geometry *g, int32_t seed, void *stream);

/* Off-diagonal second derivatives (xy, xz, yz), boundary-careful at tile
 * edges.  Replaces __mixed_partials AMR.h:1232-1266 and mixed_partials
 * AMR.h:1268-1278. */
int am_fd_mixed(const float *field, uint32_t *d2off, const am_mesh *mesh, const am_geometry *g, int32_t seed,
                void *stream);

/* Fused: gradient + full Hessian in one pass over the tile, bit-identical to
 * am_fd_d1_d2diag(seed_d) followed by am_fd_mixed(seed_m).  Replaces
 * yield_first_second_derivs AMR.h:1304-1310. */
int am_fd_all(const float *field, uint32_t *d1, uint32_t *d2diag, uint32_t *d2off, const am_mesh *mesh,
              const am_geometry *g, int32_t seed_d, int32_t seed_m, void *stream);

/* Several fields in one launch.  Per variable v: d1[v] is required unless
 * only d2off[v] is set; d2diag[v] != NULL adds the diagonal Hessian;
 * d2off[v] != NULL adds the mixed partials (seeded by seed_m[v]). */
typedef struct am_fd_batch {
    const float *field[AM_MAX_VARS];
    uint32_t *d1[AM_MAX_VARS];
    uint32_t *d2diag[AM_MAX_VARS];
    uint32_t *d2off[AM_MAX_VARS];
    int32_t seed_d[AM_MAX_VARS];
    int32_t seed_m[AM_MAX_VARS];
    int32_t n_vars;
    int32_t z_block; /* planes per workgroup: 0 = whole tile, else S, S/2 or S/4 */
    /* NULL: seed_d/seed_m are the per-launch dither seeds.  Otherwise they
     * are indices into this device array, which holds the seeds: a launch
     * recorded once in a graph then reads fresh seeds on every replay. */
    const int32_t *seed_table;
    /* output format, enum am_word_format: the reference's dithered
     * compressed_float3 (4 bytes per cell), or for the precision study three
     * fp32 (12 bytes) or three fp16 (6 bytes) per cell (S=16, P=3, order 4
     * tiles only); d1/d2diag/d2off then point to arrays of that many bytes
     * per cell */
    int32_t word_format;
    /* padded planes each workgroup prefetches into registers ahead of the
     * plane it differentiates: 0 or 1 (default), 2 or 3; cf3 words only */
    int32_t prefetch_planes;
    /* axes (bits 1 x, 2 y, 4 z; mesh->boundaries required when nonzero)
     * whose ghost cells across a face with an equal-depth neighbour (the
     * faces the RK4 stage kernels push) are read from that neighbour's
     * interior, so the tile's own ghost layers there need not be current.
     * The words equal those of a field whose ghost layers hold the pushed
     * copies. */
    int32_t direct_faces;
    /* enum am_tile_layout of every field (AM_LAYOUT_INTERIOR_FIRST takes
     * direct_faces 0, 2, 4 or 6: no x) */
    int32_t tile_layout;
    /* Upwinded advection with Kreiss-Oliger dissipation (no counterpart in
     * the reference; S=16, P=3, fd_order 4, cf3 words, prefetch 1 only):
     * adv[v] != NULL receives one float per interior cell (the words'
     * layout),
     *   [bit v of advect_mask] sum_k beta^k L_k f + (ko_sigma / 64) (1/h) sum_k K_k f
     * with beta^k the fields shift[0..2] at the cell, L_k the fourth-order
     * lopsided first derivative on the side beta^k points to (offsets -1..3
     * for beta^k > 0, else -3..1) and K_k the sixth difference (1, -6, 15,
     * -20, 15, -6, 1) along axis k, in float in a fixed order (the oracle's
     * orc_fd_adv).  Reads three ghost layers. */
    float *adv[AM_MAX_VARS];
    const float *shift[3];
    uint32_t advect_mask;
    float ko_sigma;
} am_fd_batch;
enum am_word_format { AM_WORD_CF3 = 0, AM_WORD_F32 = 1, AM_WORD_F16 = 2 };
/* bytes per cell of one word array in `format` (4, 12, 6), -1 if unknown */
int32_t am_word_bytes(int32_t format);
/* cf3 word arrays (and adv) are addressed with 32-bit byte offsets: every
 * listed slot must satisfy (slot + 1) * S^3 <= 2^30; with upwinding the
 * shift fields are too, so (slot + 1) * (the tile stride) <= 2^30. */
int am_fd_batched(const am_fd_batch *b, const am_mesh *mesh, const am_geometry *g, void *stream);

/* ------------------------------------------------- ghost cells and tree */

/* Fill the six face ghost slabs of every listed active tile from its
 * boundary target, interpolated in space (interp) and in time between old
 * and new (linear; quadratic through older/old/new when interp is
 * AM_INTERP_TRICUBIC and older != NULL).  Replaces __copy_boundaries
 * AMR.h:702-762 / 789-848 and copy_boundaries AMR.h:880-900.  Tiles whose
 * target is being filled in the same call race exactly as in the reference;
 * the C++ layer therefore calls it once per depth, coarse first. */
int am_ghost_fill(float *old_, float *new_, const float *older, const am_mesh *mesh, const am_geometry *g,
                  int32_t copy_only_new, int32_t interp, void *stream);

typedef struct am_field_batch {
    float *old_[AM_MAX_VARS];
    float *new_[AM_MAX_VARS];
    float *older[AM_MAX_VARS];
    int32_t n_vars;
    int32_t tile_layout; /* enum am_tile_layout of every field */
} am_field_batch;
int am_ghost_fill_batched(const am_field_batch *b, const am_mesh *mesh, const am_geometry *g,
                          int32_t copy_only_new, int32_t interp, void *stream);

/* old <- new on every cell of each listed active tile; interior cells with
 * a child take the mean of the child's 2x2x2 new cells (restriction).
 * Replaces __copy_new_to_old AMR.h:961-996 and copy_new_to_old
 * AMR.h:998-1051.  Uses mesh->children.  In the batched form a non-NULL
 * older[v] first receives the previous old (a third time level for
 * quadratic time interpolation; an extension). */
int am_copy_new_to_old(float *old_, const float *new_, const am_mesh *mesh, const am_geometry *g, void *stream);
int am_copy_new_to_old_batched(const am_field_batch *b, const am_mesh *mesh, const am_geometry *g, void *stream);
/* The batched copy-back for a driver whose first RK4 stage reads new rather
 * than old.  No counterpart in the reference.  As am_copy_new_to_old_batched,
 * then (a second launch, after every restriction has read its children)
 * new <- old on the restricted cells, so new equals old on the interior of
 * every active tile.  On a tile whose keep_ghosts[slot] (device, one byte
 * per slot) is 0, only interior cells of old and older are written (the
 * interior planes as whole vector runs, which take a few ghost cells
 * along): old's ghost cells are read only by interpolation into a finer
 * neighbour's face, and keep_ghosts marks the slots that are such a coarse
 * source. */
int am_copy_new_to_old_interior(const am_field_batch *b, const am_mesh *mesh, const uint8_t *keep_ghosts,
                                const am_geometry *g, void *stream);

/* Prolong parent `parent` into every cell of new child `child` at octant
 * `offset`.  Replaces __copy_to_child AMR.h:1053-1073 / copy_to_child
 * AMR.h:1075-1082.  interp: AM_INTERP_LAGRANGE or AM_INTERP_TRILINEAR. */
int am_copy_to_child(float *buf, uint32_t offset, int32_t child, int32_t parent, const am_geometry *g,
                     int32_t interp, void *stream);
/* am_copy_to_child for every non-null old_, new_ and older buffer of the
 * batch in one launch (the prolongation a new tile needs for all its fields,
 * AMR.h:1075-1115 called once per field and buffer). */
int am_copy_to_child_batched(const am_field_batch *b, uint32_t offset, int32_t child, int32_t parent,
                             const am_geometry *g, int32_t interp, void *stream);

/* Launch lists of the octree, built on the device by stream compaction
 * (the reference's GPU scan, parallel.h:10-80, applied to the launch
 * domain that its wrappers cover with every slot, AMR.h:640-644 and
 * 880-1310).  From the device copies of positions, children (8 per slot)
 * and boundaries (6 per slot) of slots 0 .. n_slots-1 it writes, in
 * ascending slot order:
 *   tiles: [every slot < n_slots][depth 0]..[depth depth_limit][the leaves]
 *          (at most 3 * n_slots entries);
 *   faces: [depth 0]..[depth depth_limit]: slot * 6 + face of every face of
 *          a tile of depth >= 1 whose target is coarser (at most 6 * n_slots);
 *   coarse_source[s] = 1 for every slot such a face reads, 0 for the other
 *          slots < max_slots;
 *   counts (device, 4 * depth_limit + 8 int32): tile list offsets and
 *          lengths (depth_limit + 3 each), then face list offsets and
 *          lengths (depth_limit + 1 each).
 * workspace: am_tile_lists_workspace(max_slots, depth_limit) int32 of device
 * memory.  Three launches on `stream`; nothing is allocated. */
int32_t am_tile_lists_workspace(int32_t max_slots, int32_t depth_limit);
int am_tile_lists(const uint32_t *positions, const int32_t *children, const uint64_t *boundaries, int32_t n_slots,
                  int32_t max_slots, int32_t depth_limit, int32_t *tiles, int32_t *faces, uint8_t *coarse_source,
                  int32_t *workspace, int32_t *counts, void *stream);

/* ------------------------------------------------------ SimpleWave */

/* __init_temp (kernel.cu:9-43), the reference's wave-test initial data:
 * f = mul (10 / (|x - a|^2 + 10) - 10 / (|x + a|^2 + 10)), a = (-25, 1/2, 0)
 * in physical units, into both fields at every cell of a listed depth-0
 * tile and at the interior cells of deeper ones.  Cell x = tile centre
 * (AMR.h:33-43) + (i + 1/2 - T/2) outer_dx / 2^depth per axis.  The
 * reference launches it for T = 32 tiles only; here any supported geometry. */
int am_init_temp(float *old_field, float *new_field, const am_mesh *mesh, const am_geometry *g, float mul,
                 void *stream);

/* __rasterise_to_image (AMR.h:1313-1353): a z-slice through the whole
 * octree as a width x height image.  Pixel (x, y) samples the point
 * (x/width - 1/2, y/height - 1/2, z_slice) of the root cube (units of
 * outer_size, z_slice in [-1/2, 1/2)) in the leaf holding it, with the
 * quadratic __interpolate_dat of that leaf (AMR.h:4, 673-699), as
 * (v + 1/2, v + 1/2, v + 1/2, 1); with boxes != 0, pixels within
 * 2 * 2^depth / width of a tile's x or y edge are (1, 0, 0, 1) (the
 * reference's DISPLAY_BOUNDING_BOXES).  raster: width * height float4
 * (16-byte aligned, device), row-major.  mesh: children only. */
int am_rasterise_slice(const float *field, const am_mesh *mesh, const am_geometry *g, uint32_t width,
                       uint32_t height, float z_slice, int32_t boxes, float *raster, void *stream);
/* __encode_img with ___rgba(float4) (image_process.h:10-54): n float4
 * pixels to RGBA8, R | G << 8 | B << 16 | A << 24, each channel
 * (uint)(clamp(c, 0, 1) * 255); a NaN channel gives 0xFFFF00FF. */
int am_encode_rgba8(const float *raster, uint32_t *pixels, int64_t n, void *stream);
/* ___rgba(float) and ___rgba(compressed_float3) (image_process.h:22-38), as
 * save_image applies them to a field or a word array: a float v gives grey
 * (c, c, c, 255), c = (uint)(clamp(v, 0, 1) * 255); a word gives
 * (r, g, b, 255) from its decoded components; NaN gives 0xFFFF00FF. */
int am_encode_rgba8_f32(const float *image, uint32_t *pixels, int64_t n, void *stream);
int am_encode_rgba8_cf3(const uint32_t *words, uint32_t *pixels, int64_t n, void *stream);

/* Refinement criterion of the SimpleWave plugin (replaces __amr_criterion_1
 * and __amr_criterion_2, SimpleWave.h:91-144): for each listed tile and
 * octant o = x + 2y + 4z, out[slot * 8 + o] = the largest over the octant's
 * cells of max(|d1| outer_dx / 2^d, |d2diag| outer_dx^2 / 4^d), from the
 * decoded gradient and diagonal-Hessian words; NaN for a slot that is unused
 * or not active at mesh->substep.  out: device, 8 floats per slot. */
int am_amr_criterion(const uint32_t *d1, const uint32_t *d2diag, const am_mesh *mesh, const am_geometry *g,
                     float *out, void *stream);

/* Heun stages of u_t = nu lap u - u u_x, nu = 0.01.  Replaces
 * __predictor_step / __corrector_step SimpleWave.h:7-46 and
 * wave_equation_predictor / _corrector SimpleWave.h:48-72. */
int am_wave_predictor(const float *old_, const uint32_t *d1, const uint32_t *d2diag, float *new_,
                      const am_mesh *mesh, const am_geometry *g, void *stream);
int am_wave_corrector(const float *old_, const uint32_t *d1, const uint32_t *d2diag, float *new_,
                      const am_mesh *mesh, const am_geometry *g, void *stream);

/* ------------------------------------------------------------ BSSN.h */

/* out = gamma^{ij} d_i d_j f from the (diag, off) Hessian words and the six
 * metric fields xx,xy,xz,yy,yz,zz.  Replaces __deriv_trace BSSN.h:7-22. */
int am_deriv_trace(const uint32_t *hdiag, const uint32_t *hoff, const float *const *metric6, float *out,
                   const am_mesh *mesh, const am_geometry *g, void *stream);

/* Christoffel symbols from the six metric-derivative words.  Word jk holds
 * (Gamma^x_jk, Gamma^y_jk, Gamma^z_jk).  compat = 1 reproduces the
 * reference, which contracts with gamma_ij instead of gamma^ij
 * (BSSN.h:82-83); compat = 0 is the true raised symbol.  Replaces
 * __compute_raised_christoffel_symbols BSSN.h:55-91 and its wrapper
 * BSSN.h:93-102. */
int am_christoffel(const float *const *metric6, const uint32_t *const *dmetric6, uint32_t *const *out6,
                   int32_t compat, const am_mesh *mesh, const am_geometry *g, int32_t seed, void *stream);

/* Ghost fill at an intermediate Runge-Kutta time: faces whose target is
 * coarser by dd levels interpolate the target's old/new at
 * (k + stage_frac) / 2^dd of its step (relative_time, AMR.h:634-639, gives
 * k and k + 1); equal-depth faces take the target's new, which holds the same
 * stage, by `equal_mode`:
 *   AM_EQUAL_FORMULA  the reference's clamped interpolation; with
 *                     stage_frac = 1 the values of am_ghost_fill_batched
 *                     (copy_only_new = 1) for finite data
 *   AM_EQUAL_COPY     an exact copy of the neighbour cell the ghost maps onto
 *   AM_EQUAL_SKIP     left alone (written by am_bssn_rhs_stage's push)
 * No counterpart in the reference, whose Heun scheme only fills at k + 1. */
enum am_equal_mode { AM_EQUAL_FORMULA = 0, AM_EQUAL_COPY = 1, AM_EQUAL_SKIP = 2 };
int am_ghost_fill_stage_batched(const am_field_batch *b, const am_mesh *mesh, const am_geometry *g, float stage_frac,
                                int32_t interp, int32_t equal_mode, void *stream);
/* The same fill over a list of faces instead of every face of mesh->tiles:
 * faces[i] = slot * 6 + face (face 0-5: +x, +y, +z, -x, -y, -z); each
 * listed face is filled exactly as am_ghost_fill_stage_batched fills it, so
 * a list holding every face whose target is coarser equals that call with
 * AM_EQUAL_SKIP.  The slots must be used slots of the mesh's octree.  No
 * counterpart in the reference, which launches one thread block per tile
 * face and leaves the faces with nothing to fill at once (AMR.h:702-762);
 * the list spares those launches. */
int am_ghost_fill_stage_faces(const am_field_batch *b, const am_mesh *mesh, const int32_t *faces, int32_t n_faces,
                              const am_geometry *g, float stage_frac, int32_t interp, int32_t equal_mode,
                              void *stream);

/* ----------------------------------------------------------- BSSN scheme
 * No counterpart in the reference (BSSN.h has no right-hand side or time
 * update, BSSN.h:143-159): the moving-puncture BSSN system with 24 fields
 *   0 chi | 1-6 gt_ij | 7 K | 8-13 At_ij | 14-16 Gt^i | 17 alpha | 18-20 beta^i | 21-23 B^i
 * (symmetric tensors in xx,xy,xz,yy,yz,zz order).  The Hessian words are for
 * chi, gt_ij, alpha, beta^i in that order (11 fields).  The conformal
 * Christoffel symbols are formed in registers from gt_ij and its gradient
 * words (am_christoffel remains for the reference's BSSN plugin). */
#define AM_BSSN_NVARS 24
#define AM_BSSN_NHESS 11

typedef struct am_bssn_params {
    float eta;            /* Gamma-driver damping */
    float shift_coeff;    /* d_t beta^i = shift_coeff B^i (+ advection), usually 3/4 */
    float chi_floor;      /* chi used in divisions is max(chi, chi_floor) */
    int32_t advect_shift; /* 1: beta^k d_k terms in the beta and B equations */
    int32_t slicing;      /* 0: 1+log, d_t alpha = -2 alpha K; 1: harmonic, -alpha^2 K */
} am_bssn_params;

typedef struct am_bssn_fields {
    const float *state[AM_BSSN_NVARS]; /* stage input, padded tiles */
    const float *old_[AM_BSSN_NVARS];  /* y_n at the start of the step */
    float *new_[AM_BSSN_NVARS];        /* stage output (interior cells written) */
    float *acc[AM_BSSN_NVARS];         /* RK4 accumulator, interior-only layout (slot * S^3 + ...) */
    const uint32_t *d1[AM_BSSN_NVARS];
    const uint32_t *d2diag[AM_BSSN_NHESS];
    const uint32_t *d2off[AM_BSSN_NHESS];
    int32_t word_format; /* enum am_word_format of d1/d2diag/d2off (non-cf3: S=16, P=3 only) */
    int32_t push_layers; /* ghost layers the push writes, innermost first: 1..P, 0 = all P */
    int32_t push_axes;   /* faces the push writes, by axis: bits 1 x, 2 y, 4 z; 0 = all three */
    int32_t tile_layout; /* enum am_tile_layout of state, old_ and new_ */
    /* NULL: the advection terms beta^k d_k use the gradient words.
     * Otherwise 24 float arrays (words' layout) holding each field's
     * upwinded advection plus Kreiss-Oliger dissipation (am_fd_batch.adv,
     * advect_mask = fields 0-17, and 18-23 with params.advect_shift), which
     * replace them: the B^i equation's d_0 Gt^i then drops adv[14+i] with
     * advect_shift, and beta^i, B^i take adv whole.  S=16, P=3, cf3 words. */
    const float *adv[AM_BSSN_NVARS];
} am_bssn_fields;

/* One classical RK4 stage on the interior of every listed active tile:
 * k = f(state); stage 0: acc = k, new = old + dt/2 k; stage 1: acc += 2k,
 * new = old + dt/2 k; stage 2: acc += 2k, new = old + dt k; stage 3:
 * new = old + dt/6 (acc + k), with dt = outer_dt / 2^depth.  With
 * push_ghosts != 0 (mesh->boundaries required) every new value within P
 * cells of a face whose neighbour has the same depth is also written into
 * that neighbour's face ghost cell: the AM_EQUAL_COPY fill, fused.  Each
 * array is addressed with 32-bit byte offsets: every slot a tile list names,
 * and every neighbour the push writes, must lie within the first 4 GiB of
 * each array (slot < 2^32 / (4 x the tile stride)), which the 118 arrays of
 * one call reach only beyond the HBM of one GPU. */
int am_bssn_rhs_stage(const am_bssn_fields *f, int32_t stage, const am_bssn_params *p, const am_mesh *mesh,
                      const am_geometry *g, int32_t push_ghosts, void *stream);

/* Hamiltonian and momentum constraints of `state` on every listed tile,
 * reduced in FP64 to one {sum H^2, sum M_i M_i, cells, max |H|} quadruple per
 * workgroup in `partial` (device, 4 * am_bssn_constraint_blocks() doubles). */
int am_bssn_constraints(const am_bssn_fields *f, const am_bssn_params *p, const am_mesh *mesh, const am_geometry *g,
                        double *partial, void *stream);
int64_t am_bssn_constraint_blocks(const am_geometry *g, int32_t n_tiles);

/* am_bssn_constraints that allocates nothing and reduces on the device:
 * out4 (device, 4 doubles) = {sum H^2, sum M_i M_i, cells, max |H|} over
 * every listed tile.  No counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc): the per-workgroup partial sums
 * live in a fixed 2 MiB device array of the library, one per GPU, held
 * until the process ends, so at most 65536 partial sums (4096 tiles of 16^3
 * cells) per call; more is refused with AM_ERR_INVALID.  Every call on one
 * GPU shares that array: calls that overlap on one GPU race. */
int am_bssn_constraints_noalloc(const am_bssn_fields *f, const am_bssn_params *p, const am_mesh *mesh,
                                const am_geometry *g, double *out4, void *stream);

/* ------------------------------------------------------ scalar wave, RK4
 * No counterpart in the reference, whose SimpleWave.h plugin is a Heun step
 * of a Burgers-type equation (SimpleWave.h:7-72): the wave equation
 * u_tt = lap u as the first-order system u_t = v, v_t = lap u, advanced by
 * the same RK4 stage update and equal-depth push as am_bssn_rhs_stage.  The
 * Laplacian is the sum of the components of u's d2diag word (the FD kernel's
 * second-derivative output). */
typedef struct am_swave_fields {
    const float *state[2]; /* u, v: stage input */
    const float *old_[2];
    float *new_[2];
    float *acc[2];            /* interior-only layout, as for am_bssn_fields */
    const uint32_t *d1_u;     /* gradient word of the stage input u */
    const uint32_t *d2diag_u; /* (u_xx, u_yy, u_zz) word of the stage input u */
    int32_t push_layers;      /* as am_bssn_fields.push_layers */
    int32_t push_axes;        /* as am_bssn_fields.push_axes */
    int32_t tile_layout;      /* as am_bssn_fields.tile_layout */
} am_swave_fields;

int am_swave_rhs_stage(const am_swave_fields *f, int32_t stage, const am_mesh *mesh, const am_geometry *g,
                       int32_t push_ghosts, void *stream);

/* Energy of `state` on every listed tile, reduced in FP64 to one
 * {sum (v^2 + |grad u|^2) h^3 / 2, sum u^2 h^3, cells, max |u|} quadruple per
 * workgroup in `partial` (4 * am_bssn_constraint_blocks() doubles; h is the
 * tile's cell size outer_dx / 2^depth, grad u comes from d1_u). */
int am_swave_energy(const am_swave_fields *f, const am_mesh *mesh, const am_geometry *g, double *partial,
                    void *stream);

/* ------------------------------------------------- learned prolongation
 * No counterpart in the reference.  A per-octant MLP predicts the 2^3 fine
 * children of a coarse cell (24 fields) from the cell's 4^3 coarse stencil
 * (offsets -1..2 per axis), as a correction on top of the tricubic
 * baseline; each field's stencil is shifted by its centre value mu and
 * scaled by sigma = max |stencil - mu| + 1e-3 |mu| + 1e-7 (a field flat to
 * within rounding must not turn its rounding noise into O(1) inputs), the
 * output is the
 * residual divided by sigma.  Layers 1536 -> 128 -> 128 -> 192 with exact
 * GELU; parameters in one fp32 array: W1[1536][128] b1[128] W2[128][128]
 * b2[128] W3[128][192] b3[192] (weights [in][out], row-major). */
#define AM_MLP_IN 1536
#define AM_MLP_HID 128
#define AM_MLP_OUT 192
int64_t am_mlp_param_count(void);

/* C = op(A) op(B) on fp32 MFMA: A(m,k) = A[m*sam + k*sak], B(k,n) =
 * B[k*sbk + n*sbn], C row-major (ldc).  epilogue 0: C = AB (+ bias[n]);
 * 1: C = AB + bias, out2 = gelu(C); 2: C = AB * gelu'(aux) (aux like C).
 * M and N multiples of 64, K of 16. */
int am_gemm_f32(const float *A, int64_t sam, int64_t sak, const float *B, int64_t sbk, int64_t sbn, float *C,
                int64_t ldc, int32_t M, int32_t N, int32_t K, const float *bias, int32_t epilogue, const float *aux,
                float *out2, void *stream);

/* Activations and gradients of one batch (batch rows each): z1 h1 z2 h2
 * dz1 dz2 are batch x 128, y dy batch x 192. */
typedef struct am_mlp_work {
    float *z1, *h1, *z2, *h2, *y, *dy, *dz2, *dz1;
} am_mlp_work;
int am_mlp_forward(const float *params, const float *x, int32_t batch, am_mlp_work *w, void *stream);
/* gradients of the loss whose output gradient is w->dy (all of grads written) */
int am_mlp_backward(const float *params, const float *x, int32_t batch, am_mlp_work *w, float *grads, void *stream);
/* dy = 2 (y - t) / (batch * 192); partial[i] = FP64 sum of (y - t)^2 over
 * block i of am_mlp_mse_blocks(batch) */
int am_mlp_mse(const float *y, const float *t, int32_t batch, float *dy, double *partial, void *stream);
int64_t am_mlp_mse_blocks(int32_t batch);
/* Adam with bias correction; step counts from 1 */
int am_mlp_adam(float *params, const float *grads, float *m, float *v, int64_t n, int32_t step, float lr,
                float beta1, float beta2, float eps, void *stream);
/* Training pairs from fine padded tiles (24 field pointers): for each
 * listed tile, every coarse cell (restriction = mean of 8, AMR.h:975-995)
 * whose stencil lies in the padded tile, am_mlp_pairs_per_tile() of them
 * per tile, tile-major then z, y, x: x (n x 1536) normalised stencils,
 * t (n x 192) normalised residuals, mu and sigma (n x 24). */
int64_t am_mlp_pairs_per_tile(const am_geometry *g);
int am_mlp_pairs(const float *const *fine, const int32_t *tiles, int32_t n_tiles, const am_geometry *g, float *x,
                 float *t, float *mu, float *sigma, void *stream);

/* The stage-time ghost fill of am_ghost_fill_stage_batched (tricubic,
 * equal_mode as there) with the learned prolongation on top, on every face
 * whose target is exactly one level coarser and not the root (the 24 BSSN
 * fields; older for all fields or none): the coarse cells of such a face's
 * ghost slab form blocks of 64; each cell's ghost children are the MLP's,
 * shifted so they average to the cell's value.  A cell whose 4^3 stencil
 * would read an edge or corner ghost of the target keeps the tricubic
 * value, and so does every cell of a block whose largest restriction
 * residual |mean(children) - value| (field units, before the shift)
 * exceeds `threshold`.  stats (device, 4 uint64, accumulated): learned
 * blocks, blocks kept tricubic, cells kept tricubic for their stencil,
 * cells evaluated. */
int am_ghost_fill_learned(const am_field_batch *b, const am_mesh *mesh, const am_geometry *g, float stage_frac,
                          const float *params, float threshold, uint64_t *stats, int32_t equal_mode, void *stream);
/* The same over a list of faces (slot * 6 + face, as am_ghost_fill_stage_
 * faces) with the gate per tile: the largest restriction residual over
 * every block of a tile's learned faces is gathered in tile_resid[slot]
 * (device, one float per slot of the octree, a workspace the call clears
 * for the tiles it touches); a tile beyond `threshold` keeps the tricubic
 * fill on all its learned faces, a tile within it takes the network's
 * values on all of them.  stats (device, 4 uint64, accumulated): faces kept
 * learned, faces refilled tricubic, cells kept tricubic for their stencil,
 * cells evaluated.  No counterpart in the reference, whose prolongation is
 * the fixed interpolation of AMR.h:1053-1115. */
int am_ghost_fill_learned_faces(const am_field_batch *b, const am_mesh *mesh, const int32_t *faces, int32_t n_faces,
                                const am_geometry *g, float stage_frac, const float *params, float threshold,
                                float *tile_resid, uint64_t *stats, int32_t equal_mode, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ADAPTIVEMESH_H */
