/*
 * adaptivemesh_host.h -- C ABI of the native host runtime: the CPU-managed
 * octree (the reference's AMR<> class, AMR.h:118-294) and the physics
 * plugins driving the kernels of adaptivemesh.h (SimpleWave.h:191-279,
 * BSSN.h:104-171).  Opaque handles; every function returns 0 or an AM_ERR_*
 * / hipError_t code unless it says otherwise.
 */
#ifndef ADAPTIVEMESH_HOST_H
#define ADAPTIVEMESH_HOST_H

#include "adaptivemesh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- octree */
typedef struct am_tree am_tree;

/* AMR(node_slots) (AMR.h:189-200).  periodic != 0 wraps the outer faces
 * instead of pointing them at the root tile.  Returns NULL on failure. */
am_tree *am_tree_create(uint32_t max_slots, int32_t periodic);
void am_tree_destroy(am_tree *t);
/* AMR::add_node (AMR.h:208-232): the new slot (>= 1) or an AM_ERR_* code */
int am_tree_add_node(am_tree *t, int32_t parent, int32_t offset);
/* AMR::remove_node (AMR.h:233-254) */
int am_tree_remove_node(am_tree *t, int32_t node);
/* AMR::increment_timer (AMR.h:180-188) */
int am_tree_increment_timer(am_tree *t);
/* AMR::find_node_idx (AMR.h:255-263) */
uint32_t am_tree_find_node_idx(const am_tree *t, uint32_t packed_pos, uint32_t max_depth);
/* Copy out the host arrays (any pointer may be NULL).  scalars receives
 * {max_depth, first_free_slot, final_used_slot, timer_helper}. */
int am_tree_export(const am_tree *t, int32_t *parent, int32_t *children, uint32_t *positions, uint64_t *boundaries,
                   uint32_t *lifetime, int32_t *scalars);
/* Upload if changed and describe the slots at `depth` (-1 = every used slot)
 * for direct kernel calls.  Needs a GPU. */
int am_tree_mesh(am_tree *t, int32_t depth, am_mesh *out);

/* ------------------------------------------------------------ checkpoints
 * Every driver saves and loads its whole state: the octree, its counter, the
 * host seed and every array its next timestep reads, so that a resumed run
 * continues bit for bit.  The reference only has raw buffer dumps
 * (io_helper.h:9-40, write_to_file / read_from_file); these files add a
 * header and a tag per array.  Layout (little-endian): am_checkpoint_header,
 * then n_sections sections of { char name[16] (NUL-padded); uint64_t bytes;
 * bytes }, in this order:
 *   "parent" int32[max_slots], "children" int32[8 max_slots],
 *   "positions" uint32[max_slots], "boundaries" uint64[6 max_slots],
 *   "lifetime" uint32[max_slots]   -- the octree arrays (AMR.h:118-132);
 *   SimpleWave: "old", "new" float[used T^3] (used = final_used_slot + 1),
 *     "d1", "d2diag" cf3 words uint32[max_slots S^3] (a re-used slot's old
 *     words feed the refinement criterion until it is first differentiated);
 *   RK4 drivers: "old", "new" and, with tricubic or learned ghosts, "older":
 *     float[n_vars][used T^3], variable-major.
 * A load must go to a driver created with the same geometry, max_slots,
 * periodic flag and options encoded in `flags`; it checks the whole file
 * (header, section names and sizes, tree consistency) before changing
 * anything, and returns AM_ERR_FORMAT or AM_ERR_IO otherwise.  Physics
 * options that `flags` does not encode (BSSN gauge parameters, learned
 * prolongation weights) are the caller's to set alike.  A save writes
 * "<path>.tmp" and renames it over <path>. */
#define AM_CHECKPOINT_MAGIC "AMCKPT\0\1"
enum am_checkpoint_kind { AM_CKPT_WAVE = 1, AM_CKPT_BSSN = 2, AM_CKPT_SWAVE = 3 };
typedef struct am_checkpoint_header {
    char magic[8];            /* AM_CHECKPOINT_MAGIC, 8 bytes */
    uint32_t kind;            /* am_checkpoint_kind */
    am_geometry geometry;
    uint32_t max_slots;
    uint32_t periodic;
    int32_t max_depth;        /* AMR::max_depth */
    uint32_t first_free_slot; /* AMR::first_free_slot */
    uint32_t final_used_slot; /* AMR::final_used_slot */
    uint32_t timer;           /* AMR::timer_helper, the sub-cycling counter */
    int32_t seed;             /* host fast_prng state */
    uint32_t n_vars;          /* fields per slot */
    uint32_t flags;           /* SimpleWave: 1 = regridding on; RK4: interp | word_format << 4 */
    uint32_t n_sections;
} am_checkpoint_header;

/* Check a checkpoint file on its own (section table, octree consistency,
 * state sections for its kind) and copy its header to *out (may be NULL).
 * Host only: needs no GPU. */
int am_checkpoint_inspect(const char *path, am_checkpoint_header *out);

/* ------------------------------------------------------------------ images
 * lodepng_encode32_file as save_image calls it (image_process.h:55-62):
 * width * height RGBA8 pixels (host, R in the low byte) to a PNG file. */
int am_write_png(const char *path, const uint32_t *pixels, uint32_t width, uint32_t height);

/* ------------------------------------------------------ SimpleWave driver */
typedef struct am_wave_sim am_wave_sim;

/* AMR<wave_AMR_data>(node_slots) with the given geometry.  NULL on failure. */
am_wave_sim *am_wave_sim_create(const am_geometry *g, uint32_t max_slots, int32_t periodic);
void am_wave_sim_destroy(am_wave_sim *s);
int am_wave_sim_add_node(am_wave_sim *s, int32_t parent, int32_t offset);
int am_wave_sim_export(const am_wave_sim *s, int32_t *parent, int32_t *children, uint32_t *positions,
                       uint64_t *boundaries, uint32_t *lifetime, int32_t *scalars);
/* which: 0 = old_field, 1 = new_field; whole padded arrays, max_slots * T^3 */
int am_wave_sim_set_field(am_wave_sim *s, int32_t which, const float *host);
int am_wave_sim_get_field(am_wave_sim *s, int32_t which, float *host);
/* wave_AMR_data phases (SimpleWave.h:214-278) */
int am_wave_sim_copy_bounds(am_wave_sim *s);
int am_wave_sim_timestep(am_wave_sim *s);
int32_t am_wave_sim_rng_state(const am_wave_sim *s);
/* modify_nodes on (1, the default, as in the reference) or off (the mesh
 * stays as built) */
int am_wave_sim_set_regrid(am_wave_sim *s, int32_t on);
/* the two average refinement pressures of the last modify_nodes (per-slot
 * maxima; octants / 8), as amr_criterion returns them (SimpleWave.h:183) */
int am_wave_sim_pressure(const am_wave_sim *s, float *out2);
int am_wave_sim_save(am_wave_sim *s, const char *path);
/* init_temp (kernel.cu:34-43) on both fields: am_init_temp over every used slot */
int am_wave_sim_init_temp(am_wave_sim *s, float mul);
/* rasterise_to_image + the encoding half of save_image (AMR.h:1355-1364,
 * image_process.h:55-62) on field `which` (0 old, 1 new): width * height
 * RGBA8 pixels into host memory; am_write_png stores them. */
int am_wave_sim_raster(am_wave_sim *s, int32_t which, uint32_t width, uint32_t height, float z_slice, int32_t boxes,
                       uint32_t *pixels);
int am_wave_sim_load(am_wave_sim *s, const char *path);

/* ------------------------------------------- RK4 driver launch machinery
 * Shared by am_bssn_sim and am_swave_sim (no counterpart in the reference):
 *   _set_launch  FD planes per workgroup (0 = whole tile, else S, S/2, S/4)
 *                and workgroup caps of persistent FD / RHS launches (0 =
 *                one workgroup per work item);
 *   _use_graph   timestep() records the coarse step once as a hipGraph and
 *                replays it (results identical to eager launches: the dither
 *                seeds are refilled on the device before every replay);
 *   _autotune    times whole coarse steps under each launch shape (one
 *                knob at a time, every trial from the same snapshot), keeps
 *                the fastest and restores the state, counter, seed and
 *                lifetimes, so the run continues unchanged. */
typedef struct am_tune_report {
    int32_t z_block, fd_grid_limit, rhs_grid_limit, n_trials;
    struct {
        int32_t kind; /* 0 FD z_block, 1 FD workgroup cap, 2 RHS workgroup cap, 3 FD prefetch planes,
                       * 4 direct-face axis mask, 5 overlap of coarser depths */
        int32_t value;
        float ms; /* best of the repetitions, one coarse step */
    } trial[32];
    int32_t fd_prefetch;
    int32_t direct_faces;
    int32_t overlap_depths;
} am_tune_report;

/* --------------------------------------------- BSSN RK4 evolution driver
 * AMR<BSSN_rk4_data>: the 24-field moving-puncture BSSN system advanced by
 * classical RK4 with sub-cycling (no counterpart in the reference, whose
 * BSSN_AMR_data computes derivatives only, BSSN.h:104-171). */
typedef struct am_bssn_sim am_bssn_sim;

typedef struct am_bssn_sim_options {
    am_bssn_params params;
    int32_t interp;   /* AM_INTERP_LAGRANGE (reference) or AM_INTERP_TRICUBIC (+ quadratic time) */
    int32_t z_block;  /* FD workgroup z extent (0 = whole tile) */
    int32_t periodic; /* outer faces wrap instead of pointing at the root */
    int32_t seed;     /* host fast_prng seed for the dither draws */
    /* enum am_word_format of the derivative words: AM_WORD_CF3 (the
     * reference's format), or AM_WORD_F32 / AM_WORD_F16 (precision study,
     * S=16, P=3, fd_order 4 only) */
    int32_t word_format;
} am_bssn_sim_options;

am_bssn_sim *am_bssn_sim_create(const am_geometry *g, uint32_t max_slots, const am_bssn_sim_options *o);
void am_bssn_sim_destroy(am_bssn_sim *s);
int am_bssn_sim_add_node(am_bssn_sim *s, int32_t parent, int32_t offset);
int am_bssn_sim_export(const am_bssn_sim *s, int32_t *parent, int32_t *children, uint32_t *positions,
                       uint64_t *boundaries, uint32_t *lifetime, int32_t *scalars);
/* which: 0 old, 1 new, 2 acc, 3 older; var 0..23; max_slots * T^3 floats on
 * the host (acc is interior-only on the device: its ghost cells read as 0) */
int am_bssn_sim_set_field(am_bssn_sim *s, int32_t which, int32_t var, const float *host);
int am_bssn_sim_get_field(am_bssn_sim *s, int32_t which, int32_t var, float *host);
/* device addresses: which 0, 1, 3 as above, in the interior-first tile
 * layout (var stride max_slots * am_tile_stride(g, AM_LAYOUT_INTERIOR_FIRST)
 * floats), 2 acc (var stride max_slots*S^3 floats), 4 d1, 5 d2diag, 6 d2off
 * words (var stride max_slots*S^3); 7 Christoffel words where a driver keeps
 * them (none: the BSSN RHS raises the symbols in registers); 8 the upwinded
 * advection + dissipation (var stride max_slots*S^3 floats; after
 * am_bssn_sim_set_dissipation) */
void *am_bssn_sim_device_ptr(am_bssn_sim *s, int32_t which);
int am_bssn_sim_fill_ghosts(am_bssn_sim *s);
int am_bssn_sim_copy_back(am_bssn_sim *s);
int am_bssn_sim_step_depth(am_bssn_sim *s, int32_t depth);
/* one RK4 stage (0-3) of every tile at `depth`: FD and Christoffel words of
 * the stage input, RHS + update, ghost fill at the stage time */
int am_bssn_sim_stage_depth(am_bssn_sim *s, int32_t depth, int32_t stage);
int am_bssn_sim_substep(am_bssn_sim *s);
int am_bssn_sim_timestep(am_bssn_sim *s);
/* {rms H, rms |M|, max |H|, cells} over the leaves (FP64) */
int am_bssn_sim_constraints(am_bssn_sim *s, double *out4);
int am_bssn_sim_sync(am_bssn_sim *s);
uint64_t am_bssn_sim_zone_cycles(const am_bssn_sim *s);
int32_t am_bssn_sim_rng_state(const am_bssn_sim *s);
int am_bssn_sim_set_rng_state(am_bssn_sim *s, int32_t seed);
/* the counter and max depth (timer_helper, read_max_depth) */
int am_bssn_sim_counter(const am_bssn_sim *s, uint32_t *timer, int32_t *max_depth);
/* per-launch HIP-event timing of the driver's kernels (fd, christoffel,
 * rhs_stage, ghost, copy_back); kernel_times returns total ms and launch
 * counts per class since profiling was (re)enabled */
int am_bssn_sim_profile(am_bssn_sim *s, int32_t enable);
int am_bssn_sim_kernel_times(am_bssn_sim *s, double *ms5, int64_t *count5);
int am_bssn_sim_set_launch(am_bssn_sim *s, int32_t z_block, int32_t fd_grid_limit, int32_t rhs_grid_limit);
/* out4: z_block, fd_grid_limit, rhs_grid_limit, fd_prefetch */
int am_bssn_sim_get_launch(const am_bssn_sim *s, int32_t *out4);
/* FD planes prefetched ahead of the differentiated plane (1-3) */
int am_bssn_sim_set_fd_prefetch(am_bssn_sim *s, int32_t planes);
/* Upwinded advection with Kreiss-Oliger dissipation (no counterpart in the
 * reference; for moving-puncture runs): upwind = 1 replaces the centred
 * advection beta^k d_k of every field by the fourth-order lopsided stencil
 * on the side the shift points to and adds sixth-order dissipation of
 * strength ko_sigma (>= 0) to every field (am_fd_batch.adv); the stage
 * kernels then push all P ghost layers.  upwind = 0, ko_sigma = 0 restores
 * the centred scheme.  Needs S=16, P=3, fd_order 4, compressed_float3
 * words; refused (AM_ERR_INVALID) otherwise.  A checkpoint records upwind
 * in its flags. */
int am_bssn_sim_set_dissipation(am_bssn_sim *s, float ko_sigma, int32_t upwind);
int am_bssn_sim_get_dissipation(const am_bssn_sim *s, float *ko_sigma, int32_t *upwind);
/* Equal-depth ghost layers the stage kernels refresh, innermost first:
 * 0 (the default) = fd_order / 2, the layers the FD stencils read; P =
 * every layer, as the reference's copy_boundaries writes them.  Layers
 * beyond keep what they held.  A value from 1 to fd_order / 2 - 1 is
 * refused (AM_ERR_INVALID): the stencils would read stale ghost layers.
 * _get_push_layers gives the layers in use. */
int am_bssn_sim_set_push_layers(am_bssn_sim *s, int32_t layers);
int am_bssn_sim_get_push_layers(const am_bssn_sim *s);
/* Direct faces: the axes (bits 1 x, 2 y, 4 z; default 4) whose equal-depth
 * faces stages 0-2 do not push; the next stage's FD reads those ghost cells
 * from the neighbours' interiors.  Stage 3 pushes every face.  The fields
 * after each fine step are the same for every mask; between stages 0-2 the
 * equal-depth ghost layers on the masked axes are not current.  The mask is
 * 0, 2, 4 or 6 (x ghost columns would be strided reads).
 * _get_direct_faces gives the mask. */
int am_bssn_sim_set_direct_faces(am_bssn_sim *s, int32_t axes);
int am_bssn_sim_get_direct_faces(const am_bssn_sim *s);
/* Overlap of depths (default 1): the copy-back and the coarser active depths
 * of a sub-step run on the driver's second stream beside the finest depth
 * (their kernels touch only their own tiles and equal-depth neighbours); the
 * finest depth's first write of new waits for the copy-back, its first
 * coarse-face fill for the coarser depths, and the two streams join before
 * the sub-step ends.  The results are those of one stream bit for bit.  With
 * per-kernel timing on, or 0 here, every launch runs on one stream. */
int am_bssn_sim_set_overlap_depths(am_bssn_sim *s, int32_t enable);
int am_bssn_sim_get_overlap_depths(const am_bssn_sim *s);
int am_bssn_sim_use_graph(am_bssn_sim *s, int32_t enable);
int am_bssn_sim_autotune(am_bssn_sim *s, int32_t reps, am_tune_report *out);
/* AM_INTERP_LEARNED: the MLP parameters (device, am_mlp_param_count()
 * floats, owned by the caller and kept alive while the driver uses them) and
 * the restriction-residual threshold of am_ghost_fill_learned; resets the
 * statistics, which _prolongation_stats copies out (4 counters, see
 * am_ghost_fill_learned). */
int am_bssn_sim_set_prolongation(am_bssn_sim *s, const float *params, float threshold);
int am_bssn_sim_prolongation_stats(am_bssn_sim *s, uint64_t *out4);
/* checkpoints (see am_checkpoint_header) */
int am_bssn_sim_save(am_bssn_sim *s, const char *path);
int am_bssn_sim_load(am_bssn_sim *s, const char *path);
/* as am_wave_sim_raster, on variable `var` of field `which` (0 old, 1 new) */
int am_bssn_sim_raster(am_bssn_sim *s, int32_t which, int32_t var, uint32_t width, uint32_t height, float z_slice,
                       int32_t boxes, uint32_t *pixels);

/* -------------------------------------------- scalar wave RK4 driver
 * AMR<swave_rk4_data>: u_tt = lap u as fields (u, v = u_t) on the same RK4
 * sub-cycling driver as the BSSN system (no counterpart in the reference:
 * its SimpleWave.h plugin is the Heun Burgers step behind am_wave_sim).
 * Every call means what the am_bssn_sim call of the same name means; var is
 * 0 (u) or 1 (v); device_ptr(4) is u's gradient word, (5) its d2diag word. */
typedef struct am_swave_sim am_swave_sim;

typedef struct am_swave_sim_options {
    int32_t interp;   /* AM_INTERP_LAGRANGE or AM_INTERP_TRICUBIC */
    int32_t z_block;  /* FD workgroup z extent (0 = whole tile) */
    int32_t periodic; /* outer faces wrap */
    int32_t seed;     /* host fast_prng seed for the dither draws */
} am_swave_sim_options;

am_swave_sim *am_swave_sim_create(const am_geometry *g, uint32_t max_slots, const am_swave_sim_options *o);
void am_swave_sim_destroy(am_swave_sim *s);
int am_swave_sim_add_node(am_swave_sim *s, int32_t parent, int32_t offset);
int am_swave_sim_export(const am_swave_sim *s, int32_t *parent, int32_t *children, uint32_t *positions,
                        uint64_t *boundaries, uint32_t *lifetime, int32_t *scalars);
int am_swave_sim_set_field(am_swave_sim *s, int32_t which, int32_t var, const float *host);
int am_swave_sim_get_field(am_swave_sim *s, int32_t which, int32_t var, float *host);
void *am_swave_sim_device_ptr(am_swave_sim *s, int32_t which);
int am_swave_sim_fill_ghosts(am_swave_sim *s);
int am_swave_sim_copy_back(am_swave_sim *s);
int am_swave_sim_step_depth(am_swave_sim *s, int32_t depth);
int am_swave_sim_stage_depth(am_swave_sim *s, int32_t depth, int32_t stage);
int am_swave_sim_substep(am_swave_sim *s);
int am_swave_sim_timestep(am_swave_sim *s);
/* {energy = sum (v^2 + |grad u|^2) h^3 / 2, sum u^2 h^3, max |u|, cells}
 * over the leaves (FP64; grad u from dithered gradient words whose draws are
 * given back, so the trajectory does not depend on how often it is called) */
int am_swave_sim_energy(am_swave_sim *s, double *out4);
int am_swave_sim_sync(am_swave_sim *s);
uint64_t am_swave_sim_zone_cycles(const am_swave_sim *s);
int32_t am_swave_sim_rng_state(const am_swave_sim *s);
int am_swave_sim_set_rng_state(am_swave_sim *s, int32_t seed);
int am_swave_sim_counter(const am_swave_sim *s, uint32_t *timer, int32_t *max_depth);
int am_swave_sim_profile(am_swave_sim *s, int32_t enable);
int am_swave_sim_kernel_times(am_swave_sim *s, double *ms5, int64_t *count5);
int am_swave_sim_set_launch(am_swave_sim *s, int32_t z_block, int32_t fd_grid_limit, int32_t rhs_grid_limit);
int am_swave_sim_get_launch(const am_swave_sim *s, int32_t *out4);
int am_swave_sim_set_fd_prefetch(am_swave_sim *s, int32_t planes);
int am_swave_sim_set_push_layers(am_swave_sim *s, int32_t layers);
int am_swave_sim_get_push_layers(const am_swave_sim *s);
int am_swave_sim_set_direct_faces(am_swave_sim *s, int32_t axes);
int am_swave_sim_get_direct_faces(const am_swave_sim *s);
int am_swave_sim_set_overlap_depths(am_swave_sim *s, int32_t enable);
int am_swave_sim_get_overlap_depths(const am_swave_sim *s);
int am_swave_sim_use_graph(am_swave_sim *s, int32_t enable);
int am_swave_sim_autotune(am_swave_sim *s, int32_t reps, am_tune_report *out);
int am_swave_sim_save(am_swave_sim *s, const char *path);
int am_swave_sim_load(am_swave_sim *s, const char *path);
int am_swave_sim_raster(am_swave_sim *s, int32_t which, int32_t var, uint32_t width, uint32_t height, float z_slice,
                        int32_t boxes, uint32_t *pixels);
/* AM_INTERP_LEARNED: the MLP parameters (device, am_mlp_param_count()
 * floats, owned by the caller and kept alive while the driver uses them) and
 * the restriction-residual threshold of am_ghost_fill_learned; resets the
 * statistics, which _prolongation_stats copies out (4 counters, see
 * am_ghost_fill_learned). */
int am_swave_sim_set_prolongation(am_swave_sim *s, const float *params, float threshold);
int am_swave_sim_prolongation_stats(am_swave_sim *s, uint64_t *out4);

#ifdef __cplusplus
}
#endif

#endif /* ADAPTIVEMESH_HOST_H */
