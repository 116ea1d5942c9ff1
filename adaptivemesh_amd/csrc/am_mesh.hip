// Tile-to-tile kernels for gfx950: ghost-cell fill with space-time
// interpolation, restriction on copy-back, and prolongation into a new child.
//
// Ghost fill: one workgroup per (tile, face, variable); its S*S threads cover
// the face and loop over the P ghost layers.  The source cells are a small
// neighbourhood of the boundary target tile, re-read by adjacent ghost cells,
// so the gathers hit L1/L2; the writes are the face slab of this tile.
// Restriction: one workgroup per (tile, z-plane, variable) streaming the
// contiguous T*T plane.
//
// Index maps and arithmetic follow AMR.h:673-762 (quadratic), 767-848
// (trilinear), 961-996 and 1053-1073 term by term with contraction off, so
// results are bit-identical to the reference's on an IEEE host.  The
// tricubic-space / quadratic-time mode is an extension.
#include "am_device.hpp"
#include "am_interp.hpp"

#pragma clang fp contract(off)

namespace am {

template <int S, int P, int MODE>
__global__ void __launch_bounds__(S *S) k_ghost_fill(FieldBatchDev b, const uint64_t *__restrict__ bnd,
                                                     const int32_t *__restrict__ tiles, uint32_t substep,
                                                     uint32_t maxd, int copy_only_new, float frac)
{
    constexpr int T = S + 2 * P;
    constexpr size_t CELLS = (size_t)T * T * T;
    const int ti = blockIdx.x / 6, dir = blockIdx.x - ti * 6;
    const int var = blockIdx.y;
    const int node = tiles[ti];
    const uint64_t bb = bnd[(size_t)node * 6 + dir];
    const uint32_t nd = (uint32_t)bb;
    const int tgt = bnd_target(nd), cd = bnd_depth(nd);
    if (tgt == -1 || cd <= 0) return;
    if (!active_depth(substep, (uint32_t)cd, maxd)) return;

    const float3 relp = cf3_unpack((uint32_t)(bb >> 32));
    const uint32_t ddiff = bnd_ddiff(nd);
    const float rel_scale = 1.f / (float)(1u << ddiff);
    // relative_time (AMR.h:634-639)
    const float tmp = (float)((substep >> (maxd - (uint32_t)cd)) & ((1u << ddiff) - 1u));
    // frac = 1 is the reference's k + 1; an equal-depth target holds the same stage
    const float tx = tmp / (float)(1u << ddiff), ty = ddiff == 0 ? 1.f : (tmp + frac) / (float)(1u << ddiff);

    float *__restrict__ dold = b.old_[var];
    float *__restrict__ dnew = b.new_[var];
    const float *__restrict__ dolder = b.older[var];
    const uint32_t a = threadIdx.x % S, bq = threadIdx.x / S;
    constexpr uint32_t uP = P, uS = S;
    const float invS = 1.f / (float)S;

    for (uint32_t l = 0; l < uP; l++) {
        uint32_t w0, w1, w2;
        switch (dir) {  // zxy / xzy swizzles of (a, b, l) (AMR.h:714-735)
        case 0: w0 = l + uP + uS; w1 = a + uP; w2 = bq + uP; break;
        case 1: w0 = a + uP; w1 = l + uP + uS; w2 = bq + uP; break;
        case 2: w0 = a + uP; w1 = bq + uP; w2 = l + uP + uS; break;
        case 3: w0 = l; w1 = a + uP; w2 = bq + uP; break;
        case 4: w0 = a + uP; w1 = l; w2 = bq + uP; break;
        default: w0 = a + uP; w1 = bq + uP; w2 = l; break;
        }
        float rp[3] = {(float)w0, (float)w1, (float)w2};
        const float rl[3] = {relp.x, relp.y, relp.z};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            rp[k] = (rp[k] - (float)uP + .5f) * invS - .5f;
            rp[k] = rp[k] * rel_scale + rl[k];
            rp[k] = (rp[k] + .5f) * (float)uS + (float)uP - 0.5f;
        }
        uint32_t ri[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (MODE == AM_INTERP_TRILINEAR)
                ri[k] = (tgt == 0) ? min((uint32_t)fmaxf(rp[k], .5f), (uint32_t)T - 2u)
                                   : min((uint32_t)fmaxf(rp[k], .5f + (float)uP), (uint32_t)T - uP - 2u);
            else if (MODE == AM_INTERP_TRICUBIC)  // 4 nodes ri-1..ri+2, interior-only on non-root targets
                ri[k] = (tgt == 0) ? min((uint32_t)fmaxf(rp[k], 1.f), (uint32_t)T - 3u)
                                   : min((uint32_t)fmaxf(rp[k], (float)(uP + 1u)), (uint32_t)T - uP - 3u);
            else
                ri[k] = (tgt == 0) ? min((uint32_t)(fmaxf(rp[k], 0.f) + .5f), (uint32_t)T - 2u)
                                   : min((uint32_t)(fmaxf(rp[k], (float)uP) + .5f), (uint32_t)T - uP - 2u);
        }
        const size_t rn = (size_t)tgt * CELLS + ri[0] + ((size_t)ri[1] + (size_t)ri[2] * T) * T;
#pragma unroll
        for (int k = 0; k < 3; k++) rp[k] -= (float)ri[k];
        const size_t wn = (size_t)node * CELLS + w0 + ((size_t)w1 + (size_t)w2 * T) * T;

        float o, nw;
        if (MODE == AM_INTERP_TRICUBIC) {
            float wx[4], wy[4], wz[4];
            lag4(rp[0], wx); lag4(rp[1], wy); lag4(rp[2], wz);
            o = interp_tricubic<T>(dold, wx, wy, wz, rn);
            nw = interp_tricubic<T>(dnew, wx, wy, wz, rn);
            if (dolder && ddiff > 0) {
                const float od = interp_tricubic<T>(dolder, wx, wy, wz, rn);
                float s = ty;
                float c0 = .5f * s * (s - 1.f), c1 = (1.f - s) * (1.f + s), c2 = .5f * s * (s + 1.f);
                const float vnew = od * c0 + o * c1 + nw * c2;
                if (!copy_only_new) {
                    s = tx;
                    c0 = .5f * s * (s - 1.f); c1 = (1.f - s) * (1.f + s); c2 = .5f * s * (s + 1.f);
                    dold[wn] = od * c0 + o * c1 + nw * c2;
                }
                dnew[wn] = vnew;
                continue;
            }
        } else if (MODE == AM_INTERP_TRILINEAR) {
            o = interp_trilinear<T>(dold, rp[0], rp[1], rp[2], rn);
            nw = interp_trilinear<T>(dnew, rp[0], rp[1], rp[2], rn);
        } else {
            o = interp_lagrange<T>(dold, rp[0], rp[1], rp[2], rn);
            nw = interp_lagrange<T>(dnew, rp[0], rp[1], rp[2], rn);
        }
        if (!copy_only_new) dold[wn] = o * (1.f - tx) + nw * tx;
        dnew[wn] = o * (1.f - ty) + nw * ty;
    }
}

// The face a workgroup of the multi-field fills handles: blockIdx.x is a
// (tile list entry, face) pair, or, with a face list, an entry slot * 6 +
// face of it.
__device__ __forceinline__ void face_of_block(const int32_t *__restrict__ tiles, const int32_t *__restrict__ faces,
                                              int &node, int &dir)
{
    if (faces) {
        const int code = faces[blockIdx.x];
        node = code / 6;
        dir = code - node * 6;
    } else {
        const int ti = blockIdx.x / 6;
        dir = blockIdx.x - ti * 6;
        node = tiles[ti];
    }
}

// Ghost fill for many fields at once (the RK4 driver's stage fills).  One
// workgroup per (tile, face, group of VPG fields); each thread owns ghost
// cells of the face slab, laid out so consecutive lanes follow the slab's
// contiguous axis, computes the index map of AMR.h:714-753 once per cell and
// applies it to every field of the group.  Equal-depth cells whose read
// position falls exactly on a target cell are copies: the quadratic formula
// there is f plus terms that are all +-0, and the time weight on `new` is 1,
// so the copy is the formula's value for finite data.  Everything else runs
// the full formula of k_ghost_fill.
template <int S, int P, int MODE, int VPG, int L>
__global__ void __launch_bounds__(256) k_ghost_fill_multi(FieldBatchDev b, int n_vars, const uint64_t *__restrict__ bnd,
                                                          const int32_t *__restrict__ tiles,
                                                          const int32_t *__restrict__ faces, uint32_t substep,
                                                          uint32_t maxd, float frac, int equal_mode, int skip_learned)
{
    constexpr int T = S + 2 * P, NC = S * S * P;
    constexpr size_t CELLS = tile_stride<S, P, L>();
    int node, dir;
    face_of_block(tiles, faces, node, dir);
    const int v0 = blockIdx.y * VPG;
    const uint64_t bb = bnd[(size_t)node * 6 + dir];
    const uint32_t nd = (uint32_t)bb;
    const int tgt = bnd_target(nd), cd = bnd_depth(nd);
    if (tgt == -1 || cd <= 0) return;
    if (!active_depth(substep, (uint32_t)cd, maxd)) return;
    const float3 relp = cf3_unpack((uint32_t)(bb >> 32));
    const uint32_t ddiff = bnd_ddiff(nd);
    if (ddiff == 0 && equal_mode == AM_EQUAL_SKIP) return;
    if (skip_learned && ddiff == 1 && tgt > 0) return;  // the learned fill's faces
    const float rel_scale = 1.f / (float)(1u << ddiff);
    const float tmp = (float)((substep >> (maxd - (uint32_t)cd)) & ((1u << ddiff) - 1u));
    const float ty = ddiff == 0 ? 1.f : (tmp + frac) / (float)(1u << ddiff);
    constexpr uint32_t uP = P, uS = S;
    const float invS = 1.f / (float)S;
    const int nv = min(VPG, n_vars - v0);

    for (int c = threadIdx.x; c < NC; c += 256) {
        // consecutive lanes follow the slab's contiguous axis: x for the
        // padded layout, the in-face axis a for the interior-first slabs
        uint32_t w0, w1, w2;
        if constexpr (L == 1) {
            const uint32_t a = c % S, bq = (c / S) % S, l = c / (S * S);
            const uint32_t g = dir < 3 ? l + uP + uS : uP - 1u - l;
            if (dir % 3 == 0) {
                w0 = g; w1 = a + uP; w2 = bq + uP;
            } else if (dir % 3 == 1) {
                w0 = a + uP; w1 = g; w2 = bq + uP;
            } else {
                w0 = a + uP; w1 = bq + uP; w2 = g;
            }
        } else if (dir == 0 || dir == 3) {  // x faces: x fastest
            const uint32_t l = c % P, a = (c / P) % S, bq = c / (P * S);
            w0 = dir == 0 ? l + uP + uS : l; w1 = a + uP; w2 = bq + uP;
        } else if (dir == 1 || dir == 4) {
            const uint32_t a = c % S, l = (c / S) % P, bq = c / (S * P);
            w0 = a + uP; w1 = dir == 1 ? l + uP + uS : l; w2 = bq + uP;
        } else {
            const uint32_t a = c % S, bq = (c / S) % S, l = c / (S * S);
            w0 = a + uP; w1 = bq + uP; w2 = dir == 2 ? l + uP + uS : l;
        }
        float rp[3] = {(float)w0, (float)w1, (float)w2};
        const float rl[3] = {relp.x, relp.y, relp.z};
#pragma unroll
        for (int k = 0; k < 3; k++) {
            rp[k] = (rp[k] - (float)uP + .5f) * invS - .5f;
            rp[k] = rp[k] * rel_scale + rl[k];
            rp[k] = (rp[k] + .5f) * (float)uS + (float)uP - 0.5f;
        }
        const size_t wn = (size_t)node * CELLS + tile_cell<S, P, L>(w0, w1, w2);
        if (ddiff == 0 && equal_mode == AM_EQUAL_COPY) {
            // the neighbour cell the ghost maps onto: rp is a whole cell index,
            // exactly for power-of-two S and within rounding otherwise (1/S
            // inexact), so it is rounded, not truncated
            const size_t src = (size_t)tgt * CELLS + tile_cell<S, P, L>((uint32_t)(rp[0] + .5f), (uint32_t)(rp[1] + .5f),
                                                                         (uint32_t)(rp[2] + .5f));
            for (int v = 0; v < nv; v++) b.new_[v0 + v][wn] = b.new_[v0 + v][src];
            continue;
        }
        uint32_t ri[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            if (MODE == AM_INTERP_TRILINEAR)
                ri[k] = (tgt == 0) ? min((uint32_t)fmaxf(rp[k], .5f), (uint32_t)T - 2u)
                                   : min((uint32_t)fmaxf(rp[k], .5f + (float)uP), (uint32_t)T - uP - 2u);
            else if (MODE == AM_INTERP_TRICUBIC)
                ri[k] = (tgt == 0) ? min((uint32_t)fmaxf(rp[k], 1.f), (uint32_t)T - 3u)
                                   : min((uint32_t)fmaxf(rp[k], (float)(uP + 1u)), (uint32_t)T - uP - 3u);
            else
                ri[k] = (tgt == 0) ? min((uint32_t)(fmaxf(rp[k], 0.f) + .5f), (uint32_t)T - 2u)
                                   : min((uint32_t)(fmaxf(rp[k], (float)uP) + .5f), (uint32_t)T - uP - 2u);
        }
        const Cells<S, P, L> cn{(size_t)tgt * CELLS, ri[0], ri[1], ri[2]};
#pragma unroll
        for (int k = 0; k < 3; k++) rp[k] -= (float)ri[k];
        if (ddiff == 0 && MODE == AM_INTERP_LAGRANGE) {
            // Equal depth: every offset is exactly 0, or 1 where the read index
            // was clamped, and the time weights are (0, 1).  Terms with a zero
            // weight are +-0 and old enters only as old * 0, so evaluating the
            // remaining terms in the formula's order gives its value.
            const bool cx = rp[0] != 0.f, cy = rp[1] != 0.f, cz = rp[2] != 0.f;
            for (int v = 0; v < nv; v++) {
                const float *d = b.new_[v0 + v];
                const float f = d[cn.at(0, 0, 0)];
                float r = f;
                float fx = 0.f, fy = 0.f, fz = 0.f;  // raw +1 neighbours where needed
                if (cx) {
                    fx = d[cn.at(1, 0, 0)];
                    const float fm = d[cn.at(-1, 0, 0)];
                    float gx = (fx - fm) * .5f;
                    const float gxx = (gx + fm) - f;
                    r = r + (gx + gxx * rp[0]) * rp[0];
                }
                if (cy) {
                    fy = d[cn.at(0, 1, 0)];
                    const float fm = d[cn.at(0, -1, 0)];
                    float gy = (fy - fm) * .5f;
                    const float gyy = (gy + fm) - f;
                    r = r + (gy + gyy * rp[1]) * rp[1];
                }
                if (cz) {
                    fz = d[cn.at(0, 0, 1)];
                    const float fm = d[cn.at(0, 0, -1)];
                    float gz = (fz - fm) * .5f;
                    const float gzz = (gz + fm) - f;
                    r = r + (gz + gzz * rp[2]) * rp[2];
                }
                if (cx && cy) r = r + (d[cn.at(1, 1, 0)] + ((f - fx) - fy)) * rp[0] * rp[1];
                if (cx && cz) r = r + (d[cn.at(1, 0, 1)] + ((f - fx) - fz)) * rp[0] * rp[2];
                if (cy && cz) r = r + (d[cn.at(0, 1, 1)] + ((f - fy) - fz)) * rp[1] * rp[2];
                b.new_[v0 + v][wn] = r;
            }
            continue;
        }
        if (ddiff == 0 && MODE == AM_INTERP_TRICUBIC && rp[0] == truncf(rp[0]) && rp[1] == truncf(rp[1]) &&
            rp[2] == truncf(rp[2])) {
            // equal depth with whole offsets (every cell for power-of-two S):
            // the 4-point weights are a unit vector and the value is the
            // target cell at that offset.  Otherwise (1/S inexact) the offsets
            // are a rounding away from whole and the formula below applies.
            const size_t src = cn.at((int)rp[0], (int)rp[1], (int)rp[2]);
            for (int v = 0; v < nv; v++) b.new_[v0 + v][wn] = b.new_[v0 + v][src];
            continue;
        }
        float wx[4], wy[4], wz[4];
        if (MODE == AM_INTERP_TRICUBIC) {
            lag4(rp[0], wx); lag4(rp[1], wy); lag4(rp[2], wz);
            for (int v = 0; v < nv; v++)
                b.new_[v0 + v][wn] =
                    tricubic_time_c<S, P, L>(b.old_[v0 + v], b.new_[v0 + v], b.older[v0 + v], ddiff, ty, wx, wy, wz, cn);
            continue;
        }
        for (int v = 0; v < nv; v++) {
            const float *dold = b.old_[v0 + v], *dnew = b.new_[v0 + v];
            float o, nw;
            if (MODE == AM_INTERP_TRILINEAR) {
                o = interp_trilinear_c<S, P, L>(dold, cn, rp[0], rp[1], rp[2]);
                nw = interp_trilinear_c<S, P, L>(dnew, cn, rp[0], rp[1], rp[2]);
            } else {
                o = interp_lagrange_c<S, P, L>(dold, cn, rp[0], rp[1], rp[2]);
                nw = interp_lagrange_c<S, P, L>(dnew, cn, rp[0], rp[1], rp[2]);
            }
            b.new_[v0 + v][wn] = o * (1.f - ty) + nw * ty;
        }
    }
}

// One job of the tricubic coarse-face fill: the 4^3 target nodes of each
// time level (row(lv, j, k, r) gives the 4 nodes along x of row j, plane k of
// level lv: 0 older, 1 old, 2 new), contracted along x, then y, then z for
// up to 2 x 2 x 2 children with the weights of their fine indices (a, nn:
// run start and length per axis), then combined in time: quadratic through
// older/old/new when `three`, else linear between old and new.
template <int S, class Row>
__device__ __forceinline__ void tricubic_children(const Row &row, bool three, const float (&s_w)[3][S][4],
                                                  const int (&a)[3], const int (&nn)[3], float ty, float c0, float c1,
                                                  float c2, float (&val)[2][2][2])
{
    float wx[2][4], wy[2][4], wz[2][4];
#pragma unroll
    for (int c = 0; c < 2; c++)
#pragma unroll
        for (int m = 0; m < 4; m++) {
            wx[c][m] = s_w[0][a[0] + min(c, nn[0] - 1)][m];
            wy[c][m] = s_w[1][a[1] + min(c, nn[1] - 1)][m];
            wz[c][m] = s_w[2][a[2] + min(c, nn[2] - 1)][m];
        }
    // time combination accumulated level by level, in the order of
    // older * c0 + old * c1 + new * c2 (or old * (1 - ty) + new * ty)
    float acc[2][2][2];
    const float coef[3] = {c0, three ? c1 : 1.f - ty, three ? c2 : ty};
#pragma unroll 1
    for (int lv = three ? 0 : 1; lv < 3; lv++) {
        float ax[2][4][4];  // child x, k, j
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int j = 0; j < 4; j++) {
                float r[4];
                row(lv, j, k, r);
#pragma unroll
                for (int cx = 0; cx < 2; cx++)
                    ax[cx][k][j] = r[0] * wx[cx][0] + r[1] * wx[cx][1] + r[2] * wx[cx][2] + r[3] * wx[cx][3];
            }
        const float cl = coef[lv];
        const bool first = lv == (three ? 0 : 1);
#pragma unroll
        for (int cx = 0; cx < 2; cx++)
#pragma unroll
            for (int cy = 0; cy < 2; cy++) {
                float accz[2] = {0.f, 0.f};
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    float ay = 0.f;
#pragma unroll
                    for (int j = 0; j < 4; j++) ay = ay + ax[cx][k][j] * wy[cy][j];
#pragma unroll
                    for (int cz = 0; cz < 2; cz++) accz[cz] = accz[cz] + ay * wz[cz][k];
                }
#pragma unroll
                for (int cz = 0; cz < 2; cz++)
                    acc[cx][cy][cz] = first ? accz[cz] * cl : acc[cx][cy][cz] + accz[cz] * cl;
            }
    }
#pragma unroll
    for (int cz = 0; cz < 2; cz++)
#pragma unroll
        for (int cy = 0; cy < 2; cy++)
#pragma unroll
            for (int cx = 0; cx < 2; cx++) val[cx][cy][cz] = acc[cx][cy][cz];
}

// Tricubic coarse-face fill, stencils shared between fine cells.  Along
// each axis the fine ghost cells of a face map onto the target's clamped
// stencil bases ri (AMR.h:736-745 positions); fine cells with the same base
// on every axis -- usually 2 x 2 x 2 children of one coarse cell -- read the
// same 4^3 nodes.  One thread reads those nodes once per time level and
// contracts them along x, then y, then z for each child's weights, which is
// the same sequence of operations interp_tricubic performs for each child
// alone, so every value is bit-identical to the per-cell fill.  Where
// clamping at tile edges gives more than 2 fine cells along an axis the
// same base, they are split into groups of at most 2, so every thread's
// work stays 2 x 2 x 2 children at most.  The target nodes a face reads
// are staged in LDS for a group of fields and every time level, each node
// read from HBM once with whole-row reads; the jobs then read LDS.  L: tile
// layout (tile_cell addresses the staged block in either).
template <int S, int P, int VPG, int L>
__global__ void __launch_bounds__(256) k_ghost_tricubic(FieldBatchDev b, int n_vars, const uint64_t *__restrict__ bnd,
                                                        const int32_t *__restrict__ tiles,
                                                        const int32_t *__restrict__ faces, uint32_t substep,
                                                        uint32_t maxd, float frac, int equal_mode, LearnedGate gate)
{
    constexpr int T = S + 2 * P, NC = S * S * P;
    constexpr size_t CELLS = tile_stride<S, P, L>();
    constexpr uint32_t uP = P, uS = S;
    int node, dir;
    face_of_block(tiles, faces, node, dir);
    const int v0 = blockIdx.y * VPG;
    const uint64_t bb = bnd[(size_t)node * 6 + dir];
    const uint32_t nd = (uint32_t)bb;
    const int tgt = bnd_target(nd), cd = bnd_depth(nd);
    if (tgt == -1 || cd <= 0) return;
    if (!active_depth(substep, (uint32_t)cd, maxd)) return;
    const float3 relp = cf3_unpack((uint32_t)(bb >> 32));
    const float rl[3] = {relp.x, relp.y, relp.z};
    const uint32_t ddiff = bnd_ddiff(nd);
    if (ddiff == 0 && equal_mode == AM_EQUAL_SKIP) return;
    // the learned fill's faces (see LearnedGate): skipped before the
    // network; after it, refilled for a tile beyond the threshold, and for
    // a tile within it only the cells the network leaves (those whose
    // coarse cell's 4^3 stencil would read an edge or corner ghost)
    const bool learned_face = ddiff == 1u && tgt > 0;
    bool network_cells_kept = false;
    if (gate.mode == 1) {
        if (threadIdx.x == 0) gate.resid[node] = 0.f;
        if (learned_face) return;
    } else if (gate.mode == 2) {  // uniform per workgroup
        if (!learned_face) return;
        network_cells_kept = gate.resid[node] <= gate.threshold;
        if (threadIdx.x == 0 && blockIdx.y == 0 && gate.stats)
            atomicAdd(&gate.stats[network_cells_kept ? 0 : 1], 1ull);
    }
    const float rel_scale = 1.f / (float)(1u << ddiff);
    const float tmp = (float)((substep >> (maxd - (uint32_t)cd)) & ((1u << ddiff) - 1u));
    const float ty = ddiff == 0 ? 1.f : (tmp + frac) / (float)(1u << ddiff);
    const float invS = 1.f / (float)S;
    const int nv = min(VPG, n_vars - v0);
    const int na = dir % 3;  // the face's normal axis
    int cnt[3], wlo[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        cnt[k] = k == na ? P : S;
        wlo[k] = k != na ? P : (dir < 3 ? P + S : 0);
    }

    if (ddiff == 0 && equal_mode == AM_EQUAL_COPY) {
        // equal depth, copy rule: the neighbour cell the ghost maps onto.
        // The read position is a whole cell index exactly for power-of-two
        // S and within rounding otherwise (1/S inexact), so it is rounded.
        // The formula rule (equal_mode AM_EQUAL_FORMULA) takes the general
        // path below, where inexact offsets weigh in as they do per cell.
        for (int c = threadIdx.x; c < NC; c += 256) {
            int ix[3];
            ix[0] = c % cnt[0];
            ix[1] = (c / cnt[0]) % cnt[1];
            ix[2] = c / (cnt[0] * cnt[1]);
            float rp[3];
            uint32_t w[3];
#pragma unroll
            for (int k = 0; k < 3; k++) {
                w[k] = (uint32_t)(wlo[k] + ix[k]);
                rp[k] = (float)w[k];
                rp[k] = (rp[k] - (float)uP + .5f) * invS - .5f;
                rp[k] = rp[k] * rel_scale + rl[k];
                rp[k] = (rp[k] + .5f) * (float)uS + (float)uP - 0.5f;
            }
            const size_t wn = (size_t)node * CELLS + tile_cell<S, P, L>(w[0], w[1], w[2]);
            const size_t src = (size_t)tgt * CELLS + tile_cell<S, P, L>((uint32_t)(rp[0] + .5f), (uint32_t)(rp[1] + .5f),
                                                                         (uint32_t)(rp[2] + .5f));
            for (int v = 0; v < nv; v++) b.new_[v0 + v][wn] = b.new_[v0 + v][src];
        }
        return;
    }

    // per axis and fine index: the clamped base and the 4-point weights
    __shared__ uint32_t s_ri[3][S];
    __shared__ float s_w[3][S][4];
    __shared__ int s_run0[3][S], s_runn[3][S], s_nruns[3];
    if (threadIdx.x < 3 * S) {
        const int k = threadIdx.x / S, i = threadIdx.x - k * S;
        if (i < cnt[k]) {
            float rp = (float)(uint32_t)(wlo[k] + i);
            rp = (rp - (float)uP + .5f) * invS - .5f;
            rp = rp * rel_scale + rl[k];
            rp = (rp + .5f) * (float)uS + (float)uP - 0.5f;
            const uint32_t ri = (tgt == 0) ? min((uint32_t)fmaxf(rp, 1.f), (uint32_t)T - 3u)
                                           : min((uint32_t)fmaxf(rp, (float)(uP + 1u)), (uint32_t)T - uP - 3u);
            rp -= (float)ri;
            lag4(rp, s_w[k][i]);
            s_ri[k][i] = ri;
        }
    }
    __syncthreads();
    if (threadIdx.x < 3) {  // runs of at most 2 equal bases (the positions rise with the index)
        const int k = threadIdx.x;
        int nr = 0;
        for (int i = 0; i < cnt[k]; i++) {
            if (i == 0 || s_ri[k][i] != s_ri[k][i - 1] || s_runn[k][nr - 1] == 2) {
                s_run0[k][nr] = i;
                s_runn[k][nr] = 1;
                nr++;
            } else {
                s_runn[k][nr - 1]++;
            }
        }
        s_nruns[k] = nr;
    }
    __syncthreads();
    const int nr0 = s_nruns[0], nr1 = s_nruns[1], nr2 = s_nruns[2];
    const int items = nr0 * nr1 * nr2;
    const bool quad = ddiff > 0;  // with older: quadratic in time
    const float c0 = .5f * ty * (ty - 1.f), c1 = (1.f - ty) * (1.f + ty), c2 = .5f * ty * (ty + 1.f);
    // one job: the run starts and lengths of its children along each axis
    auto decode = [&](int it, int (&a)[3], int (&nn)[3]) {
        const int q0 = it % nr0, q1 = (it / nr0) % nr1, q2 = it / (nr0 * nr1);
        a[0] = s_run0[0][q0], a[1] = s_run0[1][q1], a[2] = s_run0[2][q2];
        nn[0] = s_runn[0][q0], nn[1] = s_runn[1][q1], nn[2] = s_runn[2][q2];
    };
    // a fine ghost cell the network wrote: its coarse cell (nearest target
    // cell) has a 4^3 stencil (offsets -1..2) inside the target that leaves
    // the interior along one axis at most
    auto network_cell = [&](const int (&w)[3]) {
        int outside = 0;
        bool in_tile = true;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            float r = ((float)w[k] - (float)uP + .5f) * invS - .5f;
            r = r * rel_scale + rl[k];
            r = (r + .5f) * (float)uS + (float)uP - 0.5f;
            const int c = (int)floorf(r + .5f);
            in_tile = in_tile && c - 1 >= 0 && c + 2 < T;
            outside += (c - 1 < P || c + 2 >= P + S) ? 1 : 0;
        }
        return in_tile && outside <= 1;
    };
    auto store = [&](float *dst, const int (&a)[3], const int (&nn)[3], const float (&val)[2][2][2]) {
#pragma unroll
        for (int cz = 0; cz < 2; cz++)
#pragma unroll
            for (int cy = 0; cy < 2; cy++)
#pragma unroll
                for (int cx = 0; cx < 2; cx++) {
                    if (cx >= nn[0] || cy >= nn[1] || cz >= nn[2]) continue;
                    const int w[3] = {wlo[0] + a[0] + cx, wlo[1] + a[1] + cy, wlo[2] + a[2] + cz};
                    if (network_cells_kept && network_cell(w)) continue;
                    const size_t wn = (size_t)node * CELLS + tile_cell<S, P, L>((uint32_t)w[0], (uint32_t)w[1],
                                                                                 (uint32_t)w[2]);
                    dst[wn] = val[cx][cy][cz];
                }
    };

    // The target nodes the face's stencils read span, per axis, the first
    // base - 1 to the last base + 2 (the bases rise with the fine index).
    // Each group of fields stages that block for every time level in LDS,
    // each node read once with whole-row reads, and the jobs read LDS.  For
    // n fine cells along an axis the bases spread over at most
    // ceil((n - 1) / 2^ddiff) + 1 values, and over at most S - 3 (T - 3 for
    // the root) after clamping, so a block holds at most 12 x 12 x 6 nodes
    // one level coarser and (S + 4)^2 (P + 4) at equal depth (formula rule)
    // for S = 16, P = 3; the LDS holds the larger of 2 fields x 3 levels of
    // the first and 1 field x 2 levels of the second.
    constexpr int ET = (S + 1) / 2 + 4, EN = P / 2 + 5;
    constexpr int CAP = 2 * 3 * ET * ET * EN > 2 * (S + 4) * (S + 4) * (P + 4) ? 2 * 3 * ET * ET * EN
                                                                                : 2 * (S + 4) * (S + 4) * (P + 4);
    __shared__ float s_reg[CAP];
    int lo[3];
#pragma unroll
    for (int k = 0; k < 3; k++) lo[k] = (int)s_ri[k][0] - 1;
    const int E0 = (int)s_ri[0][cnt[0] - 1] + 3 - lo[0], E1 = (int)s_ri[1][cnt[1] - 1] + 3 - lo[1],
              E2 = (int)s_ri[2][cnt[2] - 1] + 3 - lo[2];
    const int EV = E0 * E1 * E2;
    const int lv0 = quad ? 0 : 1, nlev = 3 - lv0;  // levels staged: (older,) old, new
    const int per = min(nv, CAP / (nlev * EV));    // fields per group (>= 1 by the bounds above)
    // Staging: per field (uniform, so its level pointers stay in scalar
    // registers), batches of 4 elements per thread; the 4 elements' offsets
    // first (the interior-first map branches), then all their loads for
    // every level back to back, then the LDS writes.  Element c of a level
    // is (z, y, x) of the block, by divisions through float reciprocals,
    // (c + 1/2) / d rounded down, exact for c + d < 2^22.
    constexpr int BATCH = 4;
    const float inv_e0 = 1.f / (float)E0, inv_e1 = 1.f / (float)E1;
    auto qdiv = [](int i, float inv) { return (int)(((float)i + .5f) * inv); };
    const size_t tb = (size_t)tgt * CELLS;
    float *newp[VPG];
    bool older_in[VPG];
#pragma unroll
    for (int i = 0; i < VPG; i++) {
        const int vi = min(v0 + i, n_vars - 1);
        newp[i] = b.new_[vi];
        older_in[i] = b.older[vi] != nullptr;
    }
    for (int vg = 0; vg < nv; vg += per) {
        const int nf = min(per, nv - vg);
#pragma unroll 1
        for (int f = 0; f < nf; f++) {
            const int v = v0 + vg + f;
            // without an older level (linear in time; the jobs never read it)
            // the older slot loads old and stores zeros
            const bool has_older = quad && b.older[v] != nullptr;
            const float *lvp[3] = {has_older ? b.older[v] + tb : b.old_[v] + tb, b.old_[v] + tb, b.new_[v] + tb};
#pragma unroll 1
            for (int c0 = 0; c0 < EV; c0 += BATCH * 256) {
                uint32_t ofs[BATCH];
#pragma unroll
                for (int u = 0; u < BATCH; u++) {
                    int c = c0 + (int)threadIdx.x + u * 256;
                    c = c < EV ? c : 0;  // past the block: element 0, loaded and not stored
                    const int yz = qdiv(c, inv_e0), x = c - yz * E0;
                    const int z = qdiv(yz, inv_e1), y = yz - z * E1;
                    ofs[u] = tile_cell<S, P, L>((uint32_t)(lo[0] + x), (uint32_t)(lo[1] + y), (uint32_t)(lo[2] + z));
                }
                float val[3][BATCH];
#pragma unroll
                for (int l = 0; l < 3; l++)
#pragma unroll
                    for (int u = 0; u < BATCH; u++) {
                        const float x = l >= lv0 ? lvp[l][ofs[u]] : 0.f;
                        val[l][u] = l == 0 && !has_older ? 0.f : x;
                    }
#pragma unroll
                for (int l = 0; l < 3; l++) {
                    if (l < lv0) continue;
#pragma unroll
                    for (int u = 0; u < BATCH; u++) {
                        const int c = c0 + (int)threadIdx.x + u * 256;
                        if (c < EV) s_reg[(f * nlev + (l - lv0)) * EV + c] = val[l][u];
                    }
                }
            }
        }
        __syncthreads();
        for (int job = threadIdx.x; job < items * nf; job += 256) {
            const int f = job / items;
            int a[3], nn[3];
            decode(job - f * items, a, nn);
            // the job's field: its pointers picked from registers (an indexed
            // read of the argument table would be a memory round trip per job)
            float *dst = newp[0];
            bool older_ok = older_in[0];
#pragma unroll
            for (int i = 1; i < VPG; i++)
                if (vg + f == i) {
                    dst = newp[i];
                    older_ok = older_in[i];
                }
            const int off = (((int)s_ri[2][a[2]] - 1 - lo[2]) * E1 + ((int)s_ri[1][a[1]] - 1 - lo[1])) * E0 +
                            ((int)s_ri[0][a[0]] - 1 - lo[0]);
            const float *sr = s_reg + f * nlev * EV + off - lv0 * EV;
            float val[2][2][2];
            tricubic_children(
                [&](int lv, int j, int k, float (&r)[4]) {
                    const float *q = sr + lv * EV + (k * E1 + j) * E0;
                    r[0] = q[0]; r[1] = q[1]; r[2] = q[2]; r[3] = q[3];
                },
                quad && older_ok, s_w, a, nn, ty, c0, c1, c2, val);
            store(dst, a, nn, val);
        }
        __syncthreads();
    }
}

constexpr int GHOST_VPG = 6;  // fields per workgroup in the multi-field fill
constexpr int TRICUBIC_VPG = 2;  // fields per workgroup in the tricubic fill

template <int S, int P, int L>
static int launch_ghost_multi_l(const FieldBatchDev &d, int n_vars, const am_mesh *mesh, const int32_t *faces,
                                int n_faces, int interp, float frac, int equal_mode, int skip_learned,
                                LearnedGate gate, void *stream)
{
    const unsigned items = faces ? (unsigned)n_faces : (unsigned)mesh->n_tiles * 6u;
    const dim3 grid(items, (n_vars + GHOST_VPG - 1) / GHOST_VPG);
    hipStream_t st = (hipStream_t)stream;
    if (interp == AM_INTERP_LAGRANGE)
        hipLaunchKernelGGL((k_ghost_fill_multi<S, P, AM_INTERP_LAGRANGE, GHOST_VPG, L>), grid, dim3(256), 0, st, d,
                           n_vars, mesh->boundaries, mesh->tiles, faces, mesh->substep, mesh->max_depth, frac,
                           equal_mode, skip_learned);
    else if (interp == AM_INTERP_TRILINEAR)
        hipLaunchKernelGGL((k_ghost_fill_multi<S, P, AM_INTERP_TRILINEAR, GHOST_VPG, L>), grid, dim3(256), 0, st, d,
                           n_vars, mesh->boundaries, mesh->tiles, faces, mesh->substep, mesh->max_depth, frac,
                           equal_mode, skip_learned);
    else if (interp == AM_INTERP_TRICUBIC && !skip_learned)
        hipLaunchKernelGGL((k_ghost_tricubic<S, P, TRICUBIC_VPG, L>),
                           dim3(items, (n_vars + TRICUBIC_VPG - 1) / TRICUBIC_VPG), dim3(256), 0, st, d, n_vars,
                           mesh->boundaries, mesh->tiles, faces, mesh->substep, mesh->max_depth, frac, equal_mode,
                           gate);
    else if (interp == AM_INTERP_TRICUBIC)
        hipLaunchKernelGGL((k_ghost_fill_multi<S, P, AM_INTERP_TRICUBIC, GHOST_VPG, L>), grid, dim3(256), 0, st, d,
                           n_vars, mesh->boundaries, mesh->tiles, faces, mesh->substep, mesh->max_depth, frac,
                           equal_mode, skip_learned);
    else
        return set_error(AM_ERR_INVALID, "am_ghost_fill: unknown interpolation mode %d", interp);
    return check_launch("k_ghost_fill_multi");
}

template <int S, int P>
static int launch_ghost_multi(const am_field_batch *b, const am_mesh *mesh, const int32_t *faces, int n_faces,
                              int interp, float frac, int equal_mode, int skip_learned, void *stream,
                              LearnedGate gate = LearnedGate{})
{
    FieldBatchDev d;
    for (int v = 0; v < b->n_vars; v++) {
        if (!b->old_[v] || !b->new_[v]) return set_error(AM_ERR_INVALID, "am_ghost_fill: null field %d", v);
        d.old_[v] = b->old_[v];
        d.new_[v] = b->new_[v];
        d.older[v] = b->older[v];
    }
    AM_DISPATCH_LAYOUT(b->tile_layout, launch_ghost_multi_l, d, b->n_vars, mesh, faces, n_faces, interp, frac,
                       equal_mode, skip_learned, gate, stream);
}

template <int S, int P>
static int launch_ghost(const am_field_batch *b, const am_mesh *mesh, int copy_only_new, int interp, float frac,
                        void *stream)
{
    if (b->tile_layout != AM_LAYOUT_PADDED)
        return set_error(AM_ERR_INVALID, "am_ghost_fill_batched: the reference fill takes padded tiles only");
    FieldBatchDev d;
    for (int v = 0; v < b->n_vars; v++) {
        if (!b->old_[v] || !b->new_[v]) return set_error(AM_ERR_INVALID, "am_ghost_fill: null field %d", v);
        d.old_[v] = b->old_[v];
        d.new_[v] = b->new_[v];
        d.older[v] = b->older[v];
    }
    dim3 grid(mesh->n_tiles * 6, b->n_vars);
    hipStream_t st = (hipStream_t)stream;
    if (interp == AM_INTERP_LAGRANGE)
        hipLaunchKernelGGL((k_ghost_fill<S, P, AM_INTERP_LAGRANGE>), grid, dim3(S * S), 0, st, d, mesh->boundaries,
                           mesh->tiles, mesh->substep, mesh->max_depth, copy_only_new, frac);
    else if (interp == AM_INTERP_TRILINEAR)
        hipLaunchKernelGGL((k_ghost_fill<S, P, AM_INTERP_TRILINEAR>), grid, dim3(S * S), 0, st, d, mesh->boundaries,
                           mesh->tiles, mesh->substep, mesh->max_depth, copy_only_new, frac);
    else if (interp == AM_INTERP_TRICUBIC)
        hipLaunchKernelGGL((k_ghost_fill<S, P, AM_INTERP_TRICUBIC>), grid, dim3(S * S), 0, st, d, mesh->boundaries,
                           mesh->tiles, mesh->substep, mesh->max_depth, copy_only_new, frac);
    else
        return set_error(AM_ERR_INVALID, "am_ghost_fill: unknown interpolation mode %d", interp);
    return check_launch("k_ghost_fill");
}

// __copy_new_to_old (AMR.h:961-996).  One workgroup per (tile, field)
// streams the whole padded tile: planes outside the interior, and every
// plane of a tile without children, are plain 16-byte-vector copies; the
// interior planes of a refined tile take the child means cell by cell.
// With `older` set, the previous old is kept there first.  With
// `keep_ghosts` set, a tile whose flag is 0 copies its interior cells only
// (its ghost cells of old and older are left as they are).
template <int S, int P>
__global__ void __launch_bounds__(256) k_copy_new_to_old(FieldBatchDev b, const uint32_t *__restrict__ positions,
                                                         const int32_t *__restrict__ children,
                                                         const int32_t *__restrict__ tiles, uint32_t substep,
                                                         uint32_t maxd, const uint8_t *__restrict__ keep_ghosts)
{
    constexpr int T = S + 2 * P, T2 = T * T;
    constexpr size_t CELLS = (size_t)T * T2;
    const int node = tiles[blockIdx.x];
    const int depth = pos_depth(positions[node]);
    if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) return;
    float *__restrict__ dold = b.old_[blockIdx.y];
    const float *__restrict__ dnew = b.new_[blockIdx.y];
    float *__restrict__ dolder = const_cast<float *>(b.older[blockIdx.y]);
    __shared__ int ch[8];
    if (threadIdx.x < 8) ch[threadIdx.x] = children[(size_t)node * 8 + threadIdx.x];
    __syncthreads();
    bool any_child = false;
#pragma unroll
    for (int i = 0; i < 8; i++) any_child |= ch[i] != -1;
    const size_t base = (size_t)node * CELLS;
    constexpr uint32_t H = S / 2u;
    if (keep_ghosts && !keep_ghosts[node] && !any_child) {  // uniform per workgroup
        // each interior plane's rows P..P+S-1 are one contiguous run of S*T
        // floats: copied as 16-byte vectors from the vector boundary at or
        // below its start to the one at or above its end (the few extra
        // cells are ghost cells, copied as the full form would)
        for (int z = P; z < P + S; z++) {
            const int a4 = ((z * T + P) * T) / 4, b4 = ((z * T + P + S) * T + 3) / 4;
            const float4 *s4 = reinterpret_cast<const float4 *>(dnew + base);
            float4 *d4 = reinterpret_cast<float4 *>(dold + base);
            float4 *o4 = dolder ? reinterpret_cast<float4 *>(dolder + base) : nullptr;
            for (int i = a4 + (int)threadIdx.x; i < b4; i += 256) {
                if (o4) o4[i] = d4[i];
                d4[i] = s4[i];
            }
        }
        return;
    }
    if (keep_ghosts && !keep_ghosts[node]) {  // uniform per workgroup: restriction on the interior only
        for (int z = P; z < P + S; z++) {
            const uint32_t tz = ((uint32_t)z - (uint32_t)P) / H;
            for (int i = threadIdx.x; i < S * S; i += 256) {
                const int yi = i / S, xi = i - yi * S;
                const size_t o = base + ((size_t)z * T + (yi + P)) * T + (xi + P);
                if (dolder) dolder[o] = dold[o];
                const int child = any_child ? ch[(uint32_t)xi / H + ((uint32_t)yi / H) * 2u + tz * 4u] : -1;
                if (child == -1) {
                    dold[o] = dnew[o];
                    continue;
                }
                const size_t cx = ((uint32_t)xi * 2u) % (uint32_t)S + (uint32_t)P;
                const size_t cy = ((uint32_t)yi * 2u) % (uint32_t)S + (uint32_t)P;
                const size_t cz = (((uint32_t)z - (uint32_t)P) * 2u) % (uint32_t)S + (uint32_t)P;
                const size_t c = (size_t)child * CELLS + (cz * T + cy) * T + cx;
                dold[o] = (dnew[c] + dnew[c + 1] + dnew[c + T] + dnew[c + T + 1] + dnew[c + T2] + dnew[c + T2 + 1] +
                           dnew[c + T2 + T] + dnew[c + T2 + T + 1]) *
                          .125f;
            }
        }
        return;
    }
    for (int z = 0; z < T; z++) {
        const uint32_t tz = ((uint32_t)z - (uint32_t)P) / H;
        if (!any_child || tz >= 2u) {
            const float4 *s4 = reinterpret_cast<const float4 *>(dnew + base + (size_t)z * T2);
            float4 *d4 = reinterpret_cast<float4 *>(dold + base + (size_t)z * T2);
            float4 *o4 = dolder ? reinterpret_cast<float4 *>(dolder + base + (size_t)z * T2) : nullptr;
            for (int i = threadIdx.x; i < T2 / 4; i += 256) {
                if (o4) o4[i] = d4[i];
                d4[i] = s4[i];
            }
            continue;
        }
        for (int i = threadIdx.x; i < T2; i += 256) {
            const int y = i / T, x = i - y * T;
            const uint32_t tx = ((uint32_t)x - (uint32_t)P) / H, ty = ((uint32_t)y - (uint32_t)P) / H;
            const int child = (tx < 2u && ty < 2u) ? ch[tx + ty * 2u + tz * 4u] : -1;
            const size_t o = base + (size_t)z * T2 + i;
            if (dolder) dolder[o] = dold[o];
            if (child == -1) {
                dold[o] = dnew[o];
                continue;
            }
            const size_t cx = (((uint32_t)x - (uint32_t)P) * 2u) % (uint32_t)S + (uint32_t)P;
            const size_t cy = (((uint32_t)y - (uint32_t)P) * 2u) % (uint32_t)S + (uint32_t)P;
            const size_t cz = (((uint32_t)z - (uint32_t)P) * 2u) % (uint32_t)S + (uint32_t)P;
            const size_t c = (size_t)child * CELLS + (cz * T + cy) * T + cx;
            dold[o] = (dnew[c] + dnew[c + 1] + dnew[c + T] + dnew[c + T + 1] + dnew[c + T2] + dnew[c + T2 + 1] +
                       dnew[c + T2 + T] + dnew[c + T2 + T + 1]) *
                      .125f;
        }
    }
}

// After the driver's copy-back: new <- old on the interior cells of each
// active tile that a child covers, so that new equals old there too (the
// RK4 driver's first stage reads new).  A separate launch, so that every
// parent has read its children's new before any of them is changed.
template <int S, int P>
__global__ void __launch_bounds__(256) k_restricted_to_new(FieldBatchDev b, const uint32_t *__restrict__ positions,
                                                           const int32_t *__restrict__ children,
                                                           const int32_t *__restrict__ tiles, uint32_t substep,
                                                           uint32_t maxd)
{
    constexpr int T = S + 2 * P;
    constexpr size_t CELLS = (size_t)T * T * T;
    constexpr uint32_t H = S / 2u;
    const int node = tiles[blockIdx.x];
    const int depth = pos_depth(positions[node]);
    if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) return;
    __shared__ int ch[8];
    if (threadIdx.x < 8) ch[threadIdx.x] = children[(size_t)node * 8 + threadIdx.x];
    __syncthreads();
    bool any_child = false;
#pragma unroll
    for (int i = 0; i < 8; i++) any_child |= ch[i] != -1;
    if (!any_child) return;  // uniform per workgroup
    const float *dold = b.old_[blockIdx.y];
    float *dnew = b.new_[blockIdx.y];
    const size_t base = (size_t)node * CELLS;
    for (int i = threadIdx.x; i < S * S * S; i += 256) {
        const int zi = i / (S * S), yi = (i / S) % S, xi = i % S;
        if (ch[(uint32_t)xi / H + ((uint32_t)yi / H) * 2u + ((uint32_t)zi / H) * 4u] == -1) continue;
        const size_t o = base + ((size_t)(zi + P) * T + (yi + P)) * T + (xi + P);
        dnew[o] = dold[o];
    }
}

// k_copy_new_to_old and k_restricted_to_new on interior-first tiles: the
// S^3 interior block comes first in each slot, so a tile that keeps its
// ghost cells copies its whole slot and one that does not copies only that
// block, both as 16-byte vectors; interior cells covered by a child take the
// mean of its 2x2x2 interior cells, summed in the reference's order.
template <int S, int P>
__global__ void __launch_bounds__(256) k_copy_back_if(FieldBatchDev b, const uint32_t *__restrict__ positions,
                                                      const int32_t *__restrict__ children,
                                                      const int32_t *__restrict__ tiles, uint32_t substep,
                                                      uint32_t maxd, const uint8_t *__restrict__ keep_ghosts)
{
    constexpr uint32_t TS = tile_stride<S, P, 1>(), S3 = S * S * S, H = S / 2;
    static_assert(S3 % 4 == 0 && TS % 4 == 0, "vector copies");
    const int node = tiles[blockIdx.x];
    const int depth = pos_depth(positions[node]);
    if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) return;
    float *__restrict__ dold = b.old_[blockIdx.y];
    const float *__restrict__ dnew = b.new_[blockIdx.y];
    float *__restrict__ dolder = const_cast<float *>(b.older[blockIdx.y]);
    __shared__ int ch[8];
    if (threadIdx.x < 8) ch[threadIdx.x] = children[(size_t)node * 8 + threadIdx.x];
    __syncthreads();
    bool any_child = false;
#pragma unroll
    for (int i = 0; i < 8; i++) any_child |= ch[i] != -1;
    const size_t base = (size_t)node * TS;
    const bool whole = !keep_ghosts || keep_ghosts[node];
    auto copy4 = [&](uint32_t lo, uint32_t hi) {
        const float4 *s4 = reinterpret_cast<const float4 *>(dnew + base);
        float4 *d4 = reinterpret_cast<float4 *>(dold + base);
        float4 *o4 = dolder ? reinterpret_cast<float4 *>(dolder + base) : nullptr;
        for (uint32_t i = lo / 4 + threadIdx.x; i < hi / 4; i += 256) {
            if (o4) o4[i] = d4[i];
            d4[i] = s4[i];
        }
    };
    if (!any_child) {  // uniform per workgroup
        copy4(0u, whole ? TS : S3);
        return;
    }
    for (uint32_t i = threadIdx.x; i < S3; i += 256) {
        const uint32_t zi = i / (S * S), yi = (i / S) % S, xi = i % S;
        const size_t o = base + i;
        if (dolder) dolder[o] = dold[o];
        const int child = ch[xi / H + (yi / H) * 2u + (zi / H) * 4u];
        if (child == -1) {
            dold[o] = dnew[o];
            continue;
        }
        const size_t c = (size_t)child * TS + (((zi * 2u) % S) * S + (yi * 2u) % S) * S + (xi * 2u) % S;
        constexpr size_t SY = S, SZ = S * S;
        dold[o] = (dnew[c] + dnew[c + 1] + dnew[c + SY] + dnew[c + SY + 1] + dnew[c + SZ] + dnew[c + SZ + 1] +
                   dnew[c + SZ + SY] + dnew[c + SZ + SY + 1]) *
                  .125f;
    }
    if (whole) copy4(S3, TS);
}

template <int S, int P>
__global__ void __launch_bounds__(256) k_restricted_to_new_if(FieldBatchDev b, const uint32_t *__restrict__ positions,
                                                              const int32_t *__restrict__ children,
                                                              const int32_t *__restrict__ tiles, uint32_t substep,
                                                              uint32_t maxd)
{
    constexpr uint32_t TS = tile_stride<S, P, 1>(), S3 = S * S * S, H = S / 2;
    const int node = tiles[blockIdx.x];
    const int depth = pos_depth(positions[node]);
    if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) return;
    __shared__ int ch[8];
    if (threadIdx.x < 8) ch[threadIdx.x] = children[(size_t)node * 8 + threadIdx.x];
    __syncthreads();
    bool any_child = false;
#pragma unroll
    for (int i = 0; i < 8; i++) any_child |= ch[i] != -1;
    if (!any_child) return;  // uniform per workgroup
    const float *dold = b.old_[blockIdx.y];
    float *dnew = b.new_[blockIdx.y];
    const size_t base = (size_t)node * TS;
    for (uint32_t i = threadIdx.x; i < S3; i += 256) {
        const uint32_t zi = i / (S * S), yi = (i / S) % S, xi = i % S;
        if (ch[xi / H + (yi / H) * 2u + (zi / H) * 4u] == -1) continue;
        dnew[base + i] = dold[base + i];
    }
}

template <int S, int P>
static int launch_copy_back(const am_field_batch *b, const am_mesh *mesh, const uint8_t *keep_ghosts, void *stream)
{
    FieldBatchDev d;
    for (int v = 0; v < b->n_vars; v++) {
        if (!b->old_[v] || !b->new_[v]) return set_error(AM_ERR_INVALID, "am_copy_new_to_old: null field %d", v);
        if (!aligned16(b->old_[v]) || !aligned16(b->new_[v]) || (b->older[v] && !aligned16(b->older[v])))
            return set_error(AM_ERR_MISALIGNED, "am_copy_new_to_old: field %d not 16-byte aligned", v);
        d.old_[v] = b->old_[v];
        d.new_[v] = b->new_[v];
        d.older[v] = b->older[v];
    }
    dim3 grid(mesh->n_tiles, b->n_vars);
    if (b->tile_layout == AM_LAYOUT_INTERIOR_FIRST) {
        hipLaunchKernelGGL((k_copy_back_if<S, P>), grid, dim3(256), 0, (hipStream_t)stream, d, mesh->positions,
                           mesh->children, mesh->tiles, mesh->substep, mesh->max_depth, keep_ghosts);
        if (keep_ghosts)
            hipLaunchKernelGGL((k_restricted_to_new_if<S, P>), grid, dim3(256), 0, (hipStream_t)stream, d,
                               mesh->positions, mesh->children, mesh->tiles, mesh->substep, mesh->max_depth);
        return check_launch("k_copy_back_if");
    }
    if (b->tile_layout != AM_LAYOUT_PADDED)
        return set_error(AM_ERR_INVALID, "am_copy_new_to_old: tile layout %d", b->tile_layout);
    hipLaunchKernelGGL((k_copy_new_to_old<S, P>), grid, dim3(256), 0, (hipStream_t)stream, d, mesh->positions,
                       mesh->children, mesh->tiles, mesh->substep, mesh->max_depth, keep_ghosts);
    if (keep_ghosts)
        hipLaunchKernelGGL((k_restricted_to_new<S, P>), grid, dim3(256), 0, (hipStream_t)stream, d, mesh->positions,
                           mesh->children, mesh->tiles, mesh->substep, mesh->max_depth);
    return check_launch("k_copy_new_to_old");
}

// __copy_to_child (AMR.h:1053-1073); L: tile layout of the buffers
template <int S, int P, int MODE, int L>
__global__ void __launch_bounds__(256) k_copy_to_child(ChildBufs bufs, uint32_t offset, int child, int parent)
{
    constexpr int T = S + 2 * P, T2 = T * T;
    constexpr size_t CELLS = tile_stride<S, P, L>();
    float *buf = bufs.buf[blockIdx.y];
    const int z = blockIdx.x;
    const float invS = 1.f / (float)S;
    const float off[3] = {(float)(offset & 1u), (float)((offset & 2u) >> 1), (float)((offset & 4u) >> 2)};
    for (int i = threadIdx.x; i < T2; i += 256) {
        const int y = i / T, x = i - y * T;
        const uint32_t idx[3] = {(uint32_t)x, (uint32_t)y, (uint32_t)z};
        float rp[3];
        uint32_t ri[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            rp[k] = ((float)idx[k] - (float)(uint32_t)P + .5f) * invS - .5f;
            rp[k] = rp[k] * .5f + (off[k] - .5f) * .5f;
            rp[k] = (rp[k] + .5f) * (float)(uint32_t)S + (float)(uint32_t)P - .5f;
            ri[k] = (MODE == AM_INTERP_TRILINEAR) ? (uint32_t)rp[k] : (uint32_t)(rp[k] + .5f);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) rp[k] -= (float)ri[k];
        float v;
        if constexpr (L == 0) {
            const size_t rn = (size_t)parent * CELLS + ri[0] + ((size_t)ri[1] + (size_t)ri[2] * T) * T;
            v = (MODE == AM_INTERP_TRILINEAR) ? interp_trilinear<T>(buf, rp[0], rp[1], rp[2], rn)
                                              : interp_lagrange<T>(buf, rp[0], rp[1], rp[2], rn);
        } else {
            const Cells<S, P, L> cn{(size_t)parent * CELLS, ri[0], ri[1], ri[2]};
            v = (MODE == AM_INTERP_TRILINEAR) ? interp_trilinear_c<S, P, L>(buf, cn, rp[0], rp[1], rp[2])
                                              : interp_lagrange_c<S, P, L>(buf, cn, rp[0], rp[1], rp[2]);
        }
        buf[(size_t)child * CELLS + tile_cell<S, P, L>((uint32_t)x, (uint32_t)y, (uint32_t)z)] = v;
    }
}

// one launch for n buffers (grid.y = n): each buffer is prolonged from its
// own parent tile
template <int S, int P, int L>
static int launch_to_child_l(const ChildBufs &bufs, int n, uint32_t offset, int child, int parent, int interp,
                             void *stream)
{
    constexpr int T = S + 2 * P;
    if (interp == AM_INTERP_TRILINEAR)
        hipLaunchKernelGGL((k_copy_to_child<S, P, AM_INTERP_TRILINEAR, L>), dim3(T, n), dim3(256), 0,
                           (hipStream_t)stream, bufs, offset, child, parent);
    else
        hipLaunchKernelGGL((k_copy_to_child<S, P, AM_INTERP_LAGRANGE, L>), dim3(T, n), dim3(256), 0,
                           (hipStream_t)stream, bufs, offset, child, parent);
    return check_launch("k_copy_to_child");
}
template <int S, int P>
static int launch_to_child(const ChildBufs &bufs, int n, uint32_t offset, int child, int parent, int interp,
                           int layout, void *stream)
{
    AM_DISPATCH_LAYOUT(layout, launch_to_child_l, bufs, n, offset, child, parent, interp, stream);
}

static int mesh_checks(const am_mesh *mesh, const am_geometry *g, bool need_bnd, bool need_children)
{
    if (!mesh || !g) return set_error(AM_ERR_INVALID, "null mesh or geometry");
    if (!mesh->tiles || !mesh->positions) return set_error(AM_ERR_INVALID, "mesh without tiles or positions");
    if (need_bnd && !mesh->boundaries) return set_error(AM_ERR_INVALID, "mesh without boundaries");
    if (need_children && !mesh->children) return set_error(AM_ERR_INVALID, "mesh without children");
    if (mesh->n_tiles < 0) return set_error(AM_ERR_INVALID, "negative tile count");
    return AM_OK;
}

}  // namespace am

extern "C" int am_ghost_fill_batched(const am_field_batch *b, const am_mesh *mesh, const am_geometry *g,
                                     int32_t copy_only_new, int32_t interp, void *stream)
{
    if (!b) return am::set_error(AM_ERR_INVALID, "am_ghost_fill_batched: null batch");
    if (b->n_vars < 0 || b->n_vars > AM_MAX_VARS)
        return am::set_error(AM_ERR_TOO_MANY_VARS, "am_ghost_fill_batched: n_vars %d", b->n_vars);
    if (int rc = am::mesh_checks(mesh, g, true, false)) return rc;
    if (b->n_vars == 0 || mesh->n_tiles == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_ghost, b, mesh, copy_only_new, interp, 1.f, stream);
}

extern "C" int am_ghost_fill_stage_batched(const am_field_batch *b, const am_mesh *mesh, const am_geometry *g,
                                           float stage_frac, int32_t interp, int32_t equal_mode, void *stream)
{
    if (equal_mode < AM_EQUAL_FORMULA || equal_mode > AM_EQUAL_SKIP)
        return am::set_error(AM_ERR_INVALID, "am_ghost_fill_stage_batched: equal_mode %d", equal_mode);
    if (!b) return am::set_error(AM_ERR_INVALID, "am_ghost_fill_stage_batched: null batch");
    if (b->n_vars < 0 || b->n_vars > AM_MAX_VARS)
        return am::set_error(AM_ERR_TOO_MANY_VARS, "am_ghost_fill_stage_batched: n_vars %d", b->n_vars);
    if (!(stage_frac >= 0.f && stage_frac <= 1.f))
        return am::set_error(AM_ERR_INVALID, "am_ghost_fill_stage_batched: stage_frac %g", (double)stage_frac);
    if (int rc = am::mesh_checks(mesh, g, true, false)) return rc;
    if (b->n_vars == 0 || mesh->n_tiles == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_ghost_multi, b, mesh, nullptr, 0, interp, stage_frac, equal_mode, 0, stream);
}

extern "C" int am_ghost_fill_stage_faces(const am_field_batch *b, const am_mesh *mesh, const int32_t *faces,
                                         int32_t n_faces, const am_geometry *g, float stage_frac, int32_t interp,
                                         int32_t equal_mode, void *stream)
{
    if (equal_mode < AM_EQUAL_FORMULA || equal_mode > AM_EQUAL_SKIP)
        return am::set_error(AM_ERR_INVALID, "am_ghost_fill_stage_faces: equal_mode %d", equal_mode);
    if (!b) return am::set_error(AM_ERR_INVALID, "am_ghost_fill_stage_faces: null batch");
    if (b->n_vars < 0 || b->n_vars > AM_MAX_VARS)
        return am::set_error(AM_ERR_TOO_MANY_VARS, "am_ghost_fill_stage_faces: n_vars %d", b->n_vars);
    if (!(stage_frac >= 0.f && stage_frac <= 1.f))
        return am::set_error(AM_ERR_INVALID, "am_ghost_fill_stage_faces: stage_frac %g", (double)stage_frac);
    if (n_faces < 0 || (n_faces > 0 && !faces))
        return am::set_error(AM_ERR_INVALID, "am_ghost_fill_stage_faces: %d faces", n_faces);
    if (int rc = am::mesh_checks(mesh, g, true, false)) return rc;
    if (b->n_vars == 0 || n_faces == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_ghost_multi, b, mesh, faces, n_faces, interp, stage_frac, equal_mode, 0, stream);
}

int am::ghost_stage_fill(const am_field_batch *b, const am_mesh *mesh, const int32_t *faces, int n_faces,
                         const am_geometry *g, float stage_frac, int equal_mode, void *stream, LearnedGate gate)
{
    AM_DISPATCH_TILE(*g, am::launch_ghost_multi, b, mesh, faces, n_faces, AM_INTERP_TRICUBIC, stage_frac, equal_mode,
                     0, stream, gate);
}

extern "C" int am_ghost_fill(float *old_, float *new_, const float *older, const am_mesh *mesh, const am_geometry *g,
                             int32_t copy_only_new, int32_t interp, void *stream)
{
    am_field_batch b = {};
    b.old_[0] = old_;
    b.new_[0] = new_;
    b.older[0] = const_cast<float *>(older);
    b.n_vars = 1;
    return am_ghost_fill_batched(&b, mesh, g, copy_only_new, interp, stream);
}

extern "C" int am_copy_new_to_old_batched(const am_field_batch *b, const am_mesh *mesh, const am_geometry *g,
                                          void *stream)
{
    if (!b) return am::set_error(AM_ERR_INVALID, "am_copy_new_to_old_batched: null batch");
    if (b->n_vars < 0 || b->n_vars > AM_MAX_VARS)
        return am::set_error(AM_ERR_TOO_MANY_VARS, "am_copy_new_to_old_batched: n_vars %d", b->n_vars);
    if (int rc = am::mesh_checks(mesh, g, false, true)) return rc;
    if (b->n_vars == 0 || mesh->n_tiles == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_copy_back, b, mesh, nullptr, stream);
}

extern "C" int am_copy_new_to_old_interior(const am_field_batch *b, const am_mesh *mesh, const uint8_t *keep_ghosts,
                                           const am_geometry *g, void *stream)
{
    if (!b || !keep_ghosts) return am::set_error(AM_ERR_INVALID, "am_copy_new_to_old_interior: null argument");
    if (b->n_vars < 0 || b->n_vars > AM_MAX_VARS)
        return am::set_error(AM_ERR_TOO_MANY_VARS, "am_copy_new_to_old_interior: n_vars %d", b->n_vars);
    if (int rc = am::mesh_checks(mesh, g, false, true)) return rc;
    if (b->n_vars == 0 || mesh->n_tiles == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_copy_back, b, mesh, keep_ghosts, stream);
}

extern "C" int am_copy_new_to_old(float *old_, const float *new_, const am_mesh *mesh, const am_geometry *g,
                                  void *stream)
{
    am_field_batch b = {};
    b.old_[0] = old_;
    b.new_[0] = const_cast<float *>(new_);
    b.n_vars = 1;
    return am_copy_new_to_old_batched(&b, mesh, g, stream);
}

static int to_child_checks(uint32_t offset, int32_t child, int32_t parent, const am_geometry *g, int32_t interp)
{
    if (!g) return am::set_error(AM_ERR_INVALID, "am_copy_to_child: null geometry");
    if (offset > 7u || child < 0 || parent < 0) return am::set_error(AM_ERR_INVALID, "am_copy_to_child: bad indices");
    if (interp != AM_INTERP_LAGRANGE && interp != AM_INTERP_TRILINEAR)
        return am::set_error(AM_ERR_INVALID, "am_copy_to_child: interpolation %d not supported", interp);
    return AM_OK;
}

extern "C" int am_copy_to_child(float *buf, uint32_t offset, int32_t child, int32_t parent, const am_geometry *g,
                                int32_t interp, void *stream)
{
    if (!buf) return am::set_error(AM_ERR_INVALID, "am_copy_to_child: null buffer");
    if (int rc = to_child_checks(offset, child, parent, g, interp)) return rc;
    am::ChildBufs b{};
    b.buf[0] = buf;
    AM_DISPATCH_TILE(*g, am::launch_to_child, b, 1, offset, child, parent, interp, AM_LAYOUT_PADDED, stream);
}

extern "C" int am_copy_to_child_batched(const am_field_batch *fb, uint32_t offset, int32_t child, int32_t parent,
                                        const am_geometry *g, int32_t interp, void *stream)
{
    if (!fb) return am::set_error(AM_ERR_INVALID, "am_copy_to_child_batched: null batch");
    if (fb->n_vars < 0 || fb->n_vars > AM_MAX_VARS)
        return am::set_error(AM_ERR_TOO_MANY_VARS, "am_copy_to_child_batched: n_vars %d", fb->n_vars);
    if (int rc = to_child_checks(offset, child, parent, g, interp)) return rc;
    am::ChildBufs b{};
    int n = 0;
    for (int v = 0; v < fb->n_vars; v++) {
        if (fb->old_[v]) b.buf[n++] = fb->old_[v];
        if (fb->new_[v]) b.buf[n++] = fb->new_[v];
        if (fb->older[v]) b.buf[n++] = const_cast<float *>(fb->older[v]);
    }
    if (n == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_to_child, b, n, offset, child, parent, interp, fb->tile_layout, stream);
}
