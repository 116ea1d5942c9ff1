// Finite-difference stencils for gfx950: gradient, diagonal Hessian and the
// boundary-careful mixed partials of padded fp32 tiles, written as dithered
// compressed_float3 words.
//
// One workgroup = one (tile, variable, z-block).  Its S*S threads own the
// (x, y) columns of the tile and march up z.  The padded tile is streamed
// through LDS one full T*T plane at a time (a plane is contiguous in HBM, so
// each plane is a single coalesced 16-byte-vector read), into a ring of 2P+2
// planes: the 2P+1 planes the stencil needs plus one being filled, which lets
// every step use a single barrier.  Outputs for one plane are S*S contiguous
// words per array.
//
// Arithmetic follows the reference kernels term by term (AMR.h:1152-1266)
// with contraction disabled, so each word is bit-identical to the
// reference's on an IEEE host.
#include "am_device.hpp"
#include "am_words.hpp"

#pragma clang fp contract(off)

namespace am {

struct FdBatchDev {
    const float *field[AM_MAX_VARS];
    uint32_t *d1[AM_MAX_VARS];
    uint32_t *d2diag[AM_MAX_VARS];
    uint32_t *d2off[AM_MAX_VARS];
    int32_t seed_d[AM_MAX_VARS];
    int32_t seed_m[AM_MAX_VARS];
    const int32_t *seed_table;  // non-null: seed_d/seed_m index it
    float *adv[AM_MAX_VARS];    // upwinded advection + dissipation (ADV kernels)
    const float *shift[3];
    uint32_t advect_mask;
    float ko_sigma;
};

// Weights of the fourth-order lopsided first derivative, offsets -1..3
// (shift > 0) and -3..1 (shift <= 0), and of the sixth difference, as
// floats: the oracle's orc_fd_adv uses the same values in the same order.
__device__ constexpr float LOP_PLUS[5] = {-3.f / 12, -10.f / 12, 18.f / 12, -6.f / 12, 1.f / 12};
__device__ constexpr float LOP_MINUS[5] = {-1.f / 12, 6.f / 12, -18.f / 12, 10.f / 12, 3.f / 12};
__device__ constexpr float KO6[7] = {1.f, -6.f, 15.f, -20.f, 15.f, -6.f, 1.f};

struct MeshDev {
    const uint32_t *positions;
    const int32_t *tiles;
    uint32_t substep;
    uint32_t max_depth;
    int32_t n_items;  // tiles x z-blocks; workgroups loop over them
    // non-null: ghost cells across a face of the axes in direct_axes (bits
    // 1 x, 2 y, 4 z) with an equal-depth neighbour are read from that
    // neighbour's interior instead of the tile's own ghost layers (which
    // then need not be current)
    const uint64_t *direct;
    uint32_t direct_axes;
};

// Which of the four floats of float4 i of a padded plane are ghost cells the
// stencils read (the Q layers next to the interior) across the -x, +x, -y
// and +y faces, one 4-bit mask each in bits 0-3, 4-7, 8-11, 12-15.  Edge
// cells (ghost along both axes) are in none: no stencil reads them.
template <int S, int P, int Q>
__device__ __forceinline__ uint32_t face_ghost_masks(int i)
{
    constexpr int T = S + 2 * P;
    uint32_t m = 0u;
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const int f = 4 * i + k, r = f / T, c = f - r * T;
        const bool rin = r >= P && r < P + S, cin = c >= P && c < P + S;
        if (rin && c >= P - Q && c < P) m |= 1u << k;
        if (rin && c >= P + S && c < P + S + Q) m |= 1u << (4 + k);
        if (cin && r >= P - Q && r < P) m |= 1u << (8 + k);
        if (cin && r >= P + S && r < P + S + Q) m |= 1u << (12 + k);
    }
    return m;
}
__device__ __forceinline__ float4 take(float4 v, float4 t, uint32_t mask)
{
    if (mask & 1u) v.x = t.x;
    if (mask & 2u) v.y = t.y;
    if (mask & 4u) v.z = t.z;
    if (mask & 8u) v.w = t.w;
    return v;
}

// __mixed_partial_c / _bd (AMR.h:1221-1230) on LDS rows: c = centre, the
// other arguments are the neighbours named by their offsets.
__device__ __forceinline__ float mixed_c(float pp, float mm, float pm, float mp)
{
    return .25f * (pp + mm - pm - mp);
}
__device__ __forceinline__ float mixed_bd(float c, float pp, float mm, float m0, float p0, float zm, float zp)
{
    return c + .5f * (pp + mm - m0 - p0 - zm - zp);
}

// DIR: 0 the tile's own ghost layers only; 1 direct reads across z faces
// (a base-pointer swap per ghost plane); 2 direct reads on any axis mask
// (padded layout); 3 direct reads across y and/or z faces (interior-first
// layout: a y ghost row is a whole interior row of the neighbour).
// Separate instantiations, so the plain kernel carries none of the code.
// L: tile layout of the fields (0 padded, 1 interior-first).  The LDS ring
// holds padded planes either way; with L = 1 a plane is assembled from the
// interior plane (S*S contiguous floats, one per thread) and the Q slab rows
// of each x and y face (S contiguous floats each), a z ghost plane from the
// slab layer (or the z neighbour's interior plane), and ghost layers beyond
// the stencil's half-width Q are not loaded at all.
// ADV: also the upwinded advection + Kreiss-Oliger output (b.adv), which
// reads three ghost layers and the shift at each output cell.
template <int S, int P, int Q, int ZB, bool PERSIST, int FMT, int PF, int DIR, int L, bool ADV = false>
__global__ void __launch_bounds__(S *S) k_fd(FdBatchDev b, MeshDev m, DepthScales sc)
{
    static_assert(L == 0 || DIR != 2, "the interior-first layout takes direct y and z faces only");
    static_assert(DIR != 3 || L == 1, "direct y faces are built for the interior-first layout");
    static_assert(!ADV || P >= 3, "the lopsided and sixth-difference stencils reach 3 cells");
    constexpr int QL = ADV ? 3 : Q;  // ghost layers the stencils read
    constexpr int T = S + 2 * P, T2 = T * T, R = 2 * P + 2, NZB = S / ZB;
    constexpr size_t TS = tile_stride<S, P, L>();
    constexpr FdCoef C = fd_coef();
    constexpr int BASE = Q * Q + Q - 1;
    __shared__ float4 ring4[R][T2 / 4];
    const int var = blockIdx.y;
    const int32_t gs_d = b.seed_table ? b.seed_table[b.seed_d[var]] : b.seed_d[var];
    const int32_t gs_m = b.seed_table ? b.seed_table[b.seed_m[var]] : b.seed_m[var];

    static_assert(S % 4 == 0, "direct reads shift float4 indices by S/4");
    uint32_t gmask[(T2 / 4 + S * S - 1) / (S * S)] = {};
    if constexpr (DIR == 2) {
#pragma unroll
        for (int j = 0; j < (T2 / 4 + S * S - 1) / (S * S); j++)
            gmask[j] = (int)threadIdx.x + j * S * S < T2 / 4 ? face_ghost_masks<S, P, Q>((int)threadIdx.x + j * S * S)
                                                             : 0u;
    }

    // one work item = one (tile, z-block); a persistent launch loops over
    // them (a separate instantiation: the loop keeps kernel arguments live)
    const int item_end = PERSIST ? m.n_items : (int)blockIdx.x + 1;
    for (int item = blockIdx.x; item < item_end; item += PERSIST ? gridDim.x : 1) {
    const int tile_i = item / NZB, zb = item - tile_i * NZB;
    const int slot = m.tiles[tile_i];
    const int depth = pos_depth(m.positions[slot]);
    if (depth == -1 || !active_depth(m.substep, (uint32_t)depth, m.max_depth)) continue;  // uniform per workgroup

    const float *src = b.field[var] + (size_t)slot * TS;
    // face f (0-2: +x +y +z, 3-5: -x -y -z) -> the equal-depth neighbour's
    // tile, or null: its ghost layers are read where they lie
    const float *nb[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    if (DIR != 0) {
#pragma unroll
        for (int f = 0; f < 6; f++) {
            const uint32_t nd = (uint32_t)m.direct[(size_t)slot * 6 + f];
            const int tgt = bnd_target(nd);
            if (((m.direct_axes >> (f % 3)) & 1u) && tgt >= 0 && bnd_depth(nd) > 0 && bnd_ddiff(nd) == 0u)
                nb[f] = b.field[var] + (size_t)tgt * TS;
        }
    }
    uint32_t *o_d1 = b.d1[var], *o_d2 = b.d2diag[var], *o_off = b.d2off[var];
    const bool want_d = o_d1 != nullptr, want_m = o_off != nullptr;
    const float dx1 = sc.inv_dx[depth];
    const float dx2 = dx1 * dx1;

    const int tid = threadIdx.x;
    const int x = tid % S, y = tid / S;
    const int z0 = zb * ZB;
    const int c = (y + P) * T + (x + P);  // centre offset within a plane

    // Planes are prefetched PF iterations ahead into registers (NL float4
    // per thread per plane), so the HBM latency of the next planes overlaps
    // the compute of plane z.  At iteration z the ring holds padded planes
    // z..z+2P; the slot of plane z+2P+1 is (z-1) % R, which nobody reads in
    // iteration z, so one barrier per iteration suffices.  Only planes
    // z0..z0+ZB-1+2P are ever read.
    //
    // With direct reads, a ghost plane across a z face is the neighbour's
    // interior plane S away, and in an interior plane the ghost columns
    // (rows) across an x (y) face are the neighbour's interior columns (rows)
    // S away: the same float4 index shifted by S/4 (S*T/4), as S % 4 == 0.
    constexpr int NL = (T2 / 4 + S * S - 1) / (S * S);
    float4 pre[PF][NL];
    // interior-first layout: this thread's element of the x/y face slab rows
    // of an interior plane (4 faces x Q layers x S cells), if any: its offset
    // from the slab region (plus z * S per plane) and its place in the plane
    constexpr int NF = 4 * QL * S;
    int f_goff = -1, f_loff = 0;
    int f_yface = -1, f_yrow = 0;  // DIR 3: a y face's element = the neighbour's interior row f_yrow, cell a
    if constexpr (L == 1) {
        if (tid < NF) {
            const int k = tid / (QL * S), l = (tid / S) % QL, a = tid % S;
            const int f = k == 0 ? 3 : (k == 1 ? 0 : (k == 2 ? 4 : 1));  // -x +x -y +y
            f_goff = (f * P + l) * S * S + a;
            if (k >= 2) {
                f_yface = f;
                f_yrow = (f == 4 ? S - 1 - l : l) * S + a;
            }
            const int g = (k & 1) ? P + S + l : P - 1 - l;  // the ghost coordinate of the face axis
            f_loff = k < 2 ? (a + P) * T + g : g * T + a + P;
        }
    }
    // DIR 3: this thread's y-face element comes from the neighbour's interior
    const float *ynb = nullptr;
    if constexpr (DIR == 3) ynb = f_yface == 4 ? nb[4] : (f_yface == 1 ? nb[1] : nullptr);
    auto gload = [&](float4 *dst, int zp) {
        if constexpr (L == 1) {
            if (zp >= P && zp < P + S) {
                const float *pz = src + (size_t)(zp - P) * (S * S);
                dst[0].x = pz[tid];
                if (f_goff >= 0) {
                    if (DIR == 3 && ynb)
                        dst[0].y = ynb[(size_t)(zp - P) * (S * S) + f_yrow];
                    else
                        dst[0].y = src[(size_t)(S * S * S) + f_goff + (zp - P) * S];
                }
            } else {
                const bool low = zp < P;
                const int l = low ? P - 1 - zp : zp - P - S;
                if (l < QL) {
                    const float *pl_ = src + (size_t)(S * S * S) + (size_t)(((low ? 5 : 2) * P + l) * S * S);
                    if constexpr (DIR == 1 || DIR == 3) {
                        if (low && nb[5]) pl_ = nb[5] + (size_t)(S - 1 - l) * (S * S);
                        if (!low && nb[2]) pl_ = nb[2] + (size_t)l * (S * S);
                    }
                    dst[0].x = pl_[tid];
                }
            }
            return;
        }
        const float *base = src;
        int zs = zp;
        if constexpr (DIR != 0) {
            if (zp < P && nb[5]) {
                base = nb[5];
                zs = zp + S;
            } else if (zp >= P + S && nb[2]) {
                base = nb[2];
                zs = zp - S;
            }
        }
        const float4 *s4 = reinterpret_cast<const float4 *>(base + (size_t)zs * T2);
        const bool inner = DIR == 2 && zp >= P && zp < P + S;
#pragma unroll
        for (int j = 0; j < NL; j++) {
            const int i = tid + j * S * S;
            if (i < T2 / 4) {
                float4 v = s4[i];
                const uint32_t gm = gmask[j];
                if (inner && gm) {
                    const size_t pz = (size_t)zp * (T2 / 4);
                    if ((gm & 15u) && nb[3]) v = take(v, reinterpret_cast<const float4 *>(nb[3])[pz + i + S / 4], gm);
                    if ((gm & 240u) && nb[0])
                        v = take(v, reinterpret_cast<const float4 *>(nb[0])[pz + i - S / 4], gm >> 4);
                    if ((gm & 3840u) && nb[4])
                        v = take(v, reinterpret_cast<const float4 *>(nb[4])[pz + i + S * T / 4], gm >> 8);
                    if ((gm & 61440u) && nb[1])
                        v = take(v, reinterpret_cast<const float4 *>(nb[1])[pz + i - S * T / 4], gm >> 12);
                }
                dst[j] = v;
            }
        }
    };
    auto lstore = [&](const float4 *from, int zp) {
        if constexpr (L == 1) {
            float *d = reinterpret_cast<float *>(ring4[zp % R]);
            if (zp >= P && zp < P + S) {
                d[(tid / S + P) * T + tid % S + P] = from[0].x;
                if (f_goff >= 0) d[f_loff] = from[0].y;
            } else if ((zp < P ? P - 1 - zp : zp - P - S) < QL) {
                d[(tid / S + P) * T + tid % S + P] = from[0].x;
            }
            return;
        }
        float4 *d4 = ring4[zp % R];
#pragma unroll
        for (int j = 0; j < NL; j++) {
            const int i = tid + j * S * S;
            if (i < T2 / 4) d4[i] = from[j];
        }
    };
    for (int p = 0; p <= 2 * P; p++) {
        gload(pre[0], z0 + p);
        lstore(pre[0], z0 + p);
    }
#pragma unroll
    for (int j = 0; j < PF; j++)
        if (j + 1 < ZB) gload(pre[j], z0 + 2 * P + 1 + j);
    __syncthreads();

    // boundary offsets and their products are -1, 0 or 1: 24-bit multiplies
    const int ex = x - min(max(x, 1), S - 2);
    const int ey = y - min(max(y, 1), S - 2);
    const int cxy = __mul24(ex, ey);
    // dither seeds: the (x, y) part once, the z mix per plane
    const int32_t sxy_d = thread_seed_xy(gs_d, (uint32_t)x, (uint32_t)y);
    const int32_t sxy_m = thread_seed_xy(gs_m, (uint32_t)x, (uint32_t)y);

    // ADV: the shift at the output cell, loaded one plane ahead
    float bsh[3] = {0.f, 0.f, 0.f};
    float *o_adv = nullptr;
    float kos = 0.f;
    auto shift_at = [&](int zz) {
        const uint32_t ib = 4u * (uint32_t)((size_t)slot * TS + tile_interior<S, P, L>((uint32_t)x, (uint32_t)y, (uint32_t)zz));
#pragma unroll
        for (int k = 0; k < 3; k++) bsh[k] = *reinterpret_cast<const float *>(reinterpret_cast<const char *>(b.shift[k]) + ib);
    };
    if constexpr (ADV) {
        o_adv = b.adv[var];
        if (o_adv) {  // uniform per workgroup
            shift_at(z0);
            kos = (b.ko_sigma * (1.f / 64.f)) * dx1;
        }
    }
    const bool advect = ADV && ((b.advect_mask >> var) & 1u);

    for (int z = z0; z < z0 + ZB; z++) {
        // padded plane of output z is z+P; plane z+P+i sits in slot (z+P+i) % R
        const float *pl[2 * P + 1];
#pragma unroll
        for (int i = 0; i <= 2 * P; i++) pl[i] = reinterpret_cast<const float *>(ring4[(z + i) % R]);
        const float *ctr = pl[P];
        const uint32_t gz = (uint32_t)(slot * S + z);
        // the word index (the word arrays stay far below 2^32 cells)
        const uint32_t o = (gz * (uint32_t)S + (uint32_t)y) * (uint32_t)S + (uint32_t)x;

        // The reference sums the 2Q+1 products from 0 in offset order.  For
        // cf3 words the sums start from the first product instead: that
        // changes only the sign of a sum whose products are all zero, and
        // +0 and -0 pack to the same word (exponent field 0, and the dither
        // added to a zero of either sign gives the same value).  The study
        // formats store the value itself, so they keep the 0 + start.
        constexpr bool FROM0 = FMT != WORD_CF3;
        if (want_d) {
            float vx = ctr[c - Q] * C.d1[BASE - Q], vy = ctr[c - Q * T] * C.d1[BASE - Q],
                  vz = pl[P - Q][c] * C.d1[BASE - Q];
            if constexpr (FROM0) {
                vx = 0.f + vx; vy = 0.f + vy; vz = 0.f + vz;
            }
#pragma unroll
            for (int i = -Q + 1; i <= Q; i++) {
                vx = vx + ctr[c + i] * C.d1[BASE + i];
                vy = vy + ctr[c + i * T] * C.d1[BASE + i];
                vz = vz + pl[P + i][c] * C.d1[BASE + i];
            }
            vx = vx * dx1; vy = vy * dx1; vz = vz * dx1;
            int32_t rng = thread_seed_z(sxy_d, gz);
            if (o_d2) {  // uniform: the second-derivative sums only where they are stored
                float xx = ctr[c - Q] * C.d2[BASE - Q], yy = ctr[c - Q * T] * C.d2[BASE - Q],
                      zz = pl[P - Q][c] * C.d2[BASE - Q];
                if constexpr (FROM0) {
                    xx = 0.f + xx; yy = 0.f + yy; zz = 0.f + zz;
                }
#pragma unroll
                for (int i = -Q + 1; i <= Q; i++) {
                    xx = xx + ctr[c + i] * C.d2[BASE + i];
                    yy = yy + ctr[c + i * T] * C.d2[BASE + i];
                    zz = zz + pl[P + i][c] * C.d2[BASE + i];
                }
                word_put<FMT>(o_d2, o, xx * dx2, yy * dx2, zz * dx2, rng);
            }
            word_put<FMT>(o_d1, o, vx, vy, vz, rng);
        }
        if (want_m) {
            const int ez = z - min(max(z, 1), S - 2);
            const int cxz = __mul24(ex, ez), cyz = __mul24(ey, ez);
            const float *zm = pl[P - 1], *zp = pl[P + 1];
            float vxy, vxz, vyz;
            if (cxy == 0) {
                vxy = mixed_c(ctr[c + 1 + T], ctr[c - 1 - T], ctr[c + 1 - T], ctr[c - 1 + T]);
            } else {
                const int dx = -cxy;  // +-1 along x, dy = +T
                vxy = (float)(-cxy) * mixed_bd(ctr[c], ctr[c + dx + T], ctr[c - dx - T], ctr[c - dx], ctr[c + dx],
                                               ctr[c - T], ctr[c + T]);
            }
            if (cxz == 0) {
                vxz = mixed_c(zp[c + 1], zm[c - 1], zm[c + 1], zp[c - 1]);
            } else {
                const int dx = -cxz;  // +-1 along x, dy = one plane up
                vxz = (float)(-cxz) * mixed_bd(ctr[c], zp[c + dx], zm[c - dx], ctr[c - dx], ctr[c + dx], zm[c], zp[c]);
            }
            if (cyz == 0) {
                vyz = mixed_c(zp[c + T], zm[c - T], zm[c + T], zp[c - T]);
            } else {
                const int dy = __mul24(-cyz, T);  // +-1 along y, dz = one plane up
                vyz = (float)(-cyz) * mixed_bd(ctr[c], zp[c + dy], zm[c - dy], ctr[c - dy], ctr[c + dy], zm[c], zp[c]);
            }
            int32_t rng = thread_seed_z(sxy_m, gz);
            word_put<FMT>(o_off, o, vxy * dx2, vxz * dx2, vyz * dx2, rng);
        }
        if constexpr (ADV) {
            if (o_adv) {
                // sum_k shift^k L_k f and the sixth differences, axes x, y, z,
                // offsets in increasing order (the oracle's orc_fd_adv)
                float adv = 0.f, ko = 0.f;
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    auto fm = [&](int mm) { return k == 0 ? ctr[c + mm] : (k == 1 ? ctr[c + mm * T] : pl[P + mm][c]); };
                    float l = 0.f, kk = 0.f;
                    if (bsh[k] > 0.f) {
#pragma unroll
                        for (int mm = -1; mm <= 3; mm++) l = l + fm(mm) * LOP_PLUS[mm + 1];
                    } else {
#pragma unroll
                        for (int mm = -3; mm <= 1; mm++) l = l + fm(mm) * LOP_MINUS[mm + 3];
                    }
#pragma unroll
                    for (int mm = -3; mm <= 3; mm++) kk = kk + fm(mm) * KO6[mm + 3];
                    adv = adv + bsh[k] * (l * dx1);
                    ko = ko + kk;
                }
                *reinterpret_cast<float *>(reinterpret_cast<char *>(o_adv) + 4u * o) = (advect ? adv : 0.f) + ko * kos;
                if (z + 1 < z0 + ZB) shift_at(z + 1);
            }
        }
        if (z + 1 < z0 + ZB) {
            lstore(pre[0], z + 2 * P + 1);
#pragma unroll
            for (int j = 0; j + 1 < PF; j++)
#pragma unroll
                for (int k = 0; k < NL; k++) pre[j][k] = pre[j + 1][k];
            if (z + 1 + PF < z0 + ZB) gload(pre[PF - 1], z + 2 * P + 1 + PF);
        }
        __syncthreads();
    }
    }  // work items
}

template <int S, int P, int Q, int ZB, int DIR, int FMT, int PF, int L>
static void fd_launch_pf(bool persist, dim3 grid, hipStream_t st, const FdBatchDev &d, const MeshDev &m,
                         const DepthScales &sc)
{
    if (persist)
        hipLaunchKernelGGL((k_fd<S, P, Q, ZB, true, FMT, PF, DIR, L>), grid, dim3(S * S), 0, st, d, m, sc);
    else
        hipLaunchKernelGGL((k_fd<S, P, Q, ZB, false, FMT, PF, DIR, L>), grid, dim3(S * S), 0, st, d, m, sc);
}
template <int S, int P, int Q, int ZB, int DIR, int FMT, int L>
static void fd_launch_fmt(int pf, bool persist, dim3 grid, hipStream_t st, const FdBatchDev &d, const MeshDev &m,
                          const DepthScales &sc)
{
    if constexpr (L == 0) {  // the interior-first layout keeps the one-plane prefetch
        if (pf == 2) return fd_launch_pf<S, P, Q, ZB, DIR, FMT, 2, L>(persist, grid, st, d, m, sc);
        if (pf == 3) return fd_launch_pf<S, P, Q, ZB, DIR, FMT, 3, L>(persist, grid, st, d, m, sc);
    }
    fd_launch_pf<S, P, Q, ZB, DIR, FMT, 1, L>(persist, grid, st, d, m, sc);
}

// the fp32 / fp16 word formats (precision study) exist for the S = 16, P = 3,
// order-4 tiles only
template <int S, int P, int Q>
constexpr bool has_word_formats()
{
    return S == 16 && P == 3 && Q == 2;
}

template <int S, int P, int Q, int ZB, int DIR, int L>
static void fd_launch_dir(int fmt, int pf, bool persist, dim3 grid, hipStream_t st, const FdBatchDev &d,
                          const MeshDev &m, const DepthScales &sc)
{
    if constexpr (has_word_formats<S, P, Q>()) {
        // the study formats keep the one-plane prefetch
        if (fmt == WORD_F32) return fd_launch_pf<S, P, Q, ZB, DIR, WORD_F32, 1, L>(persist, grid, st, d, m, sc);
        if (fmt == WORD_F16) return fd_launch_pf<S, P, Q, ZB, DIR, WORD_F16, 1, L>(persist, grid, st, d, m, sc);
    }
    fd_launch_fmt<S, P, Q, ZB, DIR, WORD_CF3, L>(pf, persist, grid, st, d, m, sc);
}
template <int S, int P, int Q, int ZB, int L>
static void fd_launch(int fmt, int pf, bool persist, dim3 grid, hipStream_t st, const FdBatchDev &d,
                      const MeshDev &m, const DepthScales &sc)
{
    if (m.direct_axes == 0u) return fd_launch_dir<S, P, Q, ZB, 0, L>(fmt, pf, persist, grid, st, d, m, sc);
    if (m.direct_axes == 4u) return fd_launch_dir<S, P, Q, ZB, 1, L>(fmt, pf, persist, grid, st, d, m, sc);
    if constexpr (L == 0) fd_launch_dir<S, P, Q, ZB, 2, L>(fmt, pf, persist, grid, st, d, m, sc);
    else fd_launch_dir<S, P, Q, ZB, 3, L>(fmt, pf, persist, grid, st, d, m, sc);  // y, or y and z
}
// the upwind / dissipation kernels: S = 16, P = 3, order 4, cf3 words,
// one-plane prefetch, direct faces 0, z, or (interior-first) y and/or z
template <int S, int P, int Q, int ZB, int L>
static void fd_launch_adv_l(bool persist, dim3 grid, hipStream_t st, const FdBatchDev &d, const MeshDev &m,
                            const DepthScales &sc)
{
    if constexpr (S == 16 && P == 3 && Q == 2) {
        if (m.direct_axes == 4u) {
            if (persist)
                hipLaunchKernelGGL((k_fd<S, P, Q, ZB, true, WORD_CF3, 1, 1, L, true>), grid, dim3(S * S), 0, st, d, m, sc);
            else
                hipLaunchKernelGGL((k_fd<S, P, Q, ZB, false, WORD_CF3, 1, 1, L, true>), grid, dim3(S * S), 0, st, d, m, sc);
        } else if (L == 1 && m.direct_axes != 0u) {
            if (persist)
                hipLaunchKernelGGL((k_fd<S, P, Q, ZB, true, WORD_CF3, 1, 3, 1, true>), grid, dim3(S * S), 0, st, d, m, sc);
            else
                hipLaunchKernelGGL((k_fd<S, P, Q, ZB, false, WORD_CF3, 1, 3, 1, true>), grid, dim3(S * S), 0, st, d, m, sc);
        } else {
            if (persist)
                hipLaunchKernelGGL((k_fd<S, P, Q, ZB, true, WORD_CF3, 1, 0, L, true>), grid, dim3(S * S), 0, st, d, m, sc);
            else
                hipLaunchKernelGGL((k_fd<S, P, Q, ZB, false, WORD_CF3, 1, 0, L, true>), grid, dim3(S * S), 0, st, d, m, sc);
        }
    }
}

template <int S, int P, int Q, int ZB>
static void fd_launch_layout(int layout, int fmt, int pf, bool persist, dim3 grid, hipStream_t st,
                             const FdBatchDev &d, const MeshDev &m, const DepthScales &sc, bool adv)
{
    if (adv) {
        if (layout == AM_LAYOUT_INTERIOR_FIRST) return fd_launch_adv_l<S, P, Q, ZB, 1>(persist, grid, st, d, m, sc);
        return fd_launch_adv_l<S, P, Q, ZB, 0>(persist, grid, st, d, m, sc);
    }
    if (layout == AM_LAYOUT_INTERIOR_FIRST) return fd_launch<S, P, Q, ZB, 1>(fmt, pf, persist, grid, st, d, m, sc);
    fd_launch<S, P, Q, ZB, 0>(fmt, pf, persist, grid, st, d, m, sc);
}

template <int S, int P, int Q>
static int launch_fd(const am_fd_batch *b, const am_mesh *mesh, const am_geometry *g, void *stream)
{
    FdBatchDev d{};
    for (int v = 0; v < b->n_vars; v++) {
        d.field[v] = b->field[v];
        d.d1[v] = b->d1[v];
        d.d2diag[v] = b->d2diag[v];
        d.d2off[v] = b->d2off[v];
        d.seed_d[v] = b->seed_d[v];
        d.seed_m[v] = b->seed_m[v];
        if (b->seed_table && (b->seed_d[v] < 0 || b->seed_m[v] < 0))
            return set_error(AM_ERR_INVALID, "am_fd: negative seed index for variable %d", v);
        if (!b->field[v]) return set_error(AM_ERR_INVALID, "am_fd: null field for variable %d", v);
        if (!aligned16(b->field[v])) return set_error(AM_ERR_MISALIGNED, "am_fd: field %d not 16-byte aligned", v);
        if (!b->d1[v] && !b->d2off[v]) return set_error(AM_ERR_INVALID, "am_fd: variable %d has no output", v);
        if (!b->d1[v] && b->d2diag[v])
            return set_error(AM_ERR_INVALID, "am_fd: variable %d: diagonal Hessian needs the gradient word", v);
    }
    d.seed_table = b->seed_table;
    const DepthScales sc = depth_scales(*g);
    int zb = b->z_block;
    if (zb == 0) zb = S;  // whole tile per workgroup: no plane is read twice
    hipStream_t st = (hipStream_t)stream;
    // grid.x = work items, capped by the persistent limit (split over the
    // variables on grid.y)
    auto items = [&](int nzb) {
        return MeshDev{mesh->positions, mesh->tiles, mesh->substep, mesh->max_depth, mesh->n_tiles * nzb,
                       b->direct_faces ? mesh->boundaries : nullptr, (uint32_t)b->direct_faces};
    };
    if (b->direct_faces < 0 || b->direct_faces > 7)
        return set_error(AM_ERR_INVALID, "am_fd: direct_faces %d (axis mask 0..7)", b->direct_faces);
    if (b->direct_faces && !mesh->boundaries)
        return set_error(AM_ERR_INVALID, "am_fd: direct_faces needs mesh->boundaries");
    auto gx = [&](int nzb) {
        const int n = mesh->n_tiles * nzb;
        if (mesh->grid_limit <= 0) return n;
        const int lim = (mesh->grid_limit + b->n_vars - 1) / b->n_vars;
        return lim < n ? (lim > 0 ? lim : 1) : n;
    };
    const bool persist = mesh->grid_limit > 0;
    const int fmt = b->word_format;
    const int pf = b->prefetch_planes <= 1 ? 1 : b->prefetch_planes;
    if (pf > 3) return set_error(AM_ERR_INVALID, "am_fd: prefetch_planes %d (0-3)", b->prefetch_planes);
    if (fmt != WORD_CF3 && !(has_word_formats<S, P, Q>() && (fmt == WORD_F32 || fmt == WORD_F16)))
        return set_error(AM_ERR_INVALID, "am_fd: word format %d needs S=16, P=3, fd_order 4", fmt);
    const int lay = b->tile_layout;
    if (lay != AM_LAYOUT_PADDED && lay != AM_LAYOUT_INTERIOR_FIRST)
        return set_error(AM_ERR_INVALID, "am_fd: tile layout %d", lay);
    bool adv = false;
    for (int v = 0; v < b->n_vars; v++) {
        d.adv[v] = b->adv[v];
        adv |= b->adv[v] != nullptr;
    }
    if (adv) {
        if (!(S == 16 && P == 3 && Q == 2) || fmt != WORD_CF3 || pf > 1 ||
            (b->direct_faces != 0 && b->direct_faces != 4 &&
             (b->tile_layout != AM_LAYOUT_INTERIOR_FIRST || (b->direct_faces & 1))))
            return set_error(AM_ERR_INVALID,
                             "am_fd: the upwind / dissipation output needs S=16, P=3, fd_order 4, cf3 words, "
                             "prefetch 1 and direct_faces 0 or 4 (or 2 and 6 in the interior-first layout)");
        if (!b->shift[0] || !b->shift[1] || !b->shift[2])
            return set_error(AM_ERR_INVALID, "am_fd: the upwind output needs the three shift fields");
        if (!(b->ko_sigma >= 0.f)) return set_error(AM_ERR_INVALID, "am_fd: ko_sigma %g", (double)b->ko_sigma);
        for (int k = 0; k < 3; k++) d.shift[k] = b->shift[k];
        d.advect_mask = b->advect_mask;
        d.ko_sigma = b->ko_sigma;
    }
    if (lay == AM_LAYOUT_INTERIOR_FIRST && (pf > 1 || (b->direct_faces & 1)))
        return set_error(AM_ERR_INVALID,
                         "am_fd: the interior-first layout takes prefetch_planes 0-1 and direct_faces 0, 2, 4 or 6 "
                         "(%d, %d)", b->prefetch_planes, b->direct_faces);
    if (zb == S) {
        fd_launch_layout<S, P, Q, S>(lay, fmt, pf, persist, dim3(gx(1), b->n_vars), st, d, items(1), sc, adv);
    } else if (zb == S / 2) {
        fd_launch_layout<S, P, Q, S / 2>(lay, fmt, pf, persist, dim3(gx(2), b->n_vars), st, d, items(2), sc, adv);
    } else if (zb == S / 4 && (S % 4) == 0) {
        fd_launch_layout<S, P, Q, S / 4>(lay, fmt, pf, persist, dim3(gx(4), b->n_vars), st, d, items(4), sc, adv);
    } else {
        return set_error(AM_ERR_INVALID, "am_fd: z_block %d must be S, S/2 or S/4", zb);
    }
    return check_launch("k_fd");
}

// table[i] = draw i + 1 of fast_prng from `seed`: the draws the host would
// make, in its order, for the launches of a recorded graph
__global__ void __launch_bounds__(64) k_seed_table(int32_t *table, int32_t n, int32_t seed)
{
    if (threadIdx.x != 0) return;
    int32_t s = seed;
    for (int32_t i = 0; i < n; i++) table[i] = prng_next(s);
}

static int fd_checks(const am_mesh *mesh, const am_geometry *g)
{
    if (!mesh || !g) return set_error(AM_ERR_INVALID, "am_fd: null mesh or geometry");
    if (!mesh->positions || !mesh->tiles) return set_error(AM_ERR_INVALID, "am_fd: null positions or tiles");
    if (mesh->n_tiles < 0) return set_error(AM_ERR_INVALID, "am_fd: negative tile count");
    return AM_OK;
}

}  // namespace am

extern "C" int am_fd_batched(const am_fd_batch *b, const am_mesh *mesh, const am_geometry *g, void *stream)
{
    if (!b) return am::set_error(AM_ERR_INVALID, "am_fd_batched: null batch");
    if (b->n_vars < 0 || b->n_vars > AM_MAX_VARS)
        return am::set_error(AM_ERR_TOO_MANY_VARS, "am_fd_batched: n_vars %d", b->n_vars);
    if (int rc = am::fd_checks(mesh, g)) return rc;
    if (b->n_vars == 0 || mesh->n_tiles == 0) return AM_OK;
    AM_DISPATCH_GEOMETRY(*g, am::launch_fd, b, mesh, g, stream);
}

static int fd_single(const float *field, uint32_t *d1, uint32_t *d2diag, uint32_t *d2off, const am_mesh *mesh,
                     const am_geometry *g, int32_t seed_d, int32_t seed_m, void *stream)
{
    am_fd_batch b = {};
    b.field[0] = field;
    b.d1[0] = d1;
    b.d2diag[0] = d2diag;
    b.d2off[0] = d2off;
    b.seed_d[0] = seed_d;
    b.seed_m[0] = seed_m;
    b.n_vars = 1;
    return am_fd_batched(&b, mesh, g, stream);
}

extern "C" int am_fd_d1(const float *field, uint32_t *d1, const am_mesh *mesh, const am_geometry *g, int32_t seed,
                        void *stream)
{
    if (!d1) return am::set_error(AM_ERR_INVALID, "am_fd_d1: null output");
    return fd_single(field, d1, nullptr, nullptr, mesh, g, seed, 0, stream);
}

extern "C" int am_fd_d1_d2diag(const float *field, uint32_t *d1, uint32_t *d2diag, const am_mesh *mesh,
                               const am_geometry *g, int32_t seed, void *stream)
{
    if (!d1 || !d2diag) return am::set_error(AM_ERR_INVALID, "am_fd_d1_d2diag: null output");
    return fd_single(field, d1, d2diag, nullptr, mesh, g, seed, 0, stream);
}

extern "C" int am_fd_mixed(const float *field, uint32_t *d2off, const am_mesh *mesh, const am_geometry *g,
                           int32_t seed, void *stream)
{
    if (!d2off) return am::set_error(AM_ERR_INVALID, "am_fd_mixed: null output");
    return fd_single(field, nullptr, nullptr, d2off, mesh, g, 0, seed, stream);
}

extern "C" int am_fd_all(const float *field, uint32_t *d1, uint32_t *d2diag, uint32_t *d2off, const am_mesh *mesh,
                         const am_geometry *g, int32_t seed_d, int32_t seed_m, void *stream)
{
    if (!d1 || !d2diag || !d2off) return am::set_error(AM_ERR_INVALID, "am_fd_all: null output");
    return fd_single(field, d1, d2diag, d2off, mesh, g, seed_d, seed_m, stream);
}

extern "C" int am_seed_table_fill(int32_t *table, int32_t n, int32_t seed, void *stream)
{
    if (!table || n < 0) return am::set_error(AM_ERR_INVALID, "am_seed_table_fill: bad table");
    if (n == 0) return AM_OK;
    hipLaunchKernelGGL(am::k_seed_table, dim3(1), dim3(64), 0, (hipStream_t)stream, table, n, seed);
    return am::check_launch("k_seed_table");
}
