// BSSN stage and constraint kernels, templated on the tile geometry and on
// the storage format of the derivative words (am_words.hpp).  Included by
// am_bssn.hip (reference compressed_float3 words, every geometry) and
// am_bssn_words.hip (fp32 / fp16 words, S = 16, P = 3 tiles), so the two
// sets of instantiations compile in parallel.
#pragma once
#include "am_device.hpp"
#include "am_rk4.hpp"
#include "am_words.hpp"

namespace am {


constexpr int NV = 24, NH = 11;

struct BssnArgsDev {
    const float *state[NV];
    const float *old_[NV];
    float *new_[NV];
    float *acc[NV];
    const void *d1[NV];  // word arrays in format FMT (am_words.hpp)
    const void *d2d[NH];
    const void *d2o[NH];
    int push_layers;
    int push_axes;  // bits 1 x, 2 y, 4 z
    const float *adv[NV];  // upwinded advection + dissipation per field (UPW kernels)
};

struct BssnParDev {
    float eta, shift_coeff, chi_floor;
    int advect_shift, slicing;
};

// symmetric storage xx,xy,xz,yy,yz,zz
__device__ constexpr int SI[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
__device__ constexpr float MULT[6] = {1.f, 2.f, 2.f, 1.f, 2.f, 1.f};

template <int FMT>
__device__ __forceinline__ void dec3(const void *base, size_t w, float *o)
{
    const float3 v = word_get<FMT>(base, w);
    o[0] = v.x; o[1] = v.y; o[2] = v.z;
}
template <int FMT>
__device__ __forceinline__ void hess6(const void *bd, const void *bo, size_t w, float *h)
{
    const float3 d = word_get<FMT>(bd, w), o = word_get<FMT>(bo, w);
    h[0] = d.x; h[1] = o.x; h[2] = o.y; h[3] = d.y; h[4] = o.z; h[5] = d.z;
}
// sum_ij a^ij b_ij for symmetric a, b
__device__ __forceinline__ float sym_dot(const float *a, const float *b)
{
    float s = a[0] * b[0];
#pragma unroll
    for (int c = 1; c < 6; c++) s += MULT[c] * a[c] * b[c];
    return s;
}

// The derivative words of one cell.  compressed_float3 words (4 bytes) are
// all loaded by load(), up front, so that a wave has its 46 word loads in
// flight together instead of one memory round trip per word decoded; the
// study formats (12 and 6 bytes per word) are read where they are used.
// The cell's loads and stores address each array as its base pointer (a
// kernel argument, in scalar registers) plus a 32-bit byte offset per lane
// that every array of the same kind shares, so no lane does 64-bit address
// arithmetic per array.  A byte offset fits 32 bits because one field array
// of a stage call stays below 4 GiB: the 24 fields' state, y_n, new and
// accumulator plus the 46 word arrays of n slots take more than 72 times one
// field array, beyond the HBM of one GPU long before a field reaches 4 GiB
// (the drivers check it when they allocate).
template <class T>
__device__ __forceinline__ T ld_at(const T *base, uint32_t byte_off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + byte_off);
}
template <class T>
__device__ __forceinline__ void st_at(T *base, uint32_t byte_off, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + byte_off) = v;
}

template <int FMT>
struct CellWords {
    uint32_t d1[NV], dd[NH], doff[NH];
    // wo: the cell's byte offset in every word array
    __device__ __forceinline__ void load(const BssnArgsDev &a, uint32_t wo)
    {
        if constexpr (FMT == WORD_CF3) {
#pragma unroll
            for (int v = 0; v < NV; v++) d1[v] = ld_at(static_cast<const uint32_t *>(a.d1[v]), wo);
#pragma unroll
            for (int h = 0; h < NH; h++) {
                dd[h] = ld_at(static_cast<const uint32_t *>(a.d2d[h]), wo);
                doff[h] = ld_at(static_cast<const uint32_t *>(a.d2o[h]), wo);
            }
        }
    }
    __device__ __forceinline__ float3 grad(const BssnArgsDev &a, size_t w, int v) const
    {
        if constexpr (FMT == WORD_CF3) return cf3_unpack(d1[v]);
        else return word_get<FMT>(a.d1[v], w);
    }
    __device__ __forceinline__ float3 diag(const BssnArgsDev &a, size_t w, int h) const
    {
        if constexpr (FMT == WORD_CF3) return cf3_unpack(dd[h]);
        else return word_get<FMT>(a.d2d[h], w);
    }
    __device__ __forceinline__ float3 off(const BssnArgsDev &a, size_t w, int h) const
    {
        if constexpr (FMT == WORD_CF3) return cf3_unpack(doff[h]);
        else return word_get<FMT>(a.d2o[h], w);
    }
};

// RHS (and optionally the constraints) at one interior cell: u holds the
// 24 state values of the cell, cw its derivative words.
// UPW: aw holds each field's upwinded advection plus dissipation (the FD's
// adv output), which takes the place of the centred beta^k d_k terms; the B
// equation's d_0 Gt then drops aw of Gt (advection and dissipation) with
// advect_shift, and beta and B take aw whole (their dissipation, and their
// advection with advect_shift).
template <int FMT, bool WANT_RHS, bool WANT_CONS, bool UPW = false>
__device__ __forceinline__ void bssn_cell(const BssnArgsDev &a, const float (&u)[NV], const CellWords<FMT> &cw, size_t w,
                                          const BssnParDev &par, float *rhs, float *Hc, float *Mc,
                                          const float *aw = nullptr)
{
#pragma clang fp contract(fast)
    auto d1v = [&](int v, float *o) {
        const float3 t = cw.grad(a, w, v);
        o[0] = t.x; o[1] = t.y; o[2] = t.z;
    };
    auto hsv = [&](int hh, float *h) {
        const float3 d = cw.diag(a, w, hh), o = cw.off(a, w, hh);
        h[0] = d.x; h[1] = o.x; h[2] = o.y; h[3] = d.y; h[4] = o.z; h[5] = d.z;
    };
    // the two reciprocals by the hardware's (1 ulp; the RHS is not
    // bit-exact with anything, and correctly rounded division costs ten
    // instructions each)
    const float chi = u[0], ichi = __builtin_amdgcn_rcpf(fmaxf(chi, par.chi_floor));
    const float K = u[7], alpha = u[17];
    const float *g = u + 1, *At = u + 8, *Gam = u + 14, *beta = u + 18, *B = u + 21;

    // inverse conformal metric
    float gu[6];
    {
        const float c0 = g[3] * g[5] - g[4] * g[4], c1 = g[2] * g[4] - g[1] * g[5], c2 = g[1] * g[4] - g[2] * g[3];
        const float idet = __builtin_amdgcn_rcpf(g[0] * c0 + g[1] * c1 + g[2] * c2);
        gu[0] = c0 * idet;
        gu[1] = c1 * idet;
        gu[2] = c2 * idet;
        gu[3] = (g[0] * g[5] - g[2] * g[2]) * idet;
        gu[4] = (g[1] * g[2] - g[0] * g[4]) * idet;
        gu[5] = (g[0] * g[3] - g[1] * g[1]) * idet;
    }
    float dchi[3], dK[3], dal[3], db[3][3], dg[6][3];
    d1v(0, dchi);
    d1v(7, dK);
    d1v(17, dal);
#pragma unroll
    for (int i = 0; i < 3; i++) d1v(18 + i, db[i]);  // db[i][k] = d_k beta^i
#pragma unroll
    for (int c = 0; c < 6; c++) d1v(1 + c, dg[c]);
    const float divb = db[0][0] + db[1][1] + db[2][2];
    auto adv = [&](const float *d) { return beta[0] * d[0] + beta[1] * d[1] + beta[2] * d[2]; };
    auto A = [&](int v, const float *d) {
        if constexpr (UPW) return aw[v];
        else return adv(d);
    };

    // Gl[i][jk] = Gt_ijk = (d_k g_ij + d_j g_ik - d_i g_jk) / 2, G[i][jk] = Gt^i_jk
    float Gl[3][6], G[3][6];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
            for (int k = j; k < 3; k++) Gl[i][SI[j][k]] = .5f * (dg[SI[i][j]][k] + dg[SI[i][k]][j] - dg[SI[j][k]][i]);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int c = 0; c < 6; c++)
            G[i][c] = gu[SI[i][0]] * Gl[0][c] + gu[SI[i][1]] * Gl[1][c] + gu[SI[i][2]] * Gl[2][c];
    float Gd[3];
#pragma unroll
    for (int i = 0; i < 3; i++) Gd[i] = sym_dot(gu, G[i]);

    // At^i_j and At^ij
    float Atm[3][3], Atu[6];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++)
            Atm[i][j] = gu[SI[i][0]] * At[SI[0][j]] + gu[SI[i][1]] * At[SI[1][j]] + gu[SI[i][2]] * At[SI[2][j]];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = i; j < 3; j++)
            Atu[SI[i][j]] = Atm[i][0] * gu[SI[0][j]] + Atm[i][1] * gu[SI[1][j]] + Atm[i][2] * gu[SI[2][j]];
    const float AA = sym_dot(At, Atu);

    // gt^kl d_l chi, then its contractions with d chi and d alpha
    float uchi[3];
#pragma unroll
    for (int k = 0; k < 3; k++) uchi[k] = gu[SI[k][0]] * dchi[0] + gu[SI[k][1]] * dchi[1] + gu[SI[k][2]] * dchi[2];
    const float dchi_dal = uchi[0] * dal[0] + uchi[1] * dal[1] + uchi[2] * dal[2];
    const float dchi2 = uchi[0] * dchi[0] + uchi[1] * dchi[1] + uchi[2] * dchi[2];

    // conformal covariant Hessians of chi and alpha
    float DDchi[6], DDalt[6];
    {
        float h[6];
        hsv(0, h);
#pragma unroll
        for (int c = 0; c < 6; c++) DDchi[c] = h[c] - (G[0][c] * dchi[0] + G[1][c] * dchi[1] + G[2][c] * dchi[2]);
        hsv(7, h);
#pragma unroll
        for (int c = 0; c < 6; c++) DDalt[c] = h[c] - (G[0][c] * dal[0] + G[1][c] * dal[1] + G[2][c] * dal[2]);
    }
    const float trDDchi = sym_dot(gu, DDchi);

    // Ricci tensor R_ij = Rt_ij + Rchi_ij
    float R[6];
    {
        float dGam[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++) d1v(14 + i, dGam[i]);
        // E[k][i][m] = gt^ml Gt^k_li, F[k][m][j] = gt^ml Gt_klj
        float E[3][3][3], F[3][3][3];
#pragma unroll
        for (int k = 0; k < 3; k++)
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    E[k][i][m] = gu[SI[m][0]] * G[k][SI[0][i]] + gu[SI[m][1]] * G[k][SI[1][i]] +
                                 gu[SI[m][2]] * G[k][SI[2][i]];
                    F[k][m][i] = gu[SI[m][0]] * Gl[k][SI[0][i]] + gu[SI[m][1]] * Gl[k][SI[1][i]] +
                                 gu[SI[m][2]] * Gl[k][SI[2][i]];
                }
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = i; j < 3; j++) {
                const int c = SI[i][j];
                float h[6];
                hsv(1 + c, h);
                float rr = -.5f * sym_dot(gu, h);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    rr += .5f * (g[SI[k][i]] * dGam[k][j] + g[SI[k][j]] * dGam[k][i]);
                    rr += .5f * Gd[k] * (Gl[i][SI[j][k]] + Gl[j][SI[i][k]]);
#pragma unroll
                    for (int m = 0; m < 3; m++) {
                        rr += E[k][i][m] * Gl[j][SI[k][m]] + E[k][j][m] * Gl[i][SI[k][m]];
                        rr += G[k][SI[i][m]] * F[k][m][j];
                    }
                }
                rr += .5f * ichi * (DDchi[c] + g[c] * trDDchi) - .25f * ichi * ichi * (dchi[i] * dchi[j] + 3.f * g[c] * dchi2);
                R[c] = rr;
            }
    }

    if (WANT_RHS) {
        rhs[0] = A(0, dchi) + (2.f / 3.f) * chi * (alpha * K - divb);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = i; j < 3; j++) {
                const int c = SI[i][j];
                float s = A(1 + c, dg[c]) - (2.f / 3.f) * g[c] * divb - 2.f * alpha * At[c];
#pragma unroll
                for (int k = 0; k < 3; k++) s += g[SI[i][k]] * db[k][j] + g[SI[j][k]] * db[k][i];
                rhs[1 + c] = s;
            }
        const float lap_alpha = chi * sym_dot(gu, DDalt) - .5f * dchi_dal;
        rhs[7] = A(7, dK) - lap_alpha + alpha * (AA + K * K * (1.f / 3.f));
        // At_ij
        float X[6];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = i; j < 3; j++) {
                const int c = SI[i][j];
                const float DDal = DDalt[c] + .5f * ichi * (dchi[i] * dal[j] + dchi[j] * dal[i] - g[c] * dchi_dal);
                X[c] = -DDal + alpha * R[c];
            }
        const float trX = sym_dot(gu, X);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = i; j < 3; j++) {
                const int c = SI[i][j];
                float dA[3];
                d1v(8 + c, dA);
                float s = A(8 + c, dA) - (2.f / 3.f) * At[c] * divb + chi * (X[c] - g[c] * trX * (1.f / 3.f)) +
                          alpha * K * At[c];
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    s += At[SI[i][k]] * db[k][j] + At[SI[j][k]] * db[k][i];
                    s -= 2.f * alpha * At[SI[i][k]] * Atm[k][j];
                }
                rhs[8 + c] = s;
            }
        // Gt^i and B^i
        float ddb[3][6];
#pragma unroll
        for (int i = 0; i < 3; i++) hsv(8 + i, ddb[i]);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float dGi[3], dBi[3];
            d1v(14 + i, dGi);
            float s = A(14 + i, dGi) + (2.f / 3.f) * Gam[i] * divb + sym_dot(gu, ddb[i]);
#pragma unroll
            for (int k = 0; k < 3; k++) s -= Gam[k] * db[i][k];
            s += 2.f * alpha * sym_dot(Atu, G[i]);
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const float graddiv = ddb[0][SI[j][0]] + ddb[1][SI[j][1]] + ddb[2][SI[j][2]];
                s += gu[SI[i][j]] * graddiv * (1.f / 3.f);
                s -= 2.f * Atu[SI[i][j]] * dal[j];
                s -= 2.f * alpha * (1.5f * ichi * Atu[SI[i][j]] * dchi[j] + (2.f / 3.f) * gu[SI[i][j]] * dK[j]);
            }
            rhs[14 + i] = s;
            d1v(21 + i, dBi);
            const float bdb[3] = {dBi[0], dBi[1], dBi[2]};
            if constexpr (UPW)
                rhs[21 + i] = aw[21 + i] + s - (par.advect_shift ? aw[14 + i] : 0.f) - par.eta * B[i];
            else
                rhs[21 + i] = (par.advect_shift ? adv(bdb) - adv(dGi) : 0.f) + s - par.eta * B[i];
            float dbi[3] = {db[i][0], db[i][1], db[i][2]};
            if constexpr (UPW)
                rhs[18 + i] = aw[18 + i] + par.shift_coeff * B[i];
            else
                rhs[18 + i] = (par.advect_shift ? adv(dbi) : 0.f) + par.shift_coeff * B[i];
        }
        rhs[17] = A(17, dal) - (par.slicing == 1 ? alpha * alpha * K : 2.f * alpha * K);
    }
    if (WANT_CONS) {
        *Hc = chi * sym_dot(gu, R) + (2.f / 3.f) * K * K - AA;
        float dA[6][3];
#pragma unroll
        for (int c = 0; c < 6; c++) d1v(8 + c, dA[c]);
#pragma unroll
        for (int i = 0; i < 3; i++) {
            float m = -(2.f / 3.f) * dK[i];
#pragma unroll
            for (int j = 0; j < 3; j++)
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    float D = dA[SI[k][i]][j];
#pragma unroll
                    for (int l = 0; l < 3; l++) D -= G[l][SI[j][k]] * At[SI[l][i]] + G[l][SI[j][i]] * At[SI[k][l]];
                    m += gu[SI[j][k]] * (D - 1.5f * ichi * At[SI[i][j]] * dchi[k]);
                }
            Mc[i] = m;
        }
    }
}

struct TileArgs {
    const uint32_t *positions;
    const int32_t *tiles;
    const uint64_t *boundaries;
    uint32_t substep, max_depth;
    int32_t n_items;  // tiles x 256-cell blocks; workgroups loop over them
};

template <int S, int P, int STAGE, bool PUSH, bool PERSIST, int FMT, int L, bool UPW = false>
__global__ void __launch_bounds__(256) k_bssn_stage(BssnArgsDev a, TileArgs m, BssnParDev par, DepthScales sc)
{
#pragma clang fp contract(fast)
    constexpr int S3 = S * S * S, BPT = (S3 + 255) / 256;
    constexpr size_t TS = tile_stride<S, P, L>();
    __shared__ float xbuf[PUSH && L == 1 && S * S == 256 ? 2 * P * NV * S : 1];
    // persistent launches loop over work items (a separate instantiation:
    // the loop keeps every kernel-argument pointer live across iterations)
    const int item_end = PERSIST ? m.n_items : (int)blockIdx.x + 1;
    for (int item = blockIdx.x; item < item_end; item += PERSIST ? gridDim.x : 1) {
        const int ti = item / BPT, cb = item - ti * BPT;
        const int node = m.tiles[ti];
        const int depth = pos_depth(m.positions[node]);
        if (depth == -1 || !active_depth(m.substep, (uint32_t)depth, m.max_depth)) continue;
        const int cell = cb * 256 + (int)threadIdx.x;
        if (cell >= S3) continue;
        const int z = cell / (S * S), y = (cell / S) % S, x = cell % S;
        // the cell's byte offsets in the field arrays and in the interior-only arrays
        const uint32_t ro = 4u * ((uint32_t)node * (uint32_t)TS + tile_interior<S, P, L>(x, y, z));
        const uint32_t wo = 4u * ((uint32_t)node * (uint32_t)S3 + (uint32_t)cell);
        const size_t w = (size_t)node * S3 + cell;
        // Loads in two batches, each issued whole before it is waited on:
        // the state and the 46 words before the RHS, y_n and acc of all 24
        // fields after it and before any store.  Left to itself the
        // compiler waits on each word where it is decoded, and it cannot
        // move a field's y_n / acc load above the previous field's stores
        // (they might alias), so every field costs a memory round trip.
        // Loading the second batch with the first instead would hold 48
        // more registers through the RHS: one wave per SIMD, and slower.
        float u[NV], yn[NV], ac[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) u[v] = ld_at(a.state[v], ro);
        CellWords<FMT> cw;
        cw.load(a, wo);
        float aw[UPW ? NV : 1];
        if constexpr (UPW) {
#pragma unroll
            for (int v = 0; v < NV; v++) aw[v] = ld_at(a.adv[v], wo);
        }
        // Stages 0 and 3: nothing moves across this point, so all 70 loads
        // of the first batch are issued before the RHS starts.  Left free,
        // the scheduler defers about 30 of them into the RHS, each deferred
        // group costing the wave another memory round trip: depth-3 stage 0
        // 270 -> 238 us, stage 3 319 -> 303 us.  Stages 1 and 2, which load
        // a second batch after the RHS, run 1-3% slower with it (322 -> 325
        // and 334 us), so they are left to the scheduler.
        if constexpr (STAGE == 0 || STAGE == 3) __builtin_amdgcn_sched_barrier(0);
        float k[NV];
        bssn_cell<FMT, true, false, UPW>(a, u, cw, w, par, k, nullptr, nullptr, aw);
        // The offsets again, as values defined here: the compiler otherwise
        // carries their 64-bit extensions from the first batch's block and,
        // seeing no 32-bit offset in this one, forms a 64-bit address per
        // load and store of the second batch (48 extra 64-bit adds per cell)
        // instead of using the scalar base + 32-bit offset form.
        uint32_t ro2 = ro, wo2 = wo;
        asm("" : "+v"(ro2), "+v"(wo2));
#pragma unroll
        for (int v = 0; v < NV; v++) {
            yn[v] = STAGE == 0 ? u[v] : ld_at(a.old_[v], ro2);
            ac[v] = STAGE == 0 ? 0.f : ld_at(a.acc[v], wo2);
        }
        rk4_combine_at<NV, STAGE>(a.acc, a.new_, k, yn, ac, ro2, wo2, sc.dt[depth]);
        if constexpr (PUSH && L == 1 && S * S == 256) {
            // y and z faces are whole slab rows per wave already; the x
            // faces' columns are transposed through LDS
            rk4_push<NV, S, P, L, true>(a.new_, k, m.boundaries, node, x, y, z, a.push_layers, a.push_axes & 6);
            rk4_push_x_rows<NV, S, P>(a.new_, k, m.boundaries, node, x, y, z, a.push_layers, a.push_axes, xbuf);
        } else if (PUSH) {
            rk4_push<NV, S, P, L, true>(a.new_, k, m.boundaries, node, x, y, z, a.push_layers, a.push_axes);
        }
    }
}

// Per-workgroup FP64 partial sums {sum H^2, sum |M|^2, cells, max |H|}.
template <int S, int P, int FMT, int L>
__global__ void __launch_bounds__(256) k_bssn_constraints(BssnArgsDev a, TileArgs m, BssnParDev par, double *partial)
{
    constexpr int S3 = S * S * S, BPT = (S3 + 255) / 256;
    const int ti = blockIdx.x / BPT, cb = blockIdx.x - ti * BPT;
    const int node = m.tiles[ti];
    const int cell = cb * 256 + (int)threadIdx.x;
    double h2 = 0.0, m2 = 0.0, cnt = 0.0, hmax = 0.0;
    if (cell < S3) {
        const int z = cell / (S * S), y = (cell / S) % S, x = cell % S;
        const uint32_t ro = 4u * ((uint32_t)node * tile_stride<S, P, L>() + tile_interior<S, P, L>(x, y, z));
        const uint32_t wo = 4u * ((uint32_t)node * (uint32_t)S3 + (uint32_t)cell);
        const size_t w = (size_t)node * S3 + cell;
        float H, M[3], u[NV];
#pragma unroll
        for (int v = 0; v < NV; v++) u[v] = ld_at(a.state[v], ro);
        CellWords<FMT> cw;
        cw.load(a, wo);
        bssn_cell<FMT, false, true>(a, u, cw, w, par, nullptr, &H, M);
        h2 = (double)H * (double)H;
        m2 = (double)M[0] * M[0] + (double)M[1] * M[1] + (double)M[2] * M[2];
        cnt = 1.0;
        hmax = fabs((double)H);
    }
    // wave64 reduction, then across the four waves through LDS
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        h2 += __shfl_down(h2, off, 64);
        m2 += __shfl_down(m2, off, 64);
        cnt += __shfl_down(cnt, off, 64);
        hmax = fmax(hmax, __shfl_down(hmax, off, 64));
    }
    __shared__ double red[4][4];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
        red[wave][0] = h2; red[wave][1] = m2; red[wave][2] = cnt; red[wave][3] = hmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
        for (int i = 0; i < 4; i++) {
            s0 += red[i][0]; s1 += red[i][1]; s2 += red[i][2]; s3 = fmax(s3, red[i][3]);
        }
        double *o = partial + 4 * (size_t)blockIdx.x;
        o[0] = s0; o[1] = s1; o[2] = s2; o[3] = s3;
    }
}

inline int fill_args(const am_bssn_fields *f, BssnArgsDev &a, bool need_stage)
{
    for (int v = 0; v < NV; v++) {
        a.state[v] = f->state[v];
        a.old_[v] = f->old_[v];
        a.new_[v] = f->new_[v];
        a.acc[v] = f->acc[v];
        a.d1[v] = f->d1[v];
        if (!a.state[v] || !a.d1[v]) return set_error(AM_ERR_INVALID, "am_bssn: null state or gradient word %d", v);
        if (need_stage && (!a.new_[v] || !a.acc[v] || !a.old_[v]))
            return set_error(AM_ERR_INVALID, "am_bssn: null old/new/acc field %d", v);
    }
    for (int h = 0; h < NH; h++) {
        a.d2d[h] = f->d2diag[h];
        a.d2o[h] = f->d2off[h];
        if (!a.d2d[h] || !a.d2o[h]) return set_error(AM_ERR_INVALID, "am_bssn: null Hessian word %d", h);
    }
    return AM_OK;
}

inline BssnParDev dev_par(const am_bssn_params *p)
{
    BssnParDev d;
    d.eta = p->eta;
    d.shift_coeff = p->shift_coeff;
    d.chi_floor = p->chi_floor;
    d.advect_shift = p->advect_shift;
    d.slicing = p->slicing;
    return d;
}

template <int S, int P, bool PUSH, bool PERSIST, int FMT, int L>
static void launch_stage_push(int stage, dim3 grid, hipStream_t st, const BssnArgsDev &a, const TileArgs &m,
                              const BssnParDev &d, const DepthScales &sc)
{
    switch (stage) {
    case 0: hipLaunchKernelGGL((k_bssn_stage<S, P, 0, PUSH, PERSIST, FMT, L>), grid, dim3(256), 0, st, a, m, d, sc); break;
    case 1: hipLaunchKernelGGL((k_bssn_stage<S, P, 1, PUSH, PERSIST, FMT, L>), grid, dim3(256), 0, st, a, m, d, sc); break;
    case 2: hipLaunchKernelGGL((k_bssn_stage<S, P, 2, PUSH, PERSIST, FMT, L>), grid, dim3(256), 0, st, a, m, d, sc); break;
    default: hipLaunchKernelGGL((k_bssn_stage<S, P, 3, PUSH, PERSIST, FMT, L>), grid, dim3(256), 0, st, a, m, d, sc); break;
    }
}
template <int S, int P, bool PUSH, bool PERSIST, int FMT, int L>
static void launch_stage_upw(int stage, dim3 grid, hipStream_t st, const BssnArgsDev &a, const TileArgs &m,
                             const BssnParDev &d, const DepthScales &sc)
{
    switch (stage) {
    case 0: hipLaunchKernelGGL((k_bssn_stage<S, P, 0, PUSH, PERSIST, FMT, L, true>), grid, dim3(256), 0, st, a, m, d, sc); break;
    case 1: hipLaunchKernelGGL((k_bssn_stage<S, P, 1, PUSH, PERSIST, FMT, L, true>), grid, dim3(256), 0, st, a, m, d, sc); break;
    case 2: hipLaunchKernelGGL((k_bssn_stage<S, P, 2, PUSH, PERSIST, FMT, L, true>), grid, dim3(256), 0, st, a, m, d, sc); break;
    default: hipLaunchKernelGGL((k_bssn_stage<S, P, 3, PUSH, PERSIST, FMT, L, true>), grid, dim3(256), 0, st, a, m, d, sc); break;
    }
}
template <int S, int P, bool PUSH, bool PERSIST, int FMT>
static void launch_stage_layout(int layout, int stage, dim3 grid, hipStream_t st, const BssnArgsDev &a,
                                const TileArgs &m, const BssnParDev &d, const DepthScales &sc, bool upw)
{
    if constexpr (FMT == WORD_CF3 && S == 16 && P == 3) {  // the upwinded form: cf3 words, S=16, P=3
        if (upw) {
            if (layout == AM_LAYOUT_INTERIOR_FIRST)
                return launch_stage_upw<S, P, PUSH, PERSIST, FMT, 1>(stage, grid, st, a, m, d, sc);
            return launch_stage_upw<S, P, PUSH, PERSIST, FMT, 0>(stage, grid, st, a, m, d, sc);
        }
    }
    if (layout == AM_LAYOUT_INTERIOR_FIRST)
        return launch_stage_push<S, P, PUSH, PERSIST, FMT, 1>(stage, grid, st, a, m, d, sc);
    launch_stage_push<S, P, PUSH, PERSIST, FMT, 0>(stage, grid, st, a, m, d, sc);
}

// one stage launch with the words in format FMT (argument checks done)
template <int S, int P, int FMT>
int launch_stage_fmt(const am_bssn_fields *f, int stage, const am_bssn_params *p, const am_mesh *mesh,
                        const am_geometry *g, int push, void *stream)
{
    constexpr int BPT = (S * S * S + 255) / 256;
    if (stage < 0 || stage > 3) return set_error(AM_ERR_INVALID, "am_bssn_rhs_stage: stage %d", stage);
    if (push && !mesh->boundaries) return set_error(AM_ERR_INVALID, "am_bssn_rhs_stage: push needs boundaries");
    BssnArgsDev a;
    if (int rc = fill_args(f, a, true)) return rc;
    if (f->push_layers < 0 || f->push_layers > P)
        return set_error(AM_ERR_INVALID, "am_bssn_rhs_stage: push_layers %d (0..%d)", f->push_layers, P);
    a.push_layers = f->push_layers ? f->push_layers : P;
    if (f->push_axes < 0 || f->push_axes > 7)
        return set_error(AM_ERR_INVALID, "%s: push_axes %d (0..7)", "am_bssn_rhs_stage", f->push_axes);
    a.push_axes = f->push_axes ? f->push_axes : 7;
    const int n_items = mesh->n_tiles * BPT;
    TileArgs m{mesh->positions, mesh->tiles, mesh->boundaries, mesh->substep, mesh->max_depth, n_items};
    const DepthScales sc = depth_scales(*g);
    const BssnParDev d = dev_par(p);
    dim3 grid(mesh->grid_limit > 0 && mesh->grid_limit < n_items ? mesh->grid_limit : n_items);
    const bool persist = grid.x < (unsigned)n_items;
    hipStream_t st = (hipStream_t)stream;
    const int lay = f->tile_layout;
    if (lay != AM_LAYOUT_PADDED && lay != AM_LAYOUT_INTERIOR_FIRST)
        return set_error(AM_ERR_INVALID, "am_bssn_rhs_stage: tile layout %d", lay);
    const bool upw = f->adv[0] != nullptr;
    if (upw) {
        if (!(FMT == WORD_CF3 && S == 16 && P == 3))
            return set_error(AM_ERR_INVALID, "am_bssn_rhs_stage: the upwinded form needs S=16, P=3 and cf3 words");
        for (int v = 0; v < NV; v++) {
            if (!f->adv[v]) return set_error(AM_ERR_INVALID, "am_bssn_rhs_stage: null adv field %d", v);
            a.adv[v] = f->adv[v];
        }
    }
    if (push && persist)
        launch_stage_layout<S, P, true, true, FMT>(lay, stage, grid, st, a, m, d, sc, upw);
    else if (push)
        launch_stage_layout<S, P, true, false, FMT>(lay, stage, grid, st, a, m, d, sc, upw);
    else if (persist)
        launch_stage_layout<S, P, false, true, FMT>(lay, stage, grid, st, a, m, d, sc, upw);
    else
        launch_stage_layout<S, P, false, false, FMT>(lay, stage, grid, st, a, m, d, sc, upw);
    return check_launch("k_bssn_stage");
}

template <int S, int P, int FMT>
int launch_cons_fmt(const am_bssn_fields *f, const am_bssn_params *p, const am_mesh *mesh, double *partial,
                       void *stream)
{
    constexpr int BPT = (S * S * S + 255) / 256;
    BssnArgsDev a;
    if (int rc = fill_args(f, a, false)) return rc;
    TileArgs m{mesh->positions, mesh->tiles, mesh->boundaries, mesh->substep, mesh->max_depth,
               mesh->n_tiles * BPT};
    const dim3 grid(mesh->n_tiles * BPT);
    if (f->tile_layout == AM_LAYOUT_INTERIOR_FIRST)
        hipLaunchKernelGGL((k_bssn_constraints<S, P, FMT, 1>), grid, dim3(256), 0, (hipStream_t)stream, a, m,
                           dev_par(p), partial);
    else if (f->tile_layout == AM_LAYOUT_PADDED)
        hipLaunchKernelGGL((k_bssn_constraints<S, P, FMT, 0>), grid, dim3(256), 0, (hipStream_t)stream, a, m,
                           dev_par(p), partial);
    else
        return set_error(AM_ERR_INVALID, "am_bssn_constraints: tile layout %d", f->tile_layout);
    return check_launch("k_bssn_constraints");
}


// the fp32 / fp16 word instantiations (S = 16, P = 3), in am_bssn_words.hip
int bssn_stage_words(int fmt, const am_bssn_fields *f, int stage, const am_bssn_params *p, const am_mesh *mesh,
                     const am_geometry *g, int push, void *stream);
int bssn_cons_words(int fmt, const am_bssn_fields *f, const am_bssn_params *p, const am_mesh *mesh,
                    double *partial, void *stream);

}  // namespace am
