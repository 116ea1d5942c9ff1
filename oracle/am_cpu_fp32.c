/*
 * am_cpu_fp32.c -- FP32 host implementation of the BSSN scheme: the CPU
 * baseline that the GPU run is timed against.
 *
 * TEST INFRASTRUCTURE ONLY.  Loaded by bench.py's cpu_baseline leg and by
 * tests/test_cpu_fp32.py; nothing in the product links or loads it.
 *
 * The same scheme as the GPU path, written for host cores: derivatives of
 * order 2q straight from the fp32 state (the reference's stencils and edge
 * formula, AMR.h:1144-1266, without the compressed words), the BSSN
 * right-hand side of bssn_scheme.inc evaluated in float, and the RK4 stage
 * combination, all in one pass per cell; OpenMP over (tile, z-plane).
 * Equal-depth face ghosts are refreshed between stages by copying the
 * neighbour's boundary layers (the GPU path's push).  Built by the Makefile
 * beside it with -O3 for x86-64-v3 (AVX2 + FMA) and contraction allowed,
 * as a host code would be; its results are checked against the FP64 oracle
 * within fp32 rounding, not bit for bit.
 */
#include "am_oracle_bssn.h"

#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

static const int SYMI[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
static const int HESS_VAR[BSSN_NH] = {0, 1, 2, 3, 4, 5, 6, 17, 18, 19, 20};

#define REAL float
#define RMAX fmaxf
#define SCHEME(n) n##_f
#include "bssn_scheme.inc"

/* centred first / second derivative coefficients of orders 2, 4, 6 */
static const float CD1F[15] = {-.5f, 0.f, .5f, 1.f / 12, -2.f / 3, 0.f, 2.f / 3, -1.f / 12,
                               -1.f / 60, 3.f / 20, -3.f / 4, 0.f, 3.f / 4, -3.f / 20, 1.f / 60};
static const float CD2F[15] = {1.f, -2.f, 1.f, -1.f / 12, 4.f / 3, -5.f / 2, 4.f / 3, -1.f / 12,
                               1.f / 90, -3.f / 20, 3.f / 2, -49.f / 18, 3.f / 2, -3.f / 20, 1.f / 90};

static orc_bssn_par par_of(const double *p)
{
    orc_bssn_par r;
    r.eta = p[0];
    r.shift_coeff = p[1];
    r.chi_floor = p[2];
    r.advect_shift = (int32_t)p[3];
    r.slicing = (int32_t)p[4];
    return r;
}

EXPORT void cpu32_set_threads(int n) { omp_set_num_threads(n); }
EXPORT int cpu32_max_threads(void) { return omp_get_max_threads(); }

/* One RK4 stage on the listed tiles (one depth: inv_dx = 2^depth /
 * outer_dx, dt its step): out = y_n + c dt k and acc per the stage, k the
 * right-hand side at `state` (24 padded fp32 fields, T^3 cells per slot).
 * sigma >= 0 with upwind = 1: upwinded advection plus Kreiss-Oliger
 * dissipation (the shift's own advection only with advect_shift). */
EXPORT void cpu32_bssn_stage(const float *const *state, const float *const *yn, float *const *acc,
                             float *const *out, const int32_t *tiles, int n_tiles, int S, int P, int q, float inv_dx,
                             float dt, int stage, const double *params, float sigma, int upwind)
{
    const orc_bssn_par par = par_of(params);
    const int T = S + 2 * P;
    const int64_t CELLS = (int64_t)T * T * T;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int z = 0; z < S; z++) {
            const int64_t node = tiles[ti];
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    const int64_t r = node * CELLS + ((int64_t)(z + P) * T + (y + P)) * T + (x + P);
                    float u[BSSN_NV], d1[BSSN_NV][3], dd[BSSN_NH][6], k[BSSN_NV], advin[BSSN_NV];
                    int h = 0;
                    for (int v = 0; v < BSSN_NV; v++) {
                        u[v] = state[v][r];
                        const int has_h = h < BSSN_NH && HESS_VAR[h] == v;
                        fd_cell_f(state[v], r, S, T, q, x, y, z, inv_dx, CD1F, CD2F, d1[v], has_h ? dd[h] : NULL);
                        h += has_h;
                    }
                    if (upwind) {
                        const float beta[3] = {u[18], u[19], u[20]};
                        for (int v = 0; v < BSSN_NV; v++)
                            advin[v] = fd_adv_f(state[v], r, T, beta, inv_dx, sigma, v < 18 || par.advect_shift);
                    }
                    bssn_point_f(u, (const float(*)[3])d1, (const float(*)[6])dd, NULL, &par, upwind ? advin : NULL,
                                 k, NULL, NULL);
                    for (int v = 0; v < BSSN_NV; v++) {
                        float a_out, o;
                        rk4_combine_f(stage, dt, yn[v][r], k[v], stage == 0 ? 0.f : acc[v][r], &a_out, &o);
                        acc[v][r] = a_out;
                        out[v][r] = o;
                    }
                }
        }
}

/* Equal-depth face ghosts of the listed tiles from their neighbours'
 * boundary layers: for a face whose neighbour has the same depth, the P
 * ghost layers over the face (S x S each) take the neighbour's P layers
 * next to the shared face.  bnd: the octree's boundary words (low 32 bits:
 * target slot in bits 0-23, depth difference in 24-27, depth in 28-31). */
EXPORT void cpu32_equal_ghosts(float *const *f, int nvars, const uint64_t *bnd, const int32_t *tiles, int n_tiles,
                               int S, int P)
{
    const int T = S + 2 * P;
    const int64_t CELLS = (int64_t)T * T * T;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int dir = 0; dir < 6; dir++) {
            const int node = tiles[ti];
            const uint32_t nd = (uint32_t)bnd[(int64_t)node * 6 + dir];
            if ((nd & (1u << 23)) || (nd >> 28) == 15u || (nd >> 28) == 0u || ((nd >> 24) & 15u)) continue;
            const int64_t tgt = nd & 16777215u;
            const int ax = dir % 3, hi = dir < 3;
            for (int v = 0; v < nvars; v++) {
                float *d = f[v];
                for (int l = 0; l < P; l++) {
                    const int gw = hi ? P + S + l : l;       /* ghost layer along the axis */
                    const int sr = hi ? P + l : S + l;       /* the neighbour's layer it copies */
                    for (int b = P; b < P + S; b++)
                        for (int a = P; a < P + S; a++) {
                            int w[3], s[3];
                            const int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
                            w[ax] = gw; s[ax] = sr;
                            w[o1] = a; s[o1] = a;
                            w[o2] = b; s[o2] = b;
                            d[node * CELLS + ((int64_t)w[2] * T + w[1]) * T + w[0]] =
                                d[tgt * CELLS + ((int64_t)s[2] * T + s[1]) * T + s[0]];
                        }
                }
            }
        }
}

/* One full RK4 step on the listed tiles: the stages read yn, a, b in turn
 * (a, b: scratch state buffers) with the equal-depth ghosts of each stage
 * output refreshed before the next stage reads it; the result is in b. */
EXPORT void cpu32_rk4_step(const float *const *yn, float *const *a, float *const *b, float *const *acc,
                           const uint64_t *bnd, const int32_t *tiles, int n_tiles, int S, int P, int q, float inv_dx,
                           float dt, const double *params, float sigma, int upwind)
{
    const float *const *in[4] = {yn, (const float *const *)a, (const float *const *)b, (const float *const *)a};
    float *const *outs[4] = {a, b, a, b};
    for (int s = 0; s < 4; s++) {
        cpu32_bssn_stage(in[s], yn, acc, outs[s], tiles, n_tiles, S, P, q, inv_dx, dt, s, params, sigma, upwind);
        cpu32_equal_ghosts(outs[s], BSSN_NV, bnd, tiles, n_tiles, S, P);
    }
}

/* Derivatives only, for comparison with the reference's kernel bodies:
 * the gradient of every field (d1[v], 3 floats per interior cell) and the
 * six second derivatives of the 11 Hessian fields (d2[h], 6 per cell). */
EXPORT void cpu32_derivatives(const float *const *state, float *const *d1, float *const *d2, const int32_t *tiles,
                              int n_tiles, int S, int P, int q, float inv_dx)
{
    const int T = S + 2 * P;
    const int64_t CELLS = (int64_t)T * T * T, IC = (int64_t)S * S * S;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int z = 0; z < S; z++) {
            const int64_t node = tiles[ti];
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    const int64_t r = node * CELLS + ((int64_t)(z + P) * T + (y + P)) * T + (x + P);
                    const int64_t w = (int64_t)ti * IC + ((int64_t)z * S + y) * S + x;
                    int h = 0;
                    for (int v = 0; v < BSSN_NV; v++) {
                        const int has_h = h < BSSN_NH && HESS_VAR[h] == v;
                        fd_cell_f(state[v], r, S, T, q, x, y, z, inv_dx, CD1F, CD2F, d1[v] + 3 * w,
                                  has_h ? d2[h] + 6 * w : NULL);
                        h += has_h;
                    }
                }
        }
}
