/*
 * am_oracle_bssn.c -- FP64 CPU oracle of the BSSN scheme.
 *
 * TEST INFRASTRUCTURE ONLY (see am_oracle.c).  The reference computes only
 * the metric derivatives, gamma^{ij} d_i d_j gamma_kl and the Christoffel
 * symbols (AdaptiveMesh/BSSN.h:7-102); the right-hand side, RK4 update and
 * constraints have no reference and are restated here from the textbook
 * moving-puncture BSSN system (Baumgarte & Shapiro 1998; Campanelli et al.
 * 2006; Alcubierre, "Introduction to 3+1 Numerical Relativity", ch. 2.8):
 *
 *   d_t chi   = b.d chi + 2/3 chi (alpha K - d_k b^k)
 *   d_t gt_ij = b.d gt_ij + gt_ik d_j b^k + gt_jk d_i b^k - 2/3 gt_ij d_k b^k - 2 alpha At_ij
 *   d_t K     = b.d K - D^i D_i alpha + alpha (At_ij At^ij + K^2/3)
 *   d_t At_ij = b.d At_ij + At_ik d_j b^k + At_jk d_i b^k - 2/3 At_ij d_k b^k
 *               + chi [-D_i D_j alpha + alpha R_ij]^TF + alpha (K At_ij - 2 At_ik At^k_j)
 *   d_t Gt^i  = b.d Gt^i - Gt^k d_k b^i + 2/3 Gt^i d_k b^k + gt^jk d_j d_k b^i
 *               + 1/3 gt^ij d_j d_k b^k - 2 At^ij d_j alpha
 *               + 2 alpha (Gt^i_jk At^jk - 3/2 At^ij d_j chi / chi - 2/3 gt^ij d_j K)
 *   d_t alpha = b.d alpha - 2 alpha K            (1+log; harmonic: -alpha^2 K)
 *   d_t b^i   = [b.d b^i] + 3/4 B^i              (Gamma-driver)
 *   d_t B^i   = [b.d B^i] + d_t Gt^i - [b.d Gt^i] - eta B^i
 *
 * with R_ij = Rt_ij + Rchi_ij,
 *   Rt_ij   = -1/2 gt^lm d_l d_m gt_ij + gt_k(i d_j) Gt^k + Gd^k Gt_(ij)k
 *             + gt^lm (Gt^k_li Gt_jkm + Gt^k_lj Gt_ikm + Gt^k_im Gt_klj)
 *   Rchi_ij = (Dt_i Dt_j chi + gt_ij Dt^l Dt_l chi) / (2 chi)
 *             - (d_i chi d_j chi + 3 gt_ij d^l chi d_l chi) / (4 chi^2)
 *   D_i D_j alpha = Dt_i Dt_j alpha + (d_i chi d_j alpha + d_j chi d_i alpha
 *                   - gt_ij gt^kl d_k chi d_l alpha) / (2 chi)
 * where Gt^i_jk enters as given (the scheme stores it as compressed words),
 * Gt_ijk = (d_k gt_ij + d_j gt_ik - d_i gt_jk)/2 comes from the metric
 * derivatives, Gd^k = gt^ij Gt^k_ij, and the evolved Gt^i is used in the Lie
 * derivative.  Constraints:
 *   H   = chi gt^ij R_ij + 2/3 K^2 - At_ij At^ij
 *   M_i = gt^jk (d_j At_ki - Gt^l_jk At_li - Gt^l_ji At_kl) - 2/3 d_i K
 *         - 3/(2 chi) gt^jk At_ij d_k chi
 *
 * Pins (tests/test_oracle_bssn.py): a stationary Kerr-Schild black hole has
 * d_t = 0 for chi, gt, K, At, Gt and H = M = 0; the harmonic gauge wave
 * (Alcubierre et al. 2004) has the analytic time derivatives; flat space
 * stays flat.
 *
 * The rest of the file is the same scheme in FP64 on tiles: FD with the
 * reference's stencils and edge formulas (AMR.h:1144-1266) without
 * quantisation, ghost fill (AMR.h:670-762 index maps, plus the tricubic
 * extension) and restriction (AMR.h:961-996) in double, the RK4 stage update,
 * and constraint sums -- the "host FP64 run of the same scheme".
 */
#include "am_oracle_bssn.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define EXPORT __attribute__((visibility("default")))

static const int SYMI[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
static const int HESS_VAR[BSSN_NH] = {0, 1, 2, 3, 4, 5, 6, 17, 18, 19, 20};

/* the right-hand side, FD cell stencils and RK4 combination in double */
#define REAL double
#define RMAX fmax
#define SCHEME(n) n
#include "bssn_scheme.inc"

static orc_bssn_par par_from(const double *p)
{
    orc_bssn_par r;
    r.eta = p[0];
    r.shift_coeff = p[1];
    r.chi_floor = p[2];
    r.advect_shift = (int32_t)p[3];
    r.slicing = (int32_t)p[4];
    return r;
}

/* Pointwise entry for the tests.  params = {eta, shift_coeff, chi_floor,
 * advect_shift, slicing}; G may be NULL (computed from the metric). */
EXPORT void orc_bssn_point(const double *u, const double *d1, const double *dd, const double *G, const double *params,
                           double *rhs, double *HM)
{
    const orc_bssn_par p = par_from(params);
    bssn_point(u, (const double(*)[3])d1, (const double(*)[6])dd, (const double(*)[3])G, &p, NULL, rhs, HM,
               HM ? HM + 1 : NULL);
}
/* The same with the advection of every field given (advin[24]: upwinded
 * advection plus dissipation, as the FD's adv output). */
EXPORT void orc_bssn_point_adv(const double *u, const double *d1, const double *dd, const double *params,
                               const double *advin, double *rhs)
{
    const orc_bssn_par p = par_from(params);
    bssn_point(u, (const double(*)[3])d1, (const double(*)[6])dd, NULL, &p, advin, rhs, NULL, NULL);
}

/* --------------------------------------------------------- tile helpers */

static inline int pos_depth(uint32_t d) { return (~d) ? (int)(d >> 27) : -1; }
static inline int active_depth(uint32_t sub, uint32_t depth, uint32_t maxd)
{
    return (sub & ((1u << (maxd - depth)) - 1u)) == 0u;
}
static inline void cf3_dec(uint32_t w, double *o)
{
    const double base = ldexp(1.0, (int)(w >> 27) - 22);
    o[0] = base * (double)(w & 255u) * ((w >> 8) & 1u ? -1.0 : 1.0);
    o[1] = base * (double)((w >> 9) & 255u) * ((w >> 17) & 1u ? -1.0 : 1.0);
    o[2] = base * (double)((w >> 18) & 255u) * ((w >> 26) & 1u ? -1.0 : 1.0);
}

/* IEEE 754 binary16 bits to double (exact: normals, subnormals, zeros;
 * the words never hold infinities or NaNs) */
static inline double half_to_double(uint16_t h)
{
    const int e = (h >> 10) & 31, m = h & 1023;
    const double v = e == 0 ? ldexp((double)m, -24) : ldexp((double)(1024 + m), e - 25);
    return (h & 0x8000u) ? -v : v;
}

/* One derivative triple of cell w from a word array in format fmt: 0 the
 * dithered compressed_float3 words, 1 three fp32, 2 three fp16 per cell
 * (the precision study's formats). */
static inline void word_dec(const void *base, int64_t w, int fmt, double *o)
{
    if (fmt == 0) {
        cf3_dec(((const uint32_t *)base)[w], o);
    } else if (fmt == 1) {
        const float *p = (const float *)base + 3 * w;
        o[0] = p[0]; o[1] = p[1]; o[2] = p[2];
    } else {
        const uint16_t *p = (const uint16_t *)base + 3 * w;
        o[0] = half_to_double(p[0]); o[1] = half_to_double(p[1]); o[2] = half_to_double(p[2]);
    }
}

/* The RK4 stage on the product's inputs: fp32 state (stage input), fp32
 * y_n and dithered words; FP64 arithmetic; acc and new are FP64 arrays in
 * the padded layout, one pointer per variable (interior cells written).
 * dtd[d] is the step of depth d.  chw == NULL: the raised Christoffel
 * symbols are computed from the metric and its gradient words (as the GPU
 * RHS does) instead of read from Christoffel words. */
static void stage_from_words(const float *const *state, const float *const *old_, const double *const *acc_in,
                             double *const *acc_out, double *const *new_out, const void *const *d1w,
                             const void *const *d2dw, const void *const *d2ow, const uint32_t *const *chw,
                             const float *const *advw, const uint32_t *pos, const int32_t *tiles, int n_tiles,
                             uint32_t substep, uint32_t maxd, int stage, const double *dtd, const double *params,
                             int S, int P, int fmt)
{
    const orc_bssn_par par = par_from(params);
    const int T = S + 2 * P;
    const int64_t CELLS = (int64_t)T * T * T, ICELLS = (int64_t)S * S * S;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int z = 0; z < S; z++) {
            const int node = tiles[ti];
            const int depth = pos_depth(pos[node]);
            if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) continue;
            const double dt = dtd[depth];
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    const int64_t r = (int64_t)node * CELLS + ((int64_t)(z + P) * T + (y + P)) * T + (x + P);
                    const int64_t w = (int64_t)node * ICELLS + ((int64_t)z * S + y) * S + x;
                    double u[BSSN_NV], d1[BSSN_NV][3], dd[BSSN_NH][6], G[6][3], k[BSSN_NV];
                    for (int v = 0; v < BSSN_NV; v++) {
                        u[v] = state[v][r];
                        word_dec(d1w[v], w, fmt, d1[v]);
                    }
                    for (int h = 0; h < BSSN_NH; h++) {
                        double a[3], b[3];
                        word_dec(d2dw[h], w, fmt, a);
                        word_dec(d2ow[h], w, fmt, b);
                        dd[h][0] = a[0]; dd[h][1] = b[0]; dd[h][2] = b[1];
                        dd[h][3] = a[1]; dd[h][4] = b[2]; dd[h][5] = a[2];
                    }
                    if (chw)
                        for (int c = 0; c < 6; c++) cf3_dec(chw[c][w], G[c]);
                    double advin[BSSN_NV];
                    if (advw)
                        for (int v = 0; v < BSSN_NV; v++) advin[v] = advw[v][w];
                    bssn_point(u, (const double(*)[3])d1, (const double(*)[6])dd,
                               chw ? (const double(*)[3])G : NULL, &par, advw ? advin : NULL, k, NULL, NULL);
                    for (int v = 0; v < BSSN_NV; v++) {
                        const double ai = stage == 0 ? 0.0 : acc_in[v][r];
                        rk4_combine(stage, dt, (double)old_[v][r], k[v], ai, &acc_out[v][r], &new_out[v][r]);
                    }
                }
        }
}

EXPORT void orc_bssn_stage_from_words(const float *const *state, const float *const *old_, const double *const *acc_in,
                                      double *const *acc_out, double *const *new_out, const uint32_t *const *d1w,
                                      const uint32_t *const *d2dw, const uint32_t *const *d2ow,
                                      const uint32_t *const *chw, const uint32_t *pos, const int32_t *tiles,
                                      int n_tiles, uint32_t substep, uint32_t maxd, int stage, const double *dtd,
                                      const double *params, int S, int P)
{
    stage_from_words(state, old_, acc_in, acc_out, new_out, (const void *const *)d1w, (const void *const *)d2dw,
                     (const void *const *)d2ow, chw, NULL, pos, tiles, n_tiles, substep, maxd, stage, dtd, params, S,
                     P, 0);
}
/* the same with the FD's advection + dissipation output of every field
 * (advw: 24 fp32 arrays, interior-only layout) in place of the centred
 * advection terms */
EXPORT void orc_bssn_stage_from_words_adv(const float *const *state, const float *const *old_,
                                          const double *const *acc_in, double *const *acc_out,
                                          double *const *new_out, const uint32_t *const *d1w,
                                          const uint32_t *const *d2dw, const uint32_t *const *d2ow,
                                          const float *const *advw, const uint32_t *pos, const int32_t *tiles,
                                          int n_tiles, uint32_t substep, uint32_t maxd, int stage, const double *dtd,
                                          const double *params, int S, int P)
{
    stage_from_words(state, old_, acc_in, acc_out, new_out, (const void *const *)d1w, (const void *const *)d2dw,
                     (const void *const *)d2ow, NULL, advw, pos, tiles, n_tiles, substep, maxd, stage, dtd, params,
                     S, P, 0);
}
/* the same with the words in format fmt (0 cf3, 1 fp32, 2 fp16 triples) */
EXPORT void orc_bssn_stage_from_words_fmt(const float *const *state, const float *const *old_,
                                          const double *const *acc_in, double *const *acc_out,
                                          double *const *new_out, const void *const *d1w, const void *const *d2dw,
                                          const void *const *d2ow, const uint32_t *pos, const int32_t *tiles,
                                          int n_tiles, uint32_t substep, uint32_t maxd, int stage, const double *dtd,
                                          const double *params, int S, int P, int fmt)
{
    stage_from_words(state, old_, acc_in, acc_out, new_out, d1w, d2dw, d2ow, NULL, NULL, pos, tiles, n_tiles,
                     substep, maxd, stage, dtd, params, S, P, fmt);
}

/* Constraint sums on the product's inputs: per listed tile,
 * out[4*ti + {0,1,2,3}] = {sum H^2, sum M_i M_i, cells, max |H|}. */
static void constraints_from_words(const float *const *state, const void *const *d1w, const void *const *d2dw,
                                   const void *const *d2ow, const uint32_t *const *chw, const int32_t *tiles,
                                   int n_tiles, const double *params, int S, int P, double *out, int fmt)
{
    const orc_bssn_par par = par_from(params);
    const int T = S + 2 * P;
    const int64_t CELLS = (int64_t)T * T * T, ICELLS = (int64_t)S * S * S;
#pragma omp parallel for schedule(static)
    for (int ti = 0; ti < n_tiles; ti++) {
        const int node = tiles[ti];
        double h2 = 0, m2 = 0, hmax = 0;
        for (int z = 0; z < S; z++)
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    const int64_t r = (int64_t)node * CELLS + ((int64_t)(z + P) * T + (y + P)) * T + (x + P);
                    const int64_t w = (int64_t)node * ICELLS + ((int64_t)z * S + y) * S + x;
                    double u[BSSN_NV], d1[BSSN_NV][3], dd[BSSN_NH][6], G[6][3], k[BSSN_NV], H, M[3];
                    for (int v = 0; v < BSSN_NV; v++) {
                        u[v] = state[v][r];
                        word_dec(d1w[v], w, fmt, d1[v]);
                    }
                    for (int h = 0; h < BSSN_NH; h++) {
                        double a[3], b[3];
                        word_dec(d2dw[h], w, fmt, a);
                        word_dec(d2ow[h], w, fmt, b);
                        dd[h][0] = a[0]; dd[h][1] = b[0]; dd[h][2] = b[1];
                        dd[h][3] = a[1]; dd[h][4] = b[2]; dd[h][5] = a[2];
                    }
                    if (chw)
                        for (int c = 0; c < 6; c++) cf3_dec(chw[c][w], G[c]);
                    bssn_point(u, (const double(*)[3])d1, (const double(*)[6])dd,
                               chw ? (const double(*)[3])G : NULL, &par, NULL, k, &H, M);
                    h2 += H * H;
                    m2 += M[0] * M[0] + M[1] * M[1] + M[2] * M[2];
                    hmax = fmax(hmax, fabs(H));
                }
        out[4 * ti] = h2;
        out[4 * ti + 1] = m2;
        out[4 * ti + 2] = (double)S * S * S;
        out[4 * ti + 3] = hmax;
    }
}

EXPORT void orc_bssn_constraints_from_words(const float *const *state, const uint32_t *const *d1w,
                                            const uint32_t *const *d2dw, const uint32_t *const *d2ow,
                                            const uint32_t *const *chw, const int32_t *tiles, int n_tiles,
                                            const double *params, int S, int P, double *out)
{
    constraints_from_words(state, (const void *const *)d1w, (const void *const *)d2dw, (const void *const *)d2ow,
                           chw, tiles, n_tiles, params, S, P, out, 0);
}
EXPORT void orc_bssn_constraints_from_words_fmt(const float *const *state, const void *const *d1w,
                                                const void *const *d2dw, const void *const *d2ow,
                                                const int32_t *tiles, int n_tiles, const double *params, int S, int P,
                                                double *out, int fmt)
{
    constraints_from_words(state, d1w, d2dw, d2ow, NULL, tiles, n_tiles, params, S, P, out, fmt);
}

/* ===================================================== FP64 scheme on tiles */

static const double CD1_EXACT[15] = {-.5, 0., .5, 1. / 12, -2. / 3, 0., 2. / 3, -1. / 12,
                                     -1. / 60, 3. / 20, -3. / 4, 0., 3. / 4, -3. / 20, 1. / 60};
static const double CD2_EXACT[15] = {1., -2., 1., -1. / 12, 4. / 3, -5. / 2, 4. / 3, -1. / 12,
                                     1. / 90, -3. / 20, 3. / 2, -49. / 18, 3. / 2, -3. / 20, 1. / 90};
/* The coefficient tables the FP64 scheme uses: the exact fractions
 * (default) or, after orc_fd64_float_coefficients(1), the same fractions
 * rounded to float (the reference's tables, AMR.h:1144-1150) -- to separate
 * the effect of those tables from that of fp32 arithmetic and storage. */
static double CD1[15], CD2[15];
static int cd_ready = 0;
EXPORT void orc_fd64_float_coefficients(int on)
{
    for (int i = 0; i < 15; i++) {
        CD1[i] = on ? (double)(float)CD1_EXACT[i] : CD1_EXACT[i];
        CD2[i] = on ? (double)(float)CD2_EXACT[i] : CD2_EXACT[i];
    }
    cd_ready = 1;
}

/* Upwinded advection and Kreiss-Oliger dissipation of the FP64 scheme
 * (fd_adv): off until orc_bssn64_set_dissipation(sigma, 1); the dissipation
 * comes only with the upwinded advection. */
static double g_ko_sigma = 0.0;
static int g_upwind = 0;
EXPORT void orc_bssn64_set_dissipation(double sigma, int upwind)
{
    g_ko_sigma = sigma;
    g_upwind = upwind;
}

/* FP64 RK4 stage with exact (unquantised) derivatives and Christoffels.
 * state/old/acc/new: 24 pointers to double arrays over the padded layout.
 * inv_dx[d] and dtd[d] per depth. */
EXPORT void orc_bssn64_stage(const double *const *state, const double *const *old_, double *const *acc,
                             double *const *new_, const uint32_t *pos, const int32_t *tiles, int n_tiles,
                             uint32_t substep, uint32_t maxd, int stage, const double *inv_dx, const double *dtd,
                             const double *params, int S, int P, int q)
{
    if (!cd_ready) orc_fd64_float_coefficients(0);
    const orc_bssn_par par = par_from(params);
    const int T = S + 2 * P;
    const int64_t CELLS = (int64_t)T * T * T;
#pragma omp parallel for collapse(2) schedule(static)
    for (int ti = 0; ti < n_tiles; ti++)
        for (int z = 0; z < S; z++) {
            const int node = tiles[ti];
            const int depth = pos_depth(pos[node]);
            if (depth == -1 || !active_depth(substep, (uint32_t)depth, maxd)) continue;
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    const int64_t r = (int64_t)node * CELLS + ((int64_t)(z + P) * T + (y + P)) * T + (x + P);
                    double u[BSSN_NV], d1[BSSN_NV][3], dd[BSSN_NH][6], k[BSSN_NV];
                    int h = 0;
                    for (int v = 0; v < BSSN_NV; v++) {
                        u[v] = state[v][r];
                        const int has_h = (h < BSSN_NH && HESS_VAR[h] == v);
                        fd_cell(state[v], r, S, T, q, x, y, z, inv_dx[depth], CD1, CD2, d1[v], has_h ? dd[h] : NULL);
                        if (has_h) h++;
                    }
                    double advin[BSSN_NV];
                    if (g_upwind) {
                        const double beta[3] = {u[18], u[19], u[20]};
                        for (int v = 0; v < BSSN_NV; v++)
                            advin[v] = fd_adv(state[v], r, T, beta, inv_dx[depth], g_ko_sigma,
                                                v < 18 || par.advect_shift);
                    }
                    bssn_point(u, (const double(*)[3])d1, (const double(*)[6])dd, NULL, &par, g_upwind ? advin : NULL,
                               k, NULL, NULL);
                    for (int v = 0; v < BSSN_NV; v++) {
                        double a_out, o;
                        rk4_combine(stage, dtd[depth], old_[v][r], k[v], stage == 0 ? 0.0 : acc[v][r], &a_out, &o);
                        acc[v][r] = a_out;
                        new_[v][r] = o;
                    }
                }
        }
}

/* FP64 constraint sums with exact derivatives (layout as above). */
EXPORT void orc_bssn64_constraints(const double *const *state, const uint32_t *pos, const int32_t *tiles,
                                   int n_tiles, const double *inv_dx, const double *params, int S, int P, int q,
                                   double *out)
{
    if (!cd_ready) orc_fd64_float_coefficients(0);
    const orc_bssn_par par = par_from(params);
    const int T = S + 2 * P;
    const int64_t CELLS = (int64_t)T * T * T;
#pragma omp parallel for schedule(static)
    for (int ti = 0; ti < n_tiles; ti++) {
        const int node = tiles[ti];
        const int depth = pos_depth(pos[node]);
        double h2 = 0, m2 = 0, hmax = 0;
        for (int z = 0; z < S; z++)
            for (int y = 0; y < S; y++)
                for (int x = 0; x < S; x++) {
                    const int64_t r = (int64_t)node * CELLS + ((int64_t)(z + P) * T + (y + P)) * T + (x + P);
                    double u[BSSN_NV], d1[BSSN_NV][3], dd[BSSN_NH][6], k[BSSN_NV], H, M[3];
                    int h = 0;
                    for (int v = 0; v < BSSN_NV; v++) {
                        u[v] = state[v][r];
                        const int has_h = (h < BSSN_NH && HESS_VAR[h] == v);
                        fd_cell(state[v], r, S, T, q, x, y, z, inv_dx[depth], CD1, CD2, d1[v], has_h ? dd[h] : NULL);
                        if (has_h) h++;
                    }
                    bssn_point(u, (const double(*)[3])d1, (const double(*)[6])dd, NULL, &par, NULL, k, &H, M);
                    h2 += H * H;
                    m2 += M[0] * M[0] + M[1] * M[1] + M[2] * M[2];
                    hmax = fmax(hmax, fabs(H));
                }
        out[4 * ti] = h2;
        out[4 * ti + 1] = m2;
        out[4 * ti + 2] = (double)S * S * S;
        out[4 * ti + 3] = hmax;
    }
}

/* ---------------------------------------------- FP64 ghost fill / copy back */

static inline double lag_q(const double *d, const double *rp, int64_t n, int T)
{
    const int64_t T2 = (int64_t)T * T;
    double f = d[n], fx = d[n + 1], fy = d[n + T], fz = d[n + T2];
    double fxx = d[n - 1], fyy = d[n - T], fzz = d[n - T2];
    double fxy = d[n + T + 1] + f - fx - fy, fxz = d[n + T2 + 1] + f - fx - fz, fyz = d[n + T2 + T] + f - fy - fz;
    fx = (fx - fxx) * .5; fy = (fy - fyy) * .5; fz = (fz - fzz) * .5;
    fxx = fx + fxx - f; fyy = fy + fyy - f; fzz = fz + fzz - f;
    return f + (fx + fxx * rp[0]) * rp[0] + (fy + fyy * rp[1]) * rp[1] + (fz + fzz * rp[2]) * rp[2] +
           fxy * rp[0] * rp[1] + fxz * rp[0] * rp[2] + fyz * rp[1] * rp[2];
}
static inline void lag4d(double t, double *w)
{
    w[0] = -t * (t - 1) * (t - 2) / 6;
    w[1] = (t + 1) * (t - 1) * (t - 2) / 2;
    w[2] = -(t + 1) * t * (t - 2) / 2;
    w[3] = (t + 1) * t * (t - 1) / 6;
}
static inline double tric(const double *d, const double *rp, int64_t n, int T)
{
    const int64_t T2 = (int64_t)T * T;
    double wx[4], wy[4], wz[4], acc = 0;
    lag4d(rp[0], wx); lag4d(rp[1], wy); lag4d(rp[2], wz);
    for (int k = 0; k < 4; k++)
        for (int j = 0; j < 4; j++)
            for (int i = 0; i < 4; i++) acc += d[n + (k - 1) * T2 + (j - 1) * T + (i - 1)] * wx[i] * wy[j] * wz[k];
    return acc;
}

/* Ghost fill of 24 double fields (the product's am_ghost_fill_stage): the
 * read index maps are computed in float exactly as the product does, the
 * interpolation in double.  mode 0 quadratic, 2 tricubic; older != NULL
 * with mode 2 gives quadratic time.  Equal-depth faces copy `new`; coarser
 * faces interpolate at (k + frac) / 2^dd. */
EXPORT void orc_ghost64(double *const *old_, double *const *new_, const double *const *older, int nvars,
                        const uint64_t *bnd, const int32_t *tiles, int n_tiles, uint32_t substep, uint32_t maxd,
                        float frac, int mode, int S, int P, int equal_copy)
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
            int tgt = (int)((nd & 16777215u) | ((nd & (1u << 23)) ? 4278190080u : 0u));
            const int cd = ((nd >> 28) == 15u) ? -1 : (int)(nd >> 28);
            if (tgt == -1 || cd <= 0 || !active_depth(substep, (uint32_t)cd, maxd)) continue;
            double relp[3];
            cf3_dec((uint32_t)(bb >> 32), relp);
            const uint32_t dd = (nd >> 24) & 15u;
            const float rel_scale = 1.f / (float)(1u << dd);
            const float tmp = (float)((substep >> (maxd - (uint32_t)cd)) & ((1u << dd) - 1u));
            const double ty = dd == 0 ? 1.0 : ((double)tmp + frac) / (double)(1u << dd);
            const uint32_t uP = (uint32_t)P, uS = (uint32_t)S;
            for (uint32_t l = 0; l < uP; l++)
                for (uint32_t b = 0; b < uS; b++)
                    for (uint32_t a = 0; a < uS; a++) {
                        uint32_t w[3];
                        switch (dir) {
                        case 0: w[0] = l + uP + uS; w[1] = a + uP; w[2] = b + uP; break;
                        case 1: w[0] = a + uP; w[1] = l + uP + uS; w[2] = b + uP; break;
                        case 2: w[0] = a + uP; w[1] = b + uP; w[2] = l + uP + uS; break;
                        case 3: w[0] = l; w[1] = a + uP; w[2] = b + uP; break;
                        case 4: w[0] = a + uP; w[1] = l; w[2] = b + uP; break;
                        default: w[0] = a + uP; w[1] = b + uP; w[2] = l; break;
                        }
                        float rp[3];
                        uint32_t ri[3];
                        for (int k = 0; k < 3; k++) {
                            rp[k] = ((float)w[k] - (float)uP + .5f) * invS - .5f;
                            rp[k] = rp[k] * rel_scale + (float)relp[k];
                            rp[k] = (rp[k] + .5f) * (float)uS + (float)uP - 0.5f;
                            uint32_t v;
                            if (mode == 2)
                                v = tgt == 0 ? (uint32_t)fmaxf(rp[k], 1.f) : (uint32_t)fmaxf(rp[k], (float)(uP + 1u));
                            else
                                v = tgt == 0 ? (uint32_t)(fmaxf(rp[k], 0.f) + .5f) : (uint32_t)(fmaxf(rp[k], (float)uP) + .5f);
                            const uint32_t hi = mode == 2 ? (tgt == 0 ? (uint32_t)T - 3u : (uint32_t)T - uP - 3u)
                                                          : (tgt == 0 ? (uint32_t)T - 2u : (uint32_t)T - uP - 2u);
                            ri[k] = v < hi ? v : hi;
                        }
                        const int64_t wn0 = (int64_t)node * cells + w[0] + ((int64_t)w[1] + (int64_t)w[2] * T) * T;
                        if (equal_copy && dd == 0) {
                            const int64_t src = (int64_t)tgt * cells + (int64_t)rp[0] +
                                                ((int64_t)rp[1] + (int64_t)rp[2] * T) * T;
                            for (int v = 0; v < nvars; v++) new_[v][wn0] = new_[v][src];
                            continue;
                        }
                        const int64_t rn = (int64_t)tgt * cells + ri[0] + ((int64_t)ri[1] + (int64_t)ri[2] * T) * T;
                        double fr[3];
                        for (int k = 0; k < 3; k++) fr[k] = (double)(rp[k] - (float)ri[k]);
                        const int64_t wn = (int64_t)node * cells + w[0] + ((int64_t)w[1] + (int64_t)w[2] * T) * T;
                        for (int v = 0; v < nvars; v++) {
                            double o, nw;
                            if (mode == 2) {
                                o = tric(old_[v], fr, rn, T);
                                nw = tric(new_[v], fr, rn, T);
                            } else {
                                o = lag_q(old_[v], fr, rn, T);
                                nw = lag_q(new_[v], fr, rn, T);
                            }
                            if (mode == 2 && older && dd > 0) {
                                const double od = tric(older[v], fr, rn, T), s = ty;
                                new_[v][wn] = od * .5 * s * (s - 1) + o * (1 - s) * (1 + s) + nw * .5 * s * (s + 1);
                            } else {
                                new_[v][wn] = o * (1 - ty) + nw * ty;
                            }
                        }
                    }
        }
}

/* copy-back with restriction on double fields (AMR.h:961-996); with
 * older != NULL the previous old is kept there first (three time levels). */
EXPORT void orc_copy_back64(double *const *old_, const double *const *new_, double *const *older, int nvars,
                            const uint32_t *pos, const int32_t *children, const int32_t *tiles, int n_tiles,
                            uint32_t substep, uint32_t maxd, int S, int P)
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
                    int64_t c = -1;
                    if (child != -1) {
                        const int64_t cx = (((uint32_t)x - (uint32_t)P) * 2u) % (uint32_t)S + (uint32_t)P;
                        const int64_t cy = (((uint32_t)y - (uint32_t)P) * 2u) % (uint32_t)S + (uint32_t)P;
                        const int64_t cz = (((uint32_t)z - (uint32_t)P) * 2u) % (uint32_t)S + (uint32_t)P;
                        c = (int64_t)child * cells + (cz * T + cy) * T + cx;
                    }
                    for (int v = 0; v < nvars; v++) {
                        if (older) older[v][o] = old_[v][o];
                        const double *n = new_[v];
                        old_[v][o] = c < 0 ? n[o]
                                           : (n[c] + n[c + 1] + n[c + T] + n[c + T + 1] + n[c + T2] + n[c + T2 + 1] +
                                              n[c + T2 + T] + n[c + T2 + T + 1]) * .125;
                    }
                }
        }
}
