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

static void inv3(const double g[3][3], double r[3][3])
{
    const double det = g[0][0] * (g[1][1] * g[2][2] - g[1][2] * g[2][1]) -
                       g[0][1] * (g[1][0] * g[2][2] - g[1][2] * g[2][0]) +
                       g[0][2] * (g[1][0] * g[2][1] - g[1][1] * g[2][0]);
    r[0][0] = (g[1][1] * g[2][2] - g[1][2] * g[2][1]) / det;
    r[0][1] = (g[0][2] * g[2][1] - g[0][1] * g[2][2]) / det;
    r[0][2] = (g[0][1] * g[1][2] - g[0][2] * g[1][1]) / det;
    r[1][0] = (g[1][2] * g[2][0] - g[1][0] * g[2][2]) / det;
    r[1][1] = (g[0][0] * g[2][2] - g[0][2] * g[2][0]) / det;
    r[1][2] = (g[0][2] * g[1][0] - g[0][0] * g[1][2]) / det;
    r[2][0] = (g[1][0] * g[2][1] - g[1][1] * g[2][0]) / det;
    r[2][1] = (g[0][1] * g[2][0] - g[0][0] * g[2][1]) / det;
    r[2][2] = (g[0][0] * g[1][1] - g[0][1] * g[1][0]) / det;
}

/* u[24]; d1[24][3] (d1[v][k] = d_k u_v); dd[11][6] symmetric Hessians in
 * xx,xy,xz,yy,yz,zz order for the HESS_VAR fields; G[6][3] with
 * G[jk][i] = Gt^i_jk (jk in xx..zz order).  G == NULL computes Gt^i_jk
 * from the metric derivatives.  Writes rhs[24] and, if H != NULL, the
 * constraints H and M[3]. */
static void bssn_point(const double *u, const double (*d1)[3], const double (*dd)[6], const double (*Gin)[3],
                       const orc_bssn_par *par, const double *advin, double *rhs, double *Hc, double *Mc)
{
    const double chi = u[0], chid = fmax(u[0], par->chi_floor);
    double gt[3][3], At[3][3], gtu[3][3], Gam[3], beta[3], B[3];
    double dchi[3], dgt[3][3][3], dK[3], dAt[3][3][3], dGam[3][3], dal[3], db[3][3];
    double ddchi[3][3], ddgt[3][3][3][3], ddal[3][3], ddb[3][3][3];
    const double K = u[7], alpha = u[17];
    for (int i = 0; i < 3; i++) {
        for (int j = 0; j < 3; j++) {
            gt[i][j] = u[1 + SYMI[i][j]];
            At[i][j] = u[8 + SYMI[i][j]];
        }
        Gam[i] = u[14 + i];
        beta[i] = u[18 + i];
        B[i] = u[21 + i];
    }
    for (int k = 0; k < 3; k++) {
        dchi[k] = d1[0][k];
        dK[k] = d1[7][k];
        dal[k] = d1[17][k];
        for (int i = 0; i < 3; i++) {
            dGam[i][k] = d1[14 + i][k];
            db[i][k] = d1[18 + i][k];
            for (int j = 0; j < 3; j++) {
                dgt[i][j][k] = d1[1 + SYMI[i][j]][k];
                dAt[i][j][k] = d1[8 + SYMI[i][j]][k];
            }
        }
    }
    for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) {
            ddchi[k][l] = dd[0][SYMI[k][l]];
            ddal[k][l] = dd[7][SYMI[k][l]];
            for (int i = 0; i < 3; i++) {
                ddb[i][k][l] = dd[8 + i][SYMI[k][l]];
                for (int j = 0; j < 3; j++) ddgt[i][j][k][l] = dd[1 + SYMI[i][j]][SYMI[k][l]];
            }
        }
    inv3(gt, gtu);

    /* Christoffels: lowered from the metric derivatives, raised as given */
    double Gl[3][3][3], G[3][3][3], Gd[3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) Gl[i][j][k] = 0.5 * (dgt[i][j][k] + dgt[i][k][j] - dgt[j][k][i]);
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                if (Gin) {
                    G[i][j][k] = Gin[SYMI[j][k]][i];
                } else {
                    double s = 0;
                    for (int l = 0; l < 3; l++) s += gtu[i][l] * Gl[l][j][k];
                    G[i][j][k] = s;
                }
            }
    for (int i = 0; i < 3; i++) {
        double s = 0;
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) s += gtu[j][k] * G[i][j][k];
        Gd[i] = s;
    }

    double divb = db[0][0] + db[1][1] + db[2][2];
    double Atu[3][3], Atm[3][3], AA = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0, m = 0;
            for (int k = 0; k < 3; k++) {
                m += gtu[i][k] * At[k][j];
                for (int l = 0; l < 3; l++) s += gtu[i][k] * gtu[j][l] * At[k][l];
            }
            Atu[i][j] = s;
            Atm[i][j] = m;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) AA += At[i][j] * Atu[i][j];

    /* advection beta^k d_k of each field: centred, from the gradients, or
     * (advin != NULL) given per field -- the upwinded advection plus the
     * Kreiss-Oliger dissipation that the FD computes (see fd64_adv) */
    double adv[BSSN_NV];
    for (int v = 0; v < BSSN_NV; v++)
        adv[v] = advin ? advin[v] : beta[0] * d1[v][0] + beta[1] * d1[v][1] + beta[2] * d1[v][2];

    /* covariant second derivatives of chi and alpha */
    double DDchi[3][3], DDalt[3][3], DDal[3][3];
    double dchi_dal = 0, dchi2 = 0;
    for (int k = 0; k < 3; k++)
        for (int l = 0; l < 3; l++) {
            dchi_dal += gtu[k][l] * dchi[k] * dal[l];
            dchi2 += gtu[k][l] * dchi[k] * dchi[l];
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double sc = 0, sa = 0;
            for (int k = 0; k < 3; k++) {
                sc += G[k][i][j] * dchi[k];
                sa += G[k][i][j] * dal[k];
            }
            DDchi[i][j] = ddchi[i][j] - sc;
            DDalt[i][j] = ddal[i][j] - sa;
            DDal[i][j] = DDalt[i][j] + (dchi[i] * dal[j] + dchi[j] * dal[i] - gt[i][j] * dchi_dal) / (2 * chid);
        }
    double trDDchi = 0, trDDalt = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            trDDchi += gtu[i][j] * DDchi[i][j];
            trDDalt += gtu[i][j] * DDalt[i][j];
        }
    const double lap_alpha = chi * trDDalt - 0.5 * dchi_dal;  /* D^i D_i alpha */

    /* Ricci tensor */
    double R[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double r = 0;
            for (int l = 0; l < 3; l++)
                for (int m = 0; m < 3; m++) r -= 0.5 * gtu[l][m] * ddgt[i][j][l][m];
            for (int k = 0; k < 3; k++) {
                r += 0.5 * (gt[k][i] * dGam[k][j] + gt[k][j] * dGam[k][i]);
                r += 0.5 * Gd[k] * (Gl[i][j][k] + Gl[j][i][k]);
            }
            for (int l = 0; l < 3; l++)
                for (int m = 0; m < 3; m++)
                    for (int k = 0; k < 3; k++)
                        r += gtu[l][m] * (G[k][l][i] * Gl[j][k][m] + G[k][l][j] * Gl[i][k][m] + G[k][i][m] * Gl[k][l][j]);
            r += (DDchi[i][j] + gt[i][j] * trDDchi) / (2 * chid) -
                 (dchi[i] * dchi[j] + 3 * gt[i][j] * dchi2) / (4 * chid * chid);
            R[i][j] = r;
        }

    /* right-hand sides */
    rhs[0] = adv[0] + (2.0 / 3.0) * chi * (alpha * K - divb);
    for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++) {
            double s = adv[1 + SYMI[i][j]] - (2.0 / 3.0) * gt[i][j] * divb - 2 * alpha * At[i][j];
            for (int k = 0; k < 3; k++) s += gt[i][k] * db[k][j] + gt[j][k] * db[k][i];
            rhs[1 + SYMI[i][j]] = s;
        }
    rhs[7] = adv[7] - lap_alpha + alpha * (AA + K * K / 3.0);
    double X[3][3], trX = 0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) X[i][j] = -DDal[i][j] + alpha * R[i][j];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) trX += gtu[i][j] * X[i][j];
    for (int i = 0; i < 3; i++)
        for (int j = i; j < 3; j++) {
            double s = adv[8 + SYMI[i][j]] - (2.0 / 3.0) * At[i][j] * divb +
                       chi * (X[i][j] - gt[i][j] * trX / 3.0) + alpha * K * At[i][j];
            for (int k = 0; k < 3; k++) {
                s += At[i][k] * db[k][j] + At[j][k] * db[k][i];
                s -= 2 * alpha * At[i][k] * Atm[k][j];
            }
            rhs[8 + SYMI[i][j]] = s;
        }
    for (int i = 0; i < 3; i++) {
        double s = adv[14 + i] + (2.0 / 3.0) * Gam[i] * divb;
        double ddiv = 0;
        for (int k = 0; k < 3; k++) s -= Gam[k] * db[i][k];
        for (int j = 0; j < 3; j++)
            for (int k = 0; k < 3; k++) {
                s += gtu[j][k] * ddb[i][j][k];
                s += 2 * alpha * G[i][j][k] * Atu[j][k];
            }
        for (int j = 0; j < 3; j++) {
            ddiv = 0;
            for (int k = 0; k < 3; k++) ddiv += ddb[k][j][k];
            s += gtu[i][j] * ddiv / 3.0;
            s -= 2 * Atu[i][j] * dal[j];
            s -= 2 * alpha * (1.5 * Atu[i][j] * dchi[j] / chid + (2.0 / 3.0) * gtu[i][j] * dK[j]);
        }
        rhs[14 + i] = s;
    }
    rhs[17] = adv[17] - (par->slicing == 1 ? alpha * alpha * K : 2 * alpha * K);
    for (int i = 0; i < 3; i++) {
        if (advin) {
            /* adv[18..23] hold the shift's advection (with advect_shift)
             * plus its dissipation; d_0 Gt in the B equation drops Gt's
             * advection term, dissipation included, with advect_shift */
            rhs[18 + i] = adv[18 + i] + par->shift_coeff * B[i];
            rhs[21 + i] = adv[21 + i] + rhs[14 + i] - (par->advect_shift ? adv[14 + i] : 0.0) - par->eta * B[i];
        } else {
            rhs[18 + i] = (par->advect_shift ? adv[18 + i] : 0.0) + par->shift_coeff * B[i];
            rhs[21 + i] = (par->advect_shift ? adv[21 + i] - adv[14 + i] : 0.0) + rhs[14 + i] - par->eta * B[i];
        }
    }

    if (Hc) {
        double Rs = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Rs += gtu[i][j] * R[i][j];
        *Hc = chi * Rs + (2.0 / 3.0) * K * K - AA;
        for (int i = 0; i < 3; i++) {
            double m = -(2.0 / 3.0) * dK[i];
            for (int j = 0; j < 3; j++)
                for (int k = 0; k < 3; k++) {
                    double D = dAt[k][i][j];
                    for (int l = 0; l < 3; l++) D -= G[l][j][k] * At[l][i] + G[l][j][i] * At[k][l];
                    m += gtu[j][k] * D;
                    m -= 1.5 / chid * gtu[j][k] * At[i][j] * dchi[k];
                }
            Mc[i] = m;
        }
    }
}

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

/* RK4 stage combination shared by every stage kernel here:
 *   stage 0: acc = k;      new = y_n + dt/2 k
 *   stage 1: acc += 2k;    new = y_n + dt/2 k
 *   stage 2: acc += 2k;    new = y_n + dt k
 *   stage 3:               new = y_n + dt/6 (acc + k)      */
static inline void rk4_combine(int stage, double dt, double yn, double k, double acc_in, double *acc_out,
                               double *out)
{
    switch (stage) {
    case 0: *acc_out = k; *out = yn + 0.5 * dt * k; break;
    case 1: *acc_out = acc_in + 2 * k; *out = yn + 0.5 * dt * k; break;
    case 2: *acc_out = acc_in + 2 * k; *out = yn + dt * k; break;
    default: *acc_out = acc_in; *out = yn + dt / 6.0 * (acc_in + k); break;
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

/* Order-2q gradient and diagonal Hessian plus the mixed partials with the
 * reference's edge formula (AMR.h:1152-1266), in double, unquantised, at
 * one interior cell.  inv_dx = 2^depth / outer_dx. */
static void fd64_cell(const double *f, int64_t rw, int S, int T, int q, int x, int y, int z, double inv_dx,
                      double *g, double *h)
{
    const int base = q * q + q - 1;
    const int64_t T2 = (int64_t)T * T;
    double vx = 0, vy = 0, vz = 0, xx = 0, yy = 0, zz = 0;
    for (int i = -q; i <= q; i++) {
        vx += f[rw + i] * CD1[base + i];
        xx += f[rw + i] * CD2[base + i];
        vy += f[rw + (int64_t)i * T] * CD1[base + i];
        yy += f[rw + (int64_t)i * T] * CD2[base + i];
        vz += f[rw + (int64_t)i * T2] * CD1[base + i];
        zz += f[rw + (int64_t)i * T2] * CD2[base + i];
    }
    g[0] = vx * inv_dx; g[1] = vy * inv_dx; g[2] = vz * inv_dx;
    if (!h) return;
    const double i2 = inv_dx * inv_dx;
    const int ex = x - (x < 1 ? 1 : (x > S - 2 ? S - 2 : x));
    const int ey = y - (y < 1 ? 1 : (y > S - 2 ? S - 2 : y));
    const int ez = z - (z < 1 ? 1 : (z > S - 2 ? S - 2 : z));
    const int c[3] = {ex * ey, ex * ez, ey * ez};
    const int64_t dA[3] = {1, 1, T}, dB[3] = {T, T2, T2};
    double m[3];
    for (int k = 0; k < 3; k++) {
        const int64_t a = dA[k], b = dB[k];
        if (c[k] == 0) {
            m[k] = .25 * (f[rw + a + b] + f[rw - a - b] - f[rw + a - b] - f[rw - a + b]);
        } else {
            const int64_t da = -c[k] * a;
            m[k] = -c[k] * (f[rw] + .5 * (f[rw + da + b] + f[rw - da - b] - f[rw - da] - f[rw + da] - f[rw - b] -
                                          f[rw + b]));
        }
    }
    h[0] = xx * i2; h[1] = m[0] * i2; h[2] = m[1] * i2;
    h[3] = yy * i2; h[4] = m[2] * i2; h[5] = zz * i2;
}

/* Upwinded advection and Kreiss-Oliger dissipation of the FP64 scheme (an
 * extension for the moving-puncture runs; no reference counterpart):
 *   adv = [advect] sum_k beta^k (D^lop_k f)  +  sigma/(64 h) sum_k (Delta^6_k f)
 * D^lop_k: fourth-order one-sided-by-one first derivative on the side
 * beta^k points to, (-3 f_-1 - 10 f_0 + 18 f_1 - 6 f_2 + f_3) / (12 h) for
 * beta^k > 0 and its mirror (-f_-3 + 6 f_-2 - 18 f_-1 + 10 f_0 + 3 f_1) /
 * (12 h) otherwise (exact on polynomials of degree <= 4); Delta^6_k: the
 * sixth difference f_-3 - 6 f_-2 + 15 f_-1 - 20 f_0 + 15 f_1 - 6 f_2 + f_3
 * (zero on polynomials of degree <= 5; -64 f on the mode (-1)^k), so the
 * term damps the highest resolvable mode at rate sigma/h per axis.  The
 * literature's form of both, e.g. Brugmann et al., Phys. Rev. D 77, 024027
 * (2008).  Off until orc_bssn64_set_dissipation(sigma, 1); the dissipation
 * comes only with the upwinded advection. */
static double g_ko_sigma = 0.0;
static int g_upwind = 0;
EXPORT void orc_bssn64_set_dissipation(double sigma, int upwind)
{
    g_ko_sigma = sigma;
    g_upwind = upwind;
}
static const double LOP_PLUS[5] = {-3. / 12, -10. / 12, 18. / 12, -6. / 12, 1. / 12};   /* offsets -1..3 */
static const double LOP_MINUS[5] = {-1. / 12, 6. / 12, -18. / 12, 10. / 12, 3. / 12};  /* offsets -3..1 */
static const double KO6[7] = {1., -6., 15., -20., 15., -6., 1.};
static double fd64_adv(const double *f, int64_t rw, int T, const double *beta, double inv_dx, double sigma,
                       int advect)
{
    const int64_t st[3] = {1, T, (int64_t)T * T};
    double adv = 0.0, ko = 0.0;
    for (int k = 0; k < 3; k++) {
        const int64_t d = st[k];
        double l = 0.0, kk = 0.0;
        if (beta[k] > 0)
            for (int m = -1; m <= 3; m++) l += LOP_PLUS[m + 1] * f[rw + m * d];
        else
            for (int m = -3; m <= 1; m++) l += LOP_MINUS[m + 3] * f[rw + m * d];
        for (int m = -3; m <= 3; m++) kk += KO6[m + 3] * f[rw + m * d];
        adv += beta[k] * l * inv_dx;
        ko += kk;
    }
    return (advect ? adv : 0.0) + sigma / 64.0 * inv_dx * ko;
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
                        fd64_cell(state[v], r, S, T, q, x, y, z, inv_dx[depth], d1[v], has_h ? dd[h] : NULL);
                        if (has_h) h++;
                    }
                    double advin[BSSN_NV];
                    if (g_upwind) {
                        const double beta[3] = {u[18], u[19], u[20]};
                        for (int v = 0; v < BSSN_NV; v++)
                            advin[v] = fd64_adv(state[v], r, T, beta, inv_dx[depth], g_ko_sigma,
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
                        fd64_cell(state[v], r, S, T, q, x, y, z, inv_dx[depth], d1[v], has_h ? dd[h] : NULL);
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
