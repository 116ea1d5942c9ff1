// Learned coarse-to-fine ghost fill for gfx950: the per-octant MLP of
// am_mlp.hip evaluated inside the ghost fill.  No counterpart in the
// reference.
//
// The plain tricubic stage fill runs first over every face; this kernel then
// overwrites the cells it predicts.  Faces whose target is one level coarser
// and not the root are covered by a box of coarse cells of the target (2 x 8
// x 8 for S = 16, P = 3); each coarse cell is one sample, and one workgroup
// (NW waves, 8 threads per sample) takes a block of 8 NW samples:
//   1. per field: gather the cell's 4^3 stencil (offsets -1..2) of the
//      target at the stage time (quadratic in time through older/old/new,
//      as the tricubic fill), normalise by the centre value mu and the
//      spread sigma, add the tricubic baseline of the 8 children, and
//      accumulate layer 1 (8 NW samples x 64 inputs x 128) on MFMA;
//   2. layers 2 and 3 on MFMA from LDS; output + baseline = normalised
//      children;
//   3. the restriction residual |mean(children) - mu| (field units) is the
//      block's self-check; the children are then shifted so their mean is
//      mu;
//   4. if the block's largest residual is within the threshold, every ghost
//      cell of its samples takes its child's value.
// A sample whose stencil would read an edge or corner ghost of the target
// (never filled) keeps the tricubic value, as does every cell of a block
// that fails the check.
#include "am_common.hpp"
#include "am_device.hpp"
#include "am_interp.hpp"

namespace am {

namespace {
constexpr int L_IN = 1536, L_HID = 128, L_OUT = 192;
constexpr int64_t L_W1 = 0, L_B1 = (int64_t)L_IN * L_HID, L_W2 = L_B1 + L_HID, L_B2 = L_W2 + (int64_t)L_HID * L_HID,
                  L_W3 = L_B2 + L_HID, L_B3 = L_W3 + (int64_t)L_HID * L_OUT;
typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu_l(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }

// 4-point Lagrange weights through -1, 0, 1, 2 at the children's offsets -1/4, +1/4
__device__ constexpr float LQ[2][4] = {{15.f / 128, 135.f / 128, -27.f / 128, 5.f / 128},
                                       {-7.f / 128, 105.f / 128, 35.f / 128, -5.f / 128}};
}  // namespace

template <int S, int P>
__device__ __forceinline__ void ghost_cell_of(int dir, int c, uint32_t &w0, uint32_t &w1, uint32_t &w2)
{
    constexpr uint32_t uP = P, uS = S;
    if (dir == 0 || dir == 3) {
        const uint32_t l = c % P, a = (c / P) % S, bq = c / (P * S);
        w0 = dir == 0 ? l + uP + uS : l; w1 = a + uP; w2 = bq + uP;
    } else if (dir == 1 || dir == 4) {
        const uint32_t a = c % S, l = (c / S) % P, bq = c / (S * P);
        w0 = a + uP; w1 = dir == 1 ? l + uP + uS : l; w2 = bq + uP;
    } else {
        const uint32_t a = c % S, bq = (c / S) % S, l = c / (S * S);
        w0 = a + uP; w1 = bq + uP; w2 = dir == 2 ? l + uP + uS : l;
    }
}

// the target coordinate of fine index w along one axis (AMR.h:736-745)
template <int S, int P>
__device__ __forceinline__ float read_pos(uint32_t w, float rel, float rel_scale)
{
    float r = ((float)w - (float)P + .5f) * (1.f / (float)S) - .5f;
    r = r * rel_scale + rel;
    return (r + .5f) * (float)S + (float)P - 0.5f;
}

// NW = 8 (64-sample blocks, 128 KB of LDS, one workgroup per CU) serves the
// per-block gate, whose blocks are 64 samples.  NW = 4 (32 samples, 64 KB)
// serves the per-tile gate, where the block size does not show in the
// result: two workgroups share a CU, so one computes while the other waits
// on its field barrier or its gathers.  Every output element goes through
// the same operations in the same order for either NW.
template <int S, int P, int L, int NW>
__global__ void __launch_bounds__(64 * NW, NW == 4 ? 3 : 1) k_ghost_learned(FieldBatchDev b, const uint64_t *__restrict__ bnd,
                                                       const int32_t *__restrict__ tiles,
                                                       const int32_t *__restrict__ faces, uint32_t substep,
                                                       uint32_t maxd, float frac, const float *__restrict__ prm,
                                                       float threshold, unsigned long long *__restrict__ stats,
                                                       int n_blocks, float *__restrict__ tile_resid)
{
    constexpr int NC = S * S * P;
    constexpr size_t CELLS = tile_stride<S, P, L>();
    constexpr int T = S + 2 * P;
    constexpr int NT = 64 * NW, NSMP = 8 * NW;  // threads, samples per block
    constexpr int RB = NSMP / 16, CT = L_HID / 16 / NW;  // row blocks of samples, hidden column tiles per wave
    static_assert(NW == 4 || NW == 8, "8 threads per sample, 16-row MFMA blocks, 128 hidden columns");
    // the input tiles (double-buffered per field) and the hidden layer share
    // their LDS: the hidden layer is written after the barrier that ends the
    // last field's reads of the inputs
    constexpr int XH = 2 * NSMP * 65 > NSMP * 129 ? 2 * NSMP * 65 : NSMP * 129;
    __shared__ float xh[XH];
    float(*const Xs)[NSMP][65] = reinterpret_cast<float(*)[NSMP][65]>(xh);
    float(*const Hs)[129] = reinterpret_cast<float(*)[129]>(xh);
    __shared__ float Bn[NSMP][193];
    __shared__ float musg[2][24][NSMP];
    __shared__ int valid[NSMP];
    __shared__ int face_resid_bits;
    __shared__ int n_invalid;

    const int face = blockIdx.x / n_blocks, blk = blockIdx.x - face * n_blocks;
    int node, dir;
    if (faces) {
        const int code = faces[face];
        node = code / 6;
        dir = code - node * 6;
    } else {
        const int ti = face / 6;
        dir = face - ti * 6;
        node = tiles[ti];
    }
    const uint64_t bb = bnd[(size_t)node * 6 + dir];
    const uint32_t nd = (uint32_t)bb;
    const int tgt = bnd_target(nd), cd = bnd_depth(nd);
    const uint32_t ddiff = bnd_ddiff(nd);
    if (tgt <= 0 || cd <= 0 || ddiff != 1u) return;  // uniform per workgroup
    if (!active_depth(substep, (uint32_t)cd, maxd)) return;
    const float3 relp = cf3_unpack((uint32_t)(bb >> 32));
    const float rl[3] = {relp.x, relp.y, relp.z};
    const float rel_scale = .5f;
    const float tmp = (float)((substep >> (maxd - (uint32_t)cd)) & 1u);
    const float ty = (tmp + frac) * .5f;
    const bool quad = b.older[0] != nullptr;
    const float tc0 = quad ? .5f * ty * (ty - 1.f) : 0.f;
    const float tc1 = quad ? (1.f - ty) * (1.f + ty) : 1.f - ty;
    const float tc2 = quad ? .5f * ty * (ty + 1.f) : ty;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // coarse cells covered by the ghost slab, per axis
    int clo[3], ncl[3];
    {
        const int n = dir % 3;
        for (int k = 0; k < 3; k++) {
            const uint32_t wlo = k != n ? (uint32_t)P : (dir < 3 ? (uint32_t)(P + S) : 0u);
            const uint32_t whi = k != n ? (uint32_t)(P + S - 1) : (dir < 3 ? (uint32_t)(2 * P + S - 1) : (uint32_t)(P - 1));
            const int a = (int)floorf(read_pos<S, P>(wlo, rl[k], rel_scale) + .5f);
            const int z = (int)floorf(read_pos<S, P>(whi, rl[k], rel_scale) + .5f);
            clo[k] = a;
            ncl[k] = z - a + 1;
        }
    }
    const int NS = ncl[0] * ncl[1] * ncl[2];
    const int chunk = blk * NSMP;
    if (chunk >= NS) return;  // uniform per workgroup
    if (tid == 0) {
        face_resid_bits = 0;
        n_invalid = 0;
    }
    const size_t tbase = (size_t)tgt * CELLS;
    const int ns = min(NSMP, NS - chunk);
    if (tid < NSMP) valid[tid] = 1;
    __syncthreads();
    floatx4 acc[RB][CT];
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int c = 0; c < CT; c++) acc[i][c] = floatx4{0.f, 0.f, 0.f, 0.f};
    // sample s and its eighth q of the 4^3 stencil: plane k = q / 2, rows
    // j = 2 (q % 2) + {0, 1}, 4 points each
    const int s = tid >> 3, q = tid & 7;
    int cc[3] = {0, 0, 0};
    if (s < ns) {
        const int sa = chunk + s;
        cc[0] = clo[0] + sa % ncl[0];
        cc[1] = clo[1] + (sa / ncl[0]) % ncl[1];
        cc[2] = clo[2] + sa / (ncl[0] * ncl[1]);
    }
    // 1. per field: the stencil at the stage time (8 points per thread),
    // normalised in registers -- mu is stencil point (1, 1, 1), held by
    // lane 2 of the sample's 8, sigma the largest deviation over the 8 --
    // then into the double-buffered LDS tile.  The next field's points and
    // this wave's layer-1 weights of the next field are loaded while this
    // field's baseline and layer-1 products run.
    float raw[8];
    bool bad = false;
    // The 8 points' offsets first (the interior-first map branches), then
    // all their loads back to back, so they are in flight together: a point
    // outside the stored cells (beyond the tile, or an edge or corner ghost)
    // loads the tile's first interior cell instead and takes 0.
    uint32_t pofs[8];
    uint32_t pin = 0u;  // bit i: point i is a stored cell
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int px = cc[0] - 1 + (i & 3), py = cc[1] - 1 + 2 * (q & 1) + (i >> 2), pz = cc[2] - 1 + (q >> 1);
        const int out = (px < P || px >= P + S) + (py < P || py >= P + S) + (pz < P || pz >= P + S);
        const bool inside = s < ns && px >= 0 && px < T && py >= 0 && py < T && pz >= 0 && pz < T && out <= 1;
        pofs[i] = tile_cell<S, P, L>(inside ? (uint32_t)px : (uint32_t)P, inside ? (uint32_t)py : (uint32_t)P,
                                     inside ? (uint32_t)pz : (uint32_t)P);
        pin |= (uint32_t)inside << i;
        bad |= s < ns && !inside;
    }
    // No branch anywhere from these loads to the field's MFMAs (the third
    // time level is read from old when there is none and not used), so the
    // scheduler can leave the combination of the next field's points until
    // after this field's work.
    auto gather = [&](int v) {
        const float *dold = b.old_[v] + tbase, *dnew = b.new_[v] + tbase;
        const float *dolder = quad ? b.older[v] + tbase : dold;
        float f_old[8], f_new[8], f_older[8];
#pragma unroll
        for (int i = 0; i < 8; i++) {
            f_old[i] = dold[pofs[i]];
            f_new[i] = dnew[pofs[i]];
            f_older[i] = dolder[pofs[i]];
        }
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float val = f_old[i] * tc1 + f_new[i] * tc2;
            const float val3 = f_older[i] * tc0 + val;
            raw[i] = (pin >> i) & 1u ? (quad ? val3 : val) : 0.f;
        }
    };
    // layer-1 weights of field v for this lane and column tile ct:
    // W1[v * 64 + 4 kk + lane / 16][16 (CT wave + ct) + lane % 16], kk = 0..15
    float wcur[CT][16], wnext[NW == 8 ? CT : 1][16];
    auto load_w1 = [&](int v, float (&w)[CT][16]) {
#pragma unroll
        for (int ct = 0; ct < CT; ct++) {
            const float *wrow =
                prm + L_W1 + (int64_t)(v * 64 + (lane >> 4)) * L_HID + (CT * wave + ct) * 16 + (lane & 15);
#pragma unroll
            for (int kk = 0; kk < 16; kk++) w[ct][kk] = wrow[(int64_t)kk * 4 * L_HID];
        }
    };
    // NW = 8 loads the next field's weights a field ahead; NW = 4, which
    // holds twice the weights per wave, loads each field's at its start
    // (their latency then overlaps the field's normalisation, barrier and
    // baseline) to stay within two waves' registers per SIMD.
    gather(0);
    if constexpr (NW == 8) load_w1(0, wnext);
    for (int v = 0; v < 24; v++) {
        const int buf = v & 1;
        if constexpr (NW == 8) {
#pragma unroll
            for (int ct = 0; ct < CT; ct++)
#pragma unroll
                for (int kk = 0; kk < 16; kk++) wcur[ct][kk] = wnext[ct][kk];
        } else {
            load_w1(v, wcur);
        }
        const float mu = __shfl(raw[5], (lane & ~7) | 2, 64);
        float dev = 0.f;
#pragma unroll
        for (int i = 0; i < 8; i++) dev = fmaxf(dev, fabsf(raw[i] - mu));
        dev = fmaxf(dev, __shfl_xor(dev, 1, 64));
        dev = fmaxf(dev, __shfl_xor(dev, 2, 64));
        dev = fmaxf(dev, __shfl_xor(dev, 4, 64));
        const float sg = dev + 1e-3f * fabsf(mu) + 1e-7f;
        const float inv = 1.f / sg;
#pragma unroll
        for (int i = 0; i < 8; i++) Xs[buf][s][q * 8 + i] = s < ns ? (raw[i] - mu) * inv : 0.f;
        if (s < ns && q == 0) {
            musg[0][v][s] = mu;
            musg[1][v][s] = sg;
        }
        if (bad) valid[s] = 0;
        // Xs[buf] is complete; every wave has finished reading Xs[buf ^ 1]
        __syncthreads();
        // the last field gathers field 23 again, unused
        gather(v + 1 < 24 ? v + 1 : v);
        if constexpr (NW == 8) load_w1(v + 1 < 24 ? v + 1 : v, wnext);
        // baseline of the 8 children (normalised units): child q of sample s
        // (the rows of absent samples hold zeros and are never read)
        // (separable: along x, then y, then z, 84 fused multiply-adds)
        {
            const int jx = q & 1, jy = (q >> 1) & 1, jz = q >> 2;
            float wx[4], wy[4], wz[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                wx[i] = LQ[jx][i];
                wy[i] = LQ[jy][i];
                wz[i] = LQ[jz][i];
            }
            float bz = 0.f;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                float by = 0.f;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    float bx = 0.f;
#pragma unroll
                    for (int i = 0; i < 4; i++) bx = fmaf(wx[i], Xs[buf][s][k * 16 + j * 4 + i], bx);
                    by = fmaf(wy[j], bx, by);
                }
                bz = fmaf(wz[k], by, bz);
            }
            Bn[s][v * 8 + q] = bz;
        }
        // layer 1 contribution of this field: Z1[:, this wave's columns] += Xs (NSMP x 64) W1[v*64.., columns]
#pragma unroll
        for (int kk = 0; kk < 16; kk++) {
            const int ka = kk * 4 + (lane >> 4);
#pragma unroll
            for (int rb = 0; rb < RB; rb++) {
                const float a = Xs[buf][rb * 16 + (lane & 15)][ka];
#pragma unroll
                for (int ct = 0; ct < CT; ct++)
                    acc[rb][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, wcur[ct][kk], acc[rb][ct], 0, 0, 0);
            }
        }
    }
    __syncthreads();
    // layer 1 epilogue
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = (CT * wave + ct) * 16 + (lane & 15);
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                Hs[rb * 16 + 4 * (lane >> 4) + r][col] = gelu_l(acc[rb][ct][r] + prm[L_B1 + col]);
    }
    __syncthreads();
    // layer 2
#pragma unroll
    for (int i = 0; i < RB; i++)
#pragma unroll
        for (int c = 0; c < CT; c++) acc[i][c] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = (CT * wave + ct) * 16 + (lane & 15);
        const float *wrow = prm + L_W2 + (int64_t)(lane >> 4) * L_HID + col;
#pragma unroll 8
        for (int kk = 0; kk < 32; kk++) {
            const int ka = kk * 4 + (lane >> 4);
            const float b0 = wrow[(int64_t)kk * 4 * L_HID];
#pragma unroll
            for (int rb = 0; rb < RB; rb++)
                acc[rb][ct] =
                    __builtin_amdgcn_mfma_f32_16x16x4f32(Hs[rb * 16 + (lane & 15)][ka], b0, acc[rb][ct], 0, 0, 0);
        }
    }
    __syncthreads();
#pragma unroll
    for (int ct = 0; ct < CT; ct++) {
        const int col = (CT * wave + ct) * 16 + (lane & 15);
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                Hs[rb * 16 + 4 * (lane >> 4) + r][col] = gelu_l(acc[rb][ct][r] + prm[L_B2 + col]);
    }
    __syncthreads();
    // layer 3: 12 output tiles of 16 columns, tiles wave, wave + NW, ..., added onto the baseline
    for (int t = wave; t < 12; t += NW) {
        const int oc = t * 16 + (lane & 15);
        floatx4 acc3[RB];
#pragma unroll
        for (int i = 0; i < RB; i++) acc3[i] = floatx4{0.f, 0.f, 0.f, 0.f};
        const float *wrow = prm + L_W3 + (int64_t)(lane >> 4) * L_OUT + oc;
#pragma unroll 8
        for (int kk = 0; kk < 32; kk++) {
            const int ka = kk * 4 + (lane >> 4);
            const float b0 = wrow[(int64_t)kk * 4 * L_OUT];
#pragma unroll
            for (int rb = 0; rb < RB; rb++)
                acc3[rb] = __builtin_amdgcn_mfma_f32_16x16x4f32(Hs[rb * 16 + (lane & 15)][ka], b0, acc3[rb], 0, 0, 0);
        }
#pragma unroll
        for (int rb = 0; rb < RB; rb++)
#pragma unroll
            for (int r = 0; r < 4; r++) Bn[rb * 16 + 4 * (lane >> 4) + r][oc] += acc3[rb][r] + prm[L_B3 + oc];
    }
    __syncthreads();
    // residual of the restriction (field units) and the shift that removes it
    if (s < ns) {
        float rmax = 0.f;
        for (int v = q * 3; v < q * 3 + 3; v++) {
            float m = 0.f;
#pragma unroll
            for (int ch = 0; ch < 8; ch++) m += Bn[s][v * 8 + ch];
            m *= .125f;
            rmax = fmaxf(rmax, fabsf(m * musg[1][v][s]));
#pragma unroll
            for (int ch = 0; ch < 8; ch++) Bn[s][v * 8 + ch] -= m;
        }
        if (valid[s]) atomicMax(&face_resid_bits, __float_as_int(rmax));  // non-negative: int order = float order
        if (q == 0 && !valid[s]) atomicAdd(&n_invalid, 1);
    }
    __syncthreads();
    // 4. the block's samples' slab cells: with the per-block gate when the
    // block is within the threshold; with the per-tile gate always, the
    // block's residual going to the tile's (a later pass refills the tiles
    // beyond the threshold)
    const bool pass = tile_resid || __int_as_float(face_resid_bits) <= threshold;
    if (pass) {
        for (int c = tid; c < NC; c += NT) {
            uint32_t w[3];
            ghost_cell_of<S, P>(dir, c, w[0], w[1], w[2]);
            int cl[3], child = 0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const float rp = read_pos<S, P>(w[k], rl[k], rel_scale);
                cl[k] = (int)floorf(rp + .5f);
                child |= (rp > (float)cl[k] ? 1 : 0) << k;
            }
            const int sa = (cl[0] - clo[0]) + ncl[0] * ((cl[1] - clo[1]) + ncl[1] * (cl[2] - clo[2]));
            if (sa < chunk || sa >= chunk + ns) continue;
            const int sl = sa - chunk;
            if (!valid[sl]) continue;
            const size_t wn = (size_t)node * CELLS + tile_cell<S, P, L>(w[0], w[1], w[2]);
            for (int v = 0; v < 24; v++)
                b.new_[v][wn] = musg[0][v][sl] + musg[1][v][sl] * Bn[sl][v * 8 + child];
        }
    }
    if (tid == 0) {
        if (tile_resid) atomicMax(reinterpret_cast<int *>(tile_resid) + node, face_resid_bits);
        if (stats) {
            if (!tile_resid) atomicAdd(&stats[pass ? 0 : 1], 1ull);
            atomicAdd(&stats[2], (unsigned long long)n_invalid);
            atomicAdd(&stats[3], (unsigned long long)ns);
        }
    }
}

template <int S, int P>
static int launch_learned(const am_field_batch *fb, const am_mesh *mesh, const int32_t *faces, int n_faces,
                          const am_geometry *g, float frac, const float *params, float threshold,
                          unsigned long long *stats, float *tile_resid, int equal_mode, void *stream)
{
    if (fb->n_vars != 24) return set_error(AM_ERR_INVALID, "am_ghost_fill_learned: needs the 24 BSSN fields");
    if (fb->tile_layout != AM_LAYOUT_PADDED && fb->tile_layout != AM_LAYOUT_INTERIOR_FIRST)
        return set_error(AM_ERR_INVALID, "am_ghost_fill_learned: tile layout %d", fb->tile_layout);
    FieldBatchDev d;
    for (int v = 0; v < 24; v++) {
        if (!fb->old_[v] || !fb->new_[v]) return set_error(AM_ERR_INVALID, "am_ghost_fill_learned: null field %d", v);
        d.old_[v] = fb->old_[v];
        d.new_[v] = fb->new_[v];
        d.older[v] = fb->older[v];
        if ((fb->older[v] == nullptr) != (fb->older[0] == nullptr))
            return set_error(AM_ERR_INVALID, "am_ghost_fill_learned: older must be given for all fields or none");
    }
    // the tricubic stage fill everywhere (with the per-tile gate: except on
    // the network's faces, clearing the tiles' residuals), then the learned
    // cells on top
    LearnedGate gate;
    if (tile_resid) {
        gate.resid = tile_resid;
        gate.mode = 1;
    }
    if (int rc = ghost_stage_fill(fb, mesh, faces, n_faces, g, frac, equal_mode, stream, gate)) return rc;
    // samples per face: at most S/2 + 1 coarse cells per tangential axis and
    // P/2 + 1 along the normal (the slab's P fine layers)
    constexpr int NS_MAX = (S / 2 + 1) * (S / 2 + 1) * (P / 2 + 1);
    const unsigned items = faces ? (unsigned)n_faces : (unsigned)mesh->n_tiles * 6u;
    auto go = [&](auto kern, int nw) {
        const int nb = (NS_MAX + 8 * nw - 1) / (8 * nw);
        hipLaunchKernelGGL(kern, dim3(items * (unsigned)nb), dim3(64 * nw), 0, (hipStream_t)stream, d,
                           mesh->boundaries, mesh->tiles, faces, mesh->substep, mesh->max_depth, frac, params,
                           threshold, stats, nb, tile_resid);
    };
    // the per-tile gate does not see the block size: half-size blocks, two workgroups per CU
    if (fb->tile_layout == AM_LAYOUT_INTERIOR_FIRST) {
        if (tile_resid) go(k_ghost_learned<S, P, 1, 4>, 4);
        else go(k_ghost_learned<S, P, 1, 8>, 8);
    } else {
        if (tile_resid) go(k_ghost_learned<S, P, 0, 4>, 4);
        else go(k_ghost_learned<S, P, 0, 8>, 8);
    }
    if (int rc = check_launch("k_ghost_learned")) return rc;
    if (!tile_resid) return AM_OK;
    // the per-tile gate: the network's faces take the tricubic fill -- all
    // their cells for tiles beyond the threshold, else the cells the network
    // leaves
    gate.mode = 2;
    gate.threshold = threshold;
    gate.stats = stats;
    return ghost_stage_fill(fb, mesh, faces, n_faces, g, frac, equal_mode, stream, gate);
}

}  // namespace am

extern "C" int am_ghost_fill_learned(const am_field_batch *b, const am_mesh *mesh, const am_geometry *g,
                                     float stage_frac, const float *params, float threshold, uint64_t *stats,
                                     int32_t equal_mode, void *stream)
{
    if (!b || !mesh || !g || !params) return am::set_error(AM_ERR_INVALID, "am_ghost_fill_learned: null argument");
    if (!mesh->boundaries || !mesh->tiles) return am::set_error(AM_ERR_INVALID, "am_ghost_fill_learned: no tiles");
    if (equal_mode < AM_EQUAL_FORMULA || equal_mode > AM_EQUAL_SKIP)
        return am::set_error(AM_ERR_INVALID, "am_ghost_fill_learned: equal_mode %d", equal_mode);
    if (!(stage_frac >= 0.f && stage_frac <= 1.f))
        return am::set_error(AM_ERR_INVALID, "am_ghost_fill_learned: stage_frac %g", (double)stage_frac);
    if (mesh->n_tiles <= 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_learned, b, mesh, nullptr, 0, g, stage_frac, params, threshold,
                     reinterpret_cast<unsigned long long *>(stats), nullptr, equal_mode, stream);
}

extern "C" int am_ghost_fill_learned_faces(const am_field_batch *b, const am_mesh *mesh, const int32_t *faces,
                                           int32_t n_faces, const am_geometry *g, float stage_frac,
                                           const float *params, float threshold, float *tile_resid, uint64_t *stats,
                                           int32_t equal_mode, void *stream)
{
    if (!b || !mesh || !g || !params || !tile_resid)
        return am::set_error(AM_ERR_INVALID, "am_ghost_fill_learned_faces: null argument");
    if (!mesh->boundaries || !mesh->tiles)
        return am::set_error(AM_ERR_INVALID, "am_ghost_fill_learned_faces: no tiles");
    if (n_faces < 0 || (n_faces > 0 && !faces))
        return am::set_error(AM_ERR_INVALID, "am_ghost_fill_learned_faces: %d faces", n_faces);
    if (equal_mode < AM_EQUAL_FORMULA || equal_mode > AM_EQUAL_SKIP)
        return am::set_error(AM_ERR_INVALID, "am_ghost_fill_learned_faces: equal_mode %d", equal_mode);
    if (!(stage_frac >= 0.f && stage_frac <= 1.f))
        return am::set_error(AM_ERR_INVALID, "am_ghost_fill_learned_faces: stage_frac %g", (double)stage_frac);
    if (n_faces == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_learned, b, mesh, faces, n_faces, g, stage_frac, params, threshold,
                     reinterpret_cast<unsigned long long *>(stats), tile_resid, equal_mode, stream);
}
