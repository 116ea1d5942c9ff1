// Scalar wave equation u_tt = lap u as the system (u, v), RK4 stage update
// and FP64 energy reduction for gfx950.  The Laplacian is read from the
// stage input's d2diag word, so one stage reads 2 state values, 2 y_n / acc
// values and one word per cell: an HBM-bound pointwise pass like the BSSN
// stage, sharing its RK4 combine and equal-depth push.
#include "am_common.hpp"
#include "am_device.hpp"
#include "am_rk4.hpp"

namespace am {

struct SwaveArgsDev {
    const float *state[2];
    const float *old_[2];
    float *new_[2];
    float *acc[2];
    const uint32_t *d1;
    const uint32_t *d2;
    int push_layers;
    int push_axes;  // bits 1 x, 2 y, 4 z
    int layout;     // am_tile_layout of state, old_, new_
};

struct SwaveTiles {
    const uint32_t *positions;
    const int32_t *tiles;
    const uint64_t *boundaries;
    uint32_t substep, max_depth;
    int32_t n_items;  // tiles x 256-cell blocks; workgroups loop over them
};

template <int S, int P, int STAGE, bool PUSH, bool PERSIST, int L>
__global__ void __launch_bounds__(256) k_swave_stage(SwaveArgsDev a, SwaveTiles m, DepthScales sc)
{
    constexpr int S3 = S * S * S, BPT = (S3 + 255) / 256;
    const int item_end = PERSIST ? m.n_items : (int)blockIdx.x + 1;
    for (int item = blockIdx.x; item < item_end; item += PERSIST ? gridDim.x : 1) {
        const int ti = item / BPT, cb = item - ti * BPT;
        const int node = m.tiles[ti];
        const int depth = pos_depth(m.positions[node]);
        if (depth == -1 || !active_depth(m.substep, (uint32_t)depth, m.max_depth)) continue;
        const int cell = cb * 256 + (int)threadIdx.x;
        if (cell >= S3) continue;
        const int z = cell / (S * S), y = (cell / S) % S, x = cell % S;
        const size_t r = (size_t)node * tile_stride<S, P, L>() + tile_interior<S, P, L>(x, y, z);
        const size_t w = (size_t)node * S3 + cell;
        const float3 h = cf3_unpack(a.d2[w]);
        float k[2] = {a.state[1][r], (h.x + h.y) + h.z};
        rk4_combine<2, STAGE>(a.state, a.old_, a.acc, a.new_, k, r, w, sc.dt[depth]);
        if (PUSH) rk4_push<2, S, P, L>(a.new_, k, m.boundaries, node, x, y, z, a.push_layers, a.push_axes);
    }
}

// Per-workgroup FP64 partials {energy, sum u^2 h^3, cells, max |u|}.
template <int S, int P, int L>
__global__ void __launch_bounds__(256) k_swave_energy(SwaveArgsDev a, SwaveTiles m, double outer_dx, double *partial)
{
    constexpr int S3 = S * S * S, BPT = (S3 + 255) / 256;
    const int ti = blockIdx.x / BPT, cb = blockIdx.x - ti * BPT;
    const int node = m.tiles[ti];
    const int cell = cb * 256 + (int)threadIdx.x;
    double e = 0.0, u2 = 0.0, cnt = 0.0, umax = 0.0;
    if (cell < S3) {
        const int z = cell / (S * S), y = (cell / S) % S, x = cell % S;
        const size_t r = (size_t)node * tile_stride<S, P, L>() + tile_interior<S, P, L>(x, y, z);
        const int depth = pos_depth(m.positions[node]);
        const double hc = outer_dx / (double)(1u << depth), vol = hc * hc * hc;
        const float3 g = cf3_unpack(a.d1[(size_t)node * S3 + cell]);
        const double u = a.state[0][r], v = a.state[1][r];
        e = 0.5 * (v * v + (double)g.x * g.x + (double)g.y * g.y + (double)g.z * g.z) * vol;
        u2 = u * u * vol;
        cnt = 1.0;
        umax = fabs(u);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_down(e, off, 64);
        u2 += __shfl_down(u2, off, 64);
        cnt += __shfl_down(cnt, off, 64);
        umax = fmax(umax, __shfl_down(umax, off, 64));
    }
    __shared__ double red[4][4];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
        red[wave][0] = e; red[wave][1] = u2; red[wave][2] = cnt; red[wave][3] = umax;
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

static int swave_args(const am_swave_fields *f, SwaveArgsDev &a, bool need_stage)
{
    for (int v = 0; v < 2; v++) {
        a.state[v] = f->state[v];
        a.old_[v] = f->old_[v];
        a.new_[v] = f->new_[v];
        a.acc[v] = f->acc[v];
        if (!a.state[v]) return set_error(AM_ERR_INVALID, "am_swave: null state field %d", v);
        if (need_stage && (!a.old_[v] || !a.new_[v] || !a.acc[v]))
            return set_error(AM_ERR_INVALID, "am_swave: null old/new/acc field %d", v);
    }
    a.d1 = f->d1_u;
    a.d2 = f->d2diag_u;
    a.layout = f->tile_layout;
    if (a.layout != AM_LAYOUT_PADDED && a.layout != AM_LAYOUT_INTERIOR_FIRST)
        return set_error(AM_ERR_INVALID, "am_swave: tile layout %d", a.layout);
    if (need_stage ? !a.d2 : !a.d1) return set_error(AM_ERR_INVALID, "am_swave: null derivative word");
    return AM_OK;
}

template <int S, int P, bool PUSH, bool PERSIST, int L>
static void swave_launch_l(int stage, dim3 grid, hipStream_t st, const SwaveArgsDev &a, const SwaveTiles &m,
                           const DepthScales &sc)
{
    switch (stage) {
    case 0: hipLaunchKernelGGL((k_swave_stage<S, P, 0, PUSH, PERSIST, L>), grid, dim3(256), 0, st, a, m, sc); break;
    case 1: hipLaunchKernelGGL((k_swave_stage<S, P, 1, PUSH, PERSIST, L>), grid, dim3(256), 0, st, a, m, sc); break;
    case 2: hipLaunchKernelGGL((k_swave_stage<S, P, 2, PUSH, PERSIST, L>), grid, dim3(256), 0, st, a, m, sc); break;
    default: hipLaunchKernelGGL((k_swave_stage<S, P, 3, PUSH, PERSIST, L>), grid, dim3(256), 0, st, a, m, sc); break;
    }
}
template <int S, int P, bool PUSH, bool PERSIST>
static void swave_launch(int stage, dim3 grid, hipStream_t st, const SwaveArgsDev &a, const SwaveTiles &m,
                         const DepthScales &sc)
{
    if (a.layout == AM_LAYOUT_INTERIOR_FIRST) return swave_launch_l<S, P, PUSH, PERSIST, 1>(stage, grid, st, a, m, sc);
    swave_launch_l<S, P, PUSH, PERSIST, 0>(stage, grid, st, a, m, sc);
}

template <int S, int P>
static int swave_stage(const am_swave_fields *f, int stage, const am_mesh *mesh, const am_geometry *g, int push,
                       void *stream)
{
    constexpr int BPT = (S * S * S + 255) / 256;
    if (stage < 0 || stage > 3) return set_error(AM_ERR_INVALID, "am_swave_rhs_stage: stage %d", stage);
    if (push && !mesh->boundaries) return set_error(AM_ERR_INVALID, "am_swave_rhs_stage: push needs boundaries");
    SwaveArgsDev a;
    if (int rc = swave_args(f, a, true)) return rc;
    if (f->push_layers < 0 || f->push_layers > P)
        return set_error(AM_ERR_INVALID, "am_swave_rhs_stage: push_layers %d (0..%d)", f->push_layers, P);
    a.push_layers = f->push_layers ? f->push_layers : P;
    if (f->push_axes < 0 || f->push_axes > 7)
        return set_error(AM_ERR_INVALID, "%s: push_axes %d (0..7)", "am_swave_rhs_stage", f->push_axes);
    a.push_axes = f->push_axes ? f->push_axes : 7;
    const int n_items = mesh->n_tiles * BPT;
    const SwaveTiles m{mesh->positions, mesh->tiles, mesh->boundaries, mesh->substep, mesh->max_depth, n_items};
    const DepthScales sc = depth_scales(*g);
    const dim3 grid(mesh->grid_limit > 0 && mesh->grid_limit < n_items ? mesh->grid_limit : n_items);
    const bool persist = grid.x < (unsigned)n_items;
    hipStream_t st = (hipStream_t)stream;
    if (push && persist)
        swave_launch<S, P, true, true>(stage, grid, st, a, m, sc);
    else if (push)
        swave_launch<S, P, true, false>(stage, grid, st, a, m, sc);
    else if (persist)
        swave_launch<S, P, false, true>(stage, grid, st, a, m, sc);
    else
        swave_launch<S, P, false, false>(stage, grid, st, a, m, sc);
    return check_launch("k_swave_stage");
}

template <int S, int P>
static int swave_energy(const am_swave_fields *f, const am_mesh *mesh, const am_geometry *g, double *partial,
                        void *stream)
{
    constexpr int BPT = (S * S * S + 255) / 256;
    SwaveArgsDev a;
    if (int rc = swave_args(f, a, false)) return rc;
    const SwaveTiles m{mesh->positions, mesh->tiles, mesh->boundaries, mesh->substep, mesh->max_depth,
                       mesh->n_tiles * BPT};
    if (a.layout == AM_LAYOUT_INTERIOR_FIRST)
        hipLaunchKernelGGL((k_swave_energy<S, P, 1>), dim3(mesh->n_tiles * BPT), dim3(256), 0, (hipStream_t)stream, a,
                           m, (double)g->outer_dx, partial);
    else
        hipLaunchKernelGGL((k_swave_energy<S, P, 0>), dim3(mesh->n_tiles * BPT), dim3(256), 0, (hipStream_t)stream, a,
                           m, (double)g->outer_dx, partial);
    return check_launch("k_swave_energy");
}

}  // namespace am

extern "C" int am_swave_rhs_stage(const am_swave_fields *f, int32_t stage, const am_mesh *mesh, const am_geometry *g,
                                  int32_t push_ghosts, void *stream)
{
    if (!f || !mesh || !g || !mesh->tiles || !mesh->positions)
        return am::set_error(AM_ERR_INVALID, "am_swave_rhs_stage: null argument");
    if (mesh->n_tiles <= 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::swave_stage, f, stage, mesh, g, push_ghosts, stream);
}

extern "C" int am_swave_energy(const am_swave_fields *f, const am_mesh *mesh, const am_geometry *g, double *partial,
                               void *stream)
{
    if (!f || !mesh || !g || !mesh->tiles || !mesh->positions || !partial)
        return am::set_error(AM_ERR_INVALID, "am_swave_energy: null argument");
    if (mesh->n_tiles <= 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::swave_energy, f, mesh, g, partial, stream);
}
