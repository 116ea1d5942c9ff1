// Pointwise interior kernels for gfx950 that follow reference kernels
// bit for bit: the SimpleWave Heun stages (SimpleWave.h:7-46), the metric
// trace of a Hessian (BSSN.h:7-22) and the Christoffel symbols
// (BSSN.h:55-91).  One workgroup = one (tile, z-plane); its S*S threads read
// S*S contiguous interior words and the matching padded-tile row segments.
#include "am_device.hpp"

#pragma clang fp contract(off)

namespace am {

struct PointArgs {
    const uint32_t *positions;
    const int32_t *tiles;
    uint32_t substep;
    uint32_t max_depth;
};

// __predictor_step / __corrector_step (SimpleWave.h:7-46)
template <int S, int P, bool CORRECTOR>
__global__ void __launch_bounds__(S *S) k_wave_stage(const float *__restrict__ uold, const uint32_t *__restrict__ d1,
                                                     const uint32_t *__restrict__ d2, float *__restrict__ unew,
                                                     PointArgs m, float outer_dt)
{
    constexpr int T = S + 2 * P;
    const int ti = blockIdx.x / S, z = blockIdx.x - ti * S;
    const int node = m.tiles[ti];
    const int depth = pos_depth(m.positions[node]);
    if (depth == -1 || !active_depth(m.substep, (uint32_t)depth, m.max_depth)) return;
    const int x = threadIdx.x % S, y = threadIdx.x / S;
    const size_t i = ((size_t)node * S + z) * (S * S) + threadIdx.x;
    const float nu = 0.01f;
    const float xd = cf3_unpack(d1[i]).x;
    const float3 h = cf3_unpack(d2[i]);
    const float lap = h.x * nu + h.y * nu + h.z * nu;
    const size_t r = (((size_t)node * T + z) * T + y) * T + x + (size_t)P * (1 + T + T * T);
    float u = uold[r];
    const float div = (float)(1u << depth);
    if (!CORRECTOR) {
        unew[r] = u + (lap - u * xd) * outer_dt / div;
    } else {
        u = u + (lap - u * xd) * outer_dt / div;
        unew[r] = (unew[r] + u) * .5f;
    }
}

template <int S, int P>
static int launch_wave(const float *uold, const uint32_t *d1, const uint32_t *d2, float *unew, const am_mesh *mesh,
                       const am_geometry *g, bool corrector, void *stream)
{
    PointArgs m{mesh->positions, mesh->tiles, mesh->substep, mesh->max_depth};
    dim3 grid(mesh->n_tiles * S);
    if (corrector)
        hipLaunchKernelGGL((k_wave_stage<S, P, true>), grid, dim3(S * S), 0, (hipStream_t)stream, uold, d1, d2, unew,
                           m, g->outer_dt);
    else
        hipLaunchKernelGGL((k_wave_stage<S, P, false>), grid, dim3(S * S), 0, (hipStream_t)stream, uold, d1, d2, unew,
                           m, g->outer_dt);
    return check_launch("k_wave_stage");
}

// float3x3_sym::inverse (tensors.h:208-225)
__device__ __forceinline__ void sym_inverse(const float *m, float *r)
{
    const float xx = m[0], xy = m[1], xz = m[2], yy = m[3], yz = m[4], zz = m[5];
    const float det = -xz * yy * xz + xy * yz * xz + xz * xy * yz - xx * yz * yz - xy * xy * zz + xx * yy * zz;
    r[0] = -yz * yz + yy * zz;
    r[1] = yz * xz - xy * zz;
    r[2] = -yy * xz + xy * yz;
    r[3] = -xz * xz + xx * zz;
    r[4] = xy * xz - xx * yz;
    r[5] = -xy * xy + xx * yy;
#pragma unroll
    for (int k = 0; k < 6; k++) r[k] = r[k] / det;
}

struct Metric6 {
    const float *c[6];
};
struct Words6 {
    const uint32_t *c[6];
};
struct OutWords6 {
    uint32_t *c[6];
};

// __deriv_trace (BSSN.h:7-22)
template <int S, int P>
__global__ void __launch_bounds__(S *S) k_deriv_trace(const uint32_t *__restrict__ hd, const uint32_t *__restrict__ ho,
                                                      Metric6 met, float *__restrict__ out, PointArgs m)
{
    constexpr int T = S + 2 * P;
    const int ti = blockIdx.x / S, z = blockIdx.x - ti * S;
    const int node = m.tiles[ti];
    const int depth = pos_depth(m.positions[node]);
    if (depth == -1 || !active_depth(m.substep, (uint32_t)depth, m.max_depth)) return;
    const int x = threadIdx.x % S, y = threadIdx.x / S;
    const size_t r = (((size_t)node * T + z) * T + y) * T + x + (size_t)P * (1 + T + T * T);
    float g[6], inv[6];
#pragma unroll
    for (int k = 0; k < 6; k++) g[k] = met.c[k][r];
    sym_inverse(g, inv);
    const size_t i = ((size_t)node * S + z) * (S * S) + threadIdx.x;
    const float3 d = cf3_unpack(hd[i]), o = cf3_unpack(ho[i]);
    out[i] = d.x * inv[0] + d.y * inv[3] + d.z * inv[5] + (o.x * inv[1] + o.y * inv[2] + o.z * inv[4]) * 2.f;
}

// __compute_raised_christoffel_symbols (BSSN.h:55-91)
template <int S, int P, bool COMPAT>
__global__ void __launch_bounds__(S *S) k_christoffel(Metric6 met, Words6 dg, OutWords6 out, PointArgs m, int32_t gs)
{
    constexpr int T = S + 2 * P;
    const int ti = blockIdx.x / S, z = blockIdx.x - ti * S;
    const int node = m.tiles[ti];
    const int depth = pos_depth(m.positions[node]);
    if (depth == -1 || !active_depth(m.substep, (uint32_t)depth, m.max_depth)) return;
    const int x = threadIdx.x % S, y = threadIdx.x / S;
    int32_t rng = thread_seed(gs, (uint32_t)x, (uint32_t)y, (uint32_t)(node * S + z));
    const size_t i = ((size_t)node * S + z) * (S * S) + threadIdx.x;
    float v[6][3];
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const float3 t = cf3_unpack(dg.c[k][i]);
        v[k][0] = t.x; v[k][1] = t.y; v[k][2] = t.z;
    }
    // rows of the three float3x3_sym (tensors.h:138): X_l built from dgamma_{la}
    constexpr int RW[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    float X[3][6];
#pragma unroll
    for (int l = 0; l < 3; l++) {
        const float *a = v[RW[l][0]], *b = v[RW[l][1]], *c = v[RW[l][2]];
        X[l][0] = a[0];
        X[l][1] = (a[1] + b[0]) * .5f;
        X[l][2] = (a[2] + c[0]) * .5f;
        X[l][3] = b[1];
        X[l][4] = (c[1] + b[2]) * .5f;
        X[l][5] = c[2];
    }
    const size_t r = (((size_t)node * T + z) * T + y) * T + x + (size_t)P * (1 + T + T * T);
    float g[6], im[6];
#pragma unroll
    for (int k = 0; k < 6; k++) g[k] = met.c[k][r];
    if (COMPAT) {
#pragma unroll
        for (int k = 0; k < 6; k++) im[k] = g[k];
    } else {
        sym_inverse(g, im);
    }
#pragma unroll
    for (int jk = 0; jk < 6; jk++) {
        const float a0 = X[0][jk] - v[jk][0] * .5f;
        const float a1 = X[1][jk] - v[jk][1] * .5f;
        const float a2 = X[2][jk] - v[jk][2] * .5f;
        const float o0 = a0 * im[0] + a1 * im[1] + a2 * im[2];
        const float o1 = a0 * im[1] + a1 * im[3] + a2 * im[4];
        const float o2 = a0 * im[2] + a1 * im[4] + a2 * im[5];
        out.c[jk][i] = cf3_pack(o0, o1, o2, rng);
    }
}

template <int S, int P>
static int launch_trace(const uint32_t *hd, const uint32_t *ho, const float *const *met6, float *out,
                        const am_mesh *mesh, void *stream)
{
    Metric6 met;
    for (int k = 0; k < 6; k++) met.c[k] = met6[k];
    PointArgs m{mesh->positions, mesh->tiles, mesh->substep, mesh->max_depth};
    hipLaunchKernelGGL((k_deriv_trace<S, P>), dim3(mesh->n_tiles * S), dim3(S * S), 0, (hipStream_t)stream, hd, ho,
                       met, out, m);
    return check_launch("k_deriv_trace");
}

template <int S, int P>
static int launch_christoffel(const float *const *met6, const uint32_t *const *dg6, uint32_t *const *out6, int compat,
                              const am_mesh *mesh, int32_t seed, void *stream)
{
    Metric6 met;
    Words6 dg;
    OutWords6 out;
    for (int k = 0; k < 6; k++) {
        met.c[k] = met6[k];
        dg.c[k] = dg6[k];
        out.c[k] = out6[k];
    }
    PointArgs m{mesh->positions, mesh->tiles, mesh->substep, mesh->max_depth};
    if (compat)
        hipLaunchKernelGGL((k_christoffel<S, P, true>), dim3(mesh->n_tiles * S), dim3(S * S), 0, (hipStream_t)stream,
                           met, dg, out, m, seed);
    else
        hipLaunchKernelGGL((k_christoffel<S, P, false>), dim3(mesh->n_tiles * S), dim3(S * S), 0, (hipStream_t)stream,
                           met, dg, out, m, seed);
    return check_launch("k_christoffel");
}

static int point_checks(const am_mesh *mesh, const am_geometry *g)
{
    if (!mesh || !g) return set_error(AM_ERR_INVALID, "null mesh or geometry");
    if (!mesh->tiles || !mesh->positions) return set_error(AM_ERR_INVALID, "mesh without tiles or positions");
    if (mesh->n_tiles < 0) return set_error(AM_ERR_INVALID, "negative tile count");
    return AM_OK;
}

}  // namespace am

extern "C" int am_wave_predictor(const float *old_, const uint32_t *d1, const uint32_t *d2diag, float *new_,
                                 const am_mesh *mesh, const am_geometry *g, void *stream)
{
    if (int rc = am::point_checks(mesh, g)) return rc;
    if (!old_ || !d1 || !d2diag || !new_) return am::set_error(AM_ERR_INVALID, "am_wave_predictor: null buffer");
    if (mesh->n_tiles == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_wave, old_, d1, d2diag, new_, mesh, g, false, stream);
}

extern "C" int am_wave_corrector(const float *old_, const uint32_t *d1, const uint32_t *d2diag, float *new_,
                                 const am_mesh *mesh, const am_geometry *g, void *stream)
{
    if (int rc = am::point_checks(mesh, g)) return rc;
    if (!old_ || !d1 || !d2diag || !new_) return am::set_error(AM_ERR_INVALID, "am_wave_corrector: null buffer");
    if (mesh->n_tiles == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_wave, old_, d1, d2diag, new_, mesh, g, true, stream);
}

extern "C" int am_deriv_trace(const uint32_t *hdiag, const uint32_t *hoff, const float *const *metric6, float *out,
                              const am_mesh *mesh, const am_geometry *g, void *stream)
{
    if (int rc = am::point_checks(mesh, g)) return rc;
    if (!hdiag || !hoff || !metric6 || !out) return am::set_error(AM_ERR_INVALID, "am_deriv_trace: null buffer");
    for (int k = 0; k < 6; k++)
        if (!metric6[k]) return am::set_error(AM_ERR_INVALID, "am_deriv_trace: null metric component %d", k);
    if (mesh->n_tiles == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_trace, hdiag, hoff, metric6, out, mesh, stream);
}

extern "C" int am_christoffel(const float *const *metric6, const uint32_t *const *dmetric6, uint32_t *const *out6,
                              int32_t compat, const am_mesh *mesh, const am_geometry *g, int32_t seed, void *stream)
{
    if (int rc = am::point_checks(mesh, g)) return rc;
    if (!metric6 || !dmetric6 || !out6) return am::set_error(AM_ERR_INVALID, "am_christoffel: null table");
    for (int k = 0; k < 6; k++)
        if (!metric6[k] || !dmetric6[k] || !out6[k])
            return am::set_error(AM_ERR_INVALID, "am_christoffel: null component %d", k);
    if (mesh->n_tiles == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_christoffel, metric6, dmetric6, out6, compat, mesh, seed, stream);
}
