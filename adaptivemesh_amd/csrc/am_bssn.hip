// BSSN right-hand side, fused RK4 stage update and constraint norms for
// gfx950.  The reference stops at metric derivatives and Christoffel symbols
// (BSSN.h:7-102); the equations here are the moving-puncture BSSN system
// (chi form, 1+log or harmonic slicing, Gamma-driver shift with B^i), the
// same system the FP64 oracle states term by term.
//
// Inputs per interior cell: 24 fp32 state values (stage input) and the
// compressed_float3 words written by the FD kernel (a gradient for every
// field, a full Hessian for chi, gt_ij, alpha, beta^i).  The lowered
// symbols Gt_ijk come from the metric gradient words and are raised in
// registers (Gt^i_jk = gt^il Gt_ljk): no Christoffel words are read, so the
// symbols carry one quantisation (that of the gradient words), not two.
//
// One thread per interior cell, 256 cells per workgroup, every array read
// as contiguous runs (words: 1 KB per workgroup per array; state: 16-float
// row segments).  This kernel is HBM-bound (~500-700 B per cell per stage
// against ~1.5 kflop): bulk arithmetic is fp32 with FMA contraction, the
// constraint sums are reduced in FP64.
#include <mutex>

#include "am_bssn_kernels.hpp"

namespace am {

template <int S, int P>
static int launch_stage(const am_bssn_fields *f, int stage, const am_bssn_params *p, const am_mesh *mesh,
                        const am_geometry *g, int push, void *stream)
{
    if (f->word_format == WORD_CF3) return launch_stage_fmt<S, P, WORD_CF3>(f, stage, p, mesh, g, push, stream);
    if (S == 16 && P == 3) return bssn_stage_words(f->word_format, f, stage, p, mesh, g, push, stream);
    return set_error(AM_ERR_INVALID, "am_bssn_rhs_stage: word format %d needs S=16, P=3", f->word_format);
}

template <int S, int P>
static int launch_cons(const am_bssn_fields *f, const am_bssn_params *p, const am_mesh *mesh, double *partial,
                       void *stream)
{
    if (f->word_format == WORD_CF3) return launch_cons_fmt<S, P, WORD_CF3>(f, p, mesh, partial, stream);
    if (S == 16 && P == 3) return bssn_cons_words(f->word_format, f, p, mesh, partial, stream);
    return set_error(AM_ERR_INVALID, "am_bssn_constraints: word format %d needs S=16, P=3", f->word_format);
}

// ------------------------------------------------ allocation-free norms
// am_bssn_constraints_noalloc keeps the per-workgroup partial sums in a
// file-scope device array instead of a caller workspace, and reduces them on
// the device into the caller's four doubles.  Each GPU holds its own copy
// of the array (2 MiB) for the life of the process; every call on that GPU
// shares it, so calls that overlap on one GPU (two streams, two threads)
// race.  Nothing in it needs clearing: each call writes the partials it
// reads.
constexpr int64_t NOALLOC_MAX_BLOCKS = 65536;  // 4096 tiles of 16^3
static __device__ __align__(256) double scratch[4 * NOALLOC_MAX_BLOCKS];

// The device address of `scratch` on the current device, looked up once per
// device.  In host code `scratch` names a host-side stand-in, never the
// device array, so every use goes through this address.
constexpr int MAX_DEVICES = 64;
static std::mutex scratch_mu;
static void *scratch_dev[MAX_DEVICES];

static int scratch_address(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return set_error((int)e, "am_bssn_constraints_noalloc: hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES)
        return set_error(AM_ERR_INVALID, "am_bssn_constraints_noalloc: device %d beyond %d", dev, MAX_DEVICES);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void *ptr = nullptr;
        e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error((int)(e != hipSuccess ? e : hipErrorInvalidSymbol),
                             "am_bssn_constraints_noalloc: hipGetSymbolAddress: %s", hipGetErrorString(e));
        scratch_dev[dev] = ptr;
    }
    *out = static_cast<double *>(scratch_dev[dev]);
    return AM_OK;
}

// out4 = {sum of column 0, 1, 2, max of column 3} over n quadruples, summed
// in a fixed order (strided per thread, then a tree): one workgroup.
__global__ void __launch_bounds__(256) k_reduce_quads(const double *__restrict__ partial, int64_t n,
                                                      double *__restrict__ out4)
{
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 256) {
        const double *q = partial + 4 * i;
        s0 += q[0];
        s1 += q[1];
        s2 += q[2];
        s3 = fmax(s3, q[3]);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s0 += __shfl_down(s0, off, 64);
        s1 += __shfl_down(s1, off, 64);
        s2 += __shfl_down(s2, off, 64);
        s3 = fmax(s3, __shfl_down(s3, off, 64));
    }
    __shared__ double red[4][4];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    if (lane == 0) {
        red[wave][0] = s0; red[wave][1] = s1; red[wave][2] = s2; red[wave][3] = s3;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r0 = 0.0, r1 = 0.0, r2 = 0.0, r3 = 0.0;
        for (int w = 0; w < 4; w++) {
            r0 += red[w][0]; r1 += red[w][1]; r2 += red[w][2]; r3 = fmax(r3, red[w][3]);
        }
        out4[0] = r0; out4[1] = r1; out4[2] = r2; out4[3] = r3;
    }
}

template <int S, int P>
static int cons_noalloc(const am_bssn_fields *f, const am_bssn_params *p, const am_mesh *mesh, double *out4,
                        void *stream)
{
    const int64_t blocks = (int64_t)mesh->n_tiles * ((S * S * S + 255) / 256);
    if (blocks > NOALLOC_MAX_BLOCKS)
        return set_error(AM_ERR_INVALID,
                         "am_bssn_constraints_noalloc: %d tiles need %lld partial sums, the scratch holds %lld",
                         mesh->n_tiles, (long long)blocks, (long long)NOALLOC_MAX_BLOCKS);
    double *ptr = nullptr;
    if (int rc = scratch_address(&ptr)) return rc;
    if (blocks > 0)
        if (int rc = launch_cons<S, P>(f, p, mesh, ptr, stream)) return rc;
    hipLaunchKernelGGL(k_reduce_quads, dim3(1), dim3(256), 0, (hipStream_t)stream, ptr, blocks, out4);
    return check_launch("k_reduce_quads");
}

}  // namespace am

extern "C" int am_bssn_rhs_stage(const am_bssn_fields *f, int32_t stage, const am_bssn_params *p, const am_mesh *mesh,
                                 const am_geometry *g, int32_t push_ghosts, void *stream)
{
    if (!f || !p || !mesh || !g || !mesh->tiles || !mesh->positions)
        return am::set_error(AM_ERR_INVALID, "am_bssn_rhs_stage: null argument");
    if (mesh->n_tiles <= 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_stage, f, stage, p, mesh, g, push_ghosts, stream);
}

extern "C" int am_bssn_constraints(const am_bssn_fields *f, const am_bssn_params *p, const am_mesh *mesh,
                                   const am_geometry *g, double *partial, void *stream)
{
    if (!f || !p || !mesh || !g || !mesh->tiles || !partial)
        return am::set_error(AM_ERR_INVALID, "am_bssn_constraints: null argument");
    if (mesh->n_tiles <= 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_cons, f, p, mesh, partial, stream);
}

extern "C" int64_t am_bssn_constraint_blocks(const am_geometry *g, int32_t n_tiles)
{
    if (!g) return -1;
    const int64_t S = g->size_domain;
    return (int64_t)n_tiles * ((S * S * S + 255) / 256);
}

extern "C" int am_bssn_constraints_noalloc(const am_bssn_fields *f, const am_bssn_params *p, const am_mesh *mesh,
                                           const am_geometry *g, double *out4, void *stream)
{
    if (!f || !p || !mesh || !g || !out4 || (mesh->n_tiles > 0 && !mesh->tiles))
        return am::set_error(AM_ERR_INVALID, "am_bssn_constraints_noalloc: null argument");
    if (mesh->n_tiles < 0) return am::set_error(AM_ERR_INVALID, "am_bssn_constraints_noalloc: negative tile count");
    AM_DISPATCH_TILE(*g, am::cons_noalloc, f, p, mesh, out4, stream);
}
