// Learned per-octant prolongation: training pairs, the MLP
// 1536 -> 128 -> 128 -> 192 (exact GELU) forward and backward on fp32 MFMA,
// MSE loss and Adam, for gfx950.  No counterpart in the reference.
//
// GEMMs use v_mfma_f32_16x16x4_f32: a workgroup of 4 waves computes a 64x64
// tile of C, each wave a 32x32 quarter as 2x2 MFMA blocks; K advances 16 at
// a time through LDS.  Lane l of a 16x16x4 block holds A[l%16][l/16],
// B[l/16][l%16] and rows 4(l/16)..4(l/16)+3 of column l%16 of the result.
// Operand strides are arguments, so X W, dY W^T and X^T dZ are one kernel.
#include "am_common.hpp"
#include "am_device.hpp"

namespace am {

constexpr int MLP_IN = 1536, MLP_HID = 128, MLP_OUT = 192;
constexpr int64_t MLP_W1 = 0, MLP_B1 = MLP_W1 + (int64_t)MLP_IN * MLP_HID, MLP_W2 = MLP_B1 + MLP_HID,
                  MLP_B2 = MLP_W2 + (int64_t)MLP_HID * MLP_HID, MLP_W3 = MLP_B2 + MLP_HID,
                  MLP_B3 = MLP_W3 + (int64_t)MLP_HID * MLP_OUT, MLP_NPARAMS = MLP_B3 + MLP_OUT;

typedef float floatx4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float gelu_f(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z)
{
    return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * expf(-0.5f * z * z) * 0.39894228040143268f;
}

enum { EPI_PLAIN = 0, EPI_GELU = 1, EPI_DGELU = 2 };

// C[m][n] = sum_k A(m,k) B(k,n) (+ bias[n]); A(m,k) = A[m*sam + k*sak],
// B(k,n) = B[k*sbk + n*sbn], C row-major with leading dimension ldc.
//   EPI_PLAIN: C = AB (+bias)
//   EPI_GELU:  C = AB + bias (pre-activation), out2 = gelu(C)
//   EPI_DGELU: C = AB * gelu'(aux[m][n])  (aux row-major, ldc)
// M, N multiples of 64, K a multiple of 16.
template <int EPI>
__global__ void __launch_bounds__(256) k_gemm(const float *__restrict__ A, int64_t sam, int64_t sak,
                                              const float *__restrict__ B, int64_t sbk, int64_t sbn,
                                              float *__restrict__ C, int64_t ldc, int K, const float *__restrict__ bias,
                                              const float *__restrict__ aux, float *__restrict__ out2)
{
    __shared__ float As[64][17];
    __shared__ float Bs[16][65];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int64_t m0 = (int64_t)blockIdx.y * 64, n0 = (int64_t)blockIdx.x * 64;
    floatx4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < K; k0 += 16) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int idx = tid * 4 + q;
            int m, k;
            if (sak == 1) { m = idx >> 4; k = idx & 15; } else { m = idx & 63; k = idx >> 6; }
            As[m][k] = A[(m0 + m) * sam + (int64_t)(k0 + k) * sak];
            int kb, n;
            if (sbn == 1) { kb = idx >> 6; n = idx & 63; } else { kb = idx & 15; n = idx >> 4; }
            Bs[kb][n] = B[(int64_t)(k0 + kb) * sbk + (n0 + n) * sbn];
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
            const int ka = kk * 4 + (lane >> 4);
#pragma unroll
            for (int bi = 0; bi < 2; bi++) {
                const float a = As[wr * 32 + bi * 16 + (lane & 15)][ka];
#pragma unroll
                for (int bj = 0; bj < 2; bj++) {
                    const float b = Bs[ka][wc * 32 + bj * 16 + (lane & 15)];
                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc[bi][bj], 0, 0, 0);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int bi = 0; bi < 2; bi++)
#pragma unroll
        for (int bj = 0; bj < 2; bj++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int64_t row = m0 + wr * 32 + bi * 16 + 4 * (lane >> 4) + r;
                const int64_t col = n0 + wc * 32 + bj * 16 + (lane & 15);
                float c = acc[bi][bj][r];
                if (bias) c += bias[col];
                if (EPI == EPI_DGELU) c *= gelu_d(aux[row * ldc + col]);
                C[row * ldc + col] = c;
                if (EPI == EPI_GELU) out2[row * ldc + col] = gelu_f(c);
            }
}

// bias gradient: out[n] = sum_m dz[m][n].  One workgroup per 64 columns,
// 16 waves: wave w sums rows w, w + 16, ... (lanes on consecutive columns,
// 16 rows in flight), then the 16 partial sums are added in wave order, so
// the result does not depend on scheduling.
__global__ void __launch_bounds__(1024) k_colsum(const float *__restrict__ dz, int M, int N, float *__restrict__ out)
{
    __shared__ float part[16][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int n = blockIdx.x * 64 + lane;
    float s = 0.f;
    if (n < N) {
        int m = wave;
        for (; m + 15 * 16 < M; m += 16 * 16) {
            float v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) v[u] = dz[(int64_t)(m + 16 * u) * N + n];
#pragma unroll
            for (int u = 0; u < 16; u++) s += v[u];
        }
        for (; m < M; m += 16) s += dz[(int64_t)m * N + n];
    }
    part[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && n < N) {
        float t = part[0][lane];
        for (int w = 1; w < 16; w++) t += part[w][lane];
        out[n] = t;
    }
}

// dy = 2 (y - t) / count; per-workgroup FP64 sum of (y - t)^2
__global__ void __launch_bounds__(256) k_mse(const float *__restrict__ y, const float *__restrict__ t, int64_t count,
                                             float *__restrict__ dy, double *__restrict__ partial)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    double e2 = 0.0;
    if (i < count) {
        const float d = y[i] - t[i];
        dy[i] = 2.f * d / (float)count;
        e2 = (double)d * d;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) e2 += __shfl_down(e2, off, 64);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e2;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

__global__ void __launch_bounds__(256) k_adam(float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, int64_t n, float lr,
                                              float b1, float b2, float eps, float c1, float c2)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// 4-point Lagrange weights through nodes -1, 0, 1, 2 at t = -1/4 and +1/4
__device__ constexpr float LAG_Q[2][4] = {{15.f / 128, 135.f / 128, -27.f / 128, 5.f / 128},
                                          {-7.f / 128, 105.f / 128, 35.f / 128, -5.f / 128}};

struct PairFields {
    const float *f[AM_MAX_VARS];
};

// One thread per (sample, field): restricted 4^3 stencil, centre mu and
// spread sigma, normalised stencil, tricubic baseline and normalised
// residual of the 8 fine children.
template <int S, int P>
__global__ void __launch_bounds__(256) k_mlp_pairs(PairFields pf, const int32_t *__restrict__ tiles, int n_tiles,
                                                   int c_lo, int nc, float *__restrict__ X, float *__restrict__ Tg,
                                                   float *__restrict__ mu_out, float *__restrict__ sg_out)
{
    constexpr int T = S + 2 * P, T2 = T * T;
    const int64_t gid = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t per_tile = (int64_t)nc * nc * nc;
    const int64_t sample = gid / 24;
    const int v = (int)(gid - sample * 24);
    if (sample >= per_tile * n_tiles) return;
    const int ti = (int)(sample / per_tile);
    const int c = (int)(sample - (int64_t)ti * per_tile);
    const int cx = c_lo + c % nc, cy = c_lo + (c / nc) % nc, cz = c_lo + c / (nc * nc);
    const float *f = pf.f[v] + (size_t)tiles[ti] * (T * T2);
    float st[64];
#pragma unroll
    for (int k = 0; k < 4; k++)
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int z0 = 2 * (cz - 1 + k) + P, y0 = 2 * (cy - 1 + j) + P, x0 = 2 * (cx - 1 + i) + P;
                const float *q = f + (size_t)z0 * T2 + y0 * T + x0;
                st[k * 16 + j * 4 + i] =
                    ((q[0] + q[1]) + (q[T] + q[T + 1]) + (q[T2] + q[T2 + 1]) + (q[T2 + T] + q[T2 + T + 1])) * 0.125f;
            }
    const float mu = st[21];  // (1, 1, 1)
    float dev = 0.f;
#pragma unroll
    for (int i = 0; i < 64; i++) dev = fmaxf(dev, fabsf(st[i] - mu));
    const float sg = dev + 1e-3f * fabsf(mu) + 1e-7f;
    const float inv = 1.f / sg;
    float *x = X + sample * 1536 + v * 64;
#pragma unroll
    for (int i = 0; i < 64; i++) x[i] = (st[i] - mu) * inv;
    float *t = Tg + sample * 192 + v * 8;
    const float *fc = f + (size_t)(2 * cz + P) * T2 + (2 * cy + P) * T + (2 * cx + P);
#pragma unroll
    for (int ch = 0; ch < 8; ch++) {
        const int jx = ch & 1, jy = (ch >> 1) & 1, jz = ch >> 2;
        float b = 0.f;
#pragma unroll
        for (int k = 0; k < 4; k++)
#pragma unroll
            for (int j = 0; j < 4; j++)
#pragma unroll
                for (int i = 0; i < 4; i++) b += LAG_Q[jz][k] * LAG_Q[jy][j] * LAG_Q[jx][i] * st[k * 16 + j * 4 + i];
        t[ch] = (fc[(size_t)jz * T2 + jy * T + jx] - b) * inv;
    }
    mu_out[sample * 24 + v] = mu;
    sg_out[sample * 24 + v] = sg;
}

static int gemm(const float *A, int64_t sam, int64_t sak, const float *B, int64_t sbk, int64_t sbn, float *C,
                int64_t ldc, int M, int N, int K, const float *bias, int epi, const float *aux, float *out2,
                hipStream_t st)
{
    if (M % 64 || N % 64 || K % 16 || M <= 0 || N <= 0 || K <= 0)
        return set_error(AM_ERR_INVALID, "am_gemm_f32: M, N must be multiples of 64 and K of 16 (%d %d %d)", M, N,
                         K);
    if (epi == EPI_GELU && !out2) return set_error(AM_ERR_INVALID, "am_gemm_f32: GELU epilogue needs out2");
    if (epi == EPI_DGELU && !aux) return set_error(AM_ERR_INVALID, "am_gemm_f32: GELU' epilogue needs aux");
    const dim3 grid(N / 64, M / 64);
    switch (epi) {
    case EPI_GELU:
        hipLaunchKernelGGL(k_gemm<EPI_GELU>, grid, dim3(256), 0, st, A, sam, sak, B, sbk, sbn, C, ldc, K, bias, aux,
                           out2);
        break;
    case EPI_DGELU:
        hipLaunchKernelGGL(k_gemm<EPI_DGELU>, grid, dim3(256), 0, st, A, sam, sak, B, sbk, sbn, C, ldc, K, bias, aux,
                           out2);
        break;
    default:
        hipLaunchKernelGGL(k_gemm<EPI_PLAIN>, grid, dim3(256), 0, st, A, sam, sak, B, sbk, sbn, C, ldc, K, bias, aux,
                           out2);
    }
    return check_launch("k_gemm");
}

template <int S, int P>
static int launch_pairs(const float *const *fine, const int32_t *tiles, int32_t n_tiles, float *x, float *t,
                        float *mu, float *sigma, hipStream_t st)
{
    PairFields pf;
    for (int v = 0; v < 24; v++) {
        if (!fine[v]) return set_error(AM_ERR_INVALID, "am_mlp_pairs: null field %d", v);
        pf.f[v] = fine[v];
    }
    const int c_lo = 0, nc = (S + P - 6) / 2 + 1;
    const int64_t threads = (int64_t)n_tiles * nc * nc * nc * 24;
    hipLaunchKernelGGL((k_mlp_pairs<S, P>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, st, pf, tiles,
                       n_tiles, c_lo, nc, x, t, mu, sigma);
    return check_launch("k_mlp_pairs");
}

}  // namespace am

extern "C" int am_gemm_f32(const float *A, int64_t sam, int64_t sak, const float *B, int64_t sbk, int64_t sbn, float *C,
                           int64_t ldc, int32_t M, int32_t N, int32_t K, const float *bias, int32_t epilogue,
                           const float *aux, float *out2, void *stream)
{
    if (!A || !B || !C) return am::set_error(AM_ERR_INVALID, "am_gemm_f32: null operand");
    return am::gemm(A, sam, sak, B, sbk, sbn, C, ldc, M, N, K, bias, epilogue, aux, out2, (hipStream_t)stream);
}

extern "C" int64_t am_mlp_param_count(void) { return am::MLP_NPARAMS; }

extern "C" int am_mlp_forward(const float *params, const float *x, int32_t batch, am_mlp_work *w, void *stream)
{
    using namespace am;
    if (!params || !x || !w || !w->z1 || !w->h1 || !w->z2 || !w->h2 || !w->y)
        return set_error(AM_ERR_INVALID, "am_mlp_forward: null argument");
    hipStream_t st = (hipStream_t)stream;
    if (int rc = gemm(x, MLP_IN, 1, params + MLP_W1, MLP_HID, 1, w->z1, MLP_HID, batch, MLP_HID, MLP_IN,
                      params + MLP_B1, EPI_GELU, nullptr, w->h1, st))
        return rc;
    if (int rc = gemm(w->h1, MLP_HID, 1, params + MLP_W2, MLP_HID, 1, w->z2, MLP_HID, batch, MLP_HID, MLP_HID,
                      params + MLP_B2, EPI_GELU, nullptr, w->h2, st))
        return rc;
    return gemm(w->h2, MLP_HID, 1, params + MLP_W3, MLP_OUT, 1, w->y, MLP_OUT, batch, MLP_OUT, MLP_HID,
                params + MLP_B3, EPI_PLAIN, nullptr, nullptr, st);
}

extern "C" int am_mlp_backward(const float *params, const float *x, int32_t batch, am_mlp_work *w, float *grads,
                               void *stream)
{
    using namespace am;
    if (!params || !x || !w || !grads || !w->dy || !w->dz2 || !w->dz1)
        return set_error(AM_ERR_INVALID, "am_mlp_backward: null argument");
    hipStream_t st = (hipStream_t)stream;
    // layer 3: dW3 = H2^T dY, db3 = colsum dY, dZ2 = (dY W3^T) * gelu'(Z2)
    if (int rc = gemm(w->h2, 1, MLP_HID, w->dy, MLP_OUT, 1, grads + MLP_W3, MLP_OUT, MLP_HID, MLP_OUT, batch, nullptr,
                      EPI_PLAIN, nullptr, nullptr, st))
        return rc;
    hipLaunchKernelGGL(k_colsum, dim3((MLP_OUT + 63) / 64), dim3(1024), 0, st, w->dy, batch, MLP_OUT,
                       grads + MLP_B3);
    if (int rc = gemm(w->dy, MLP_OUT, 1, params + MLP_W3, 1, MLP_OUT, w->dz2, MLP_HID, batch, MLP_HID, MLP_OUT,
                      nullptr, EPI_DGELU, w->z2, nullptr, st))
        return rc;
    // layer 2
    if (int rc = gemm(w->h1, 1, MLP_HID, w->dz2, MLP_HID, 1, grads + MLP_W2, MLP_HID, MLP_HID, MLP_HID, batch,
                      nullptr, EPI_PLAIN, nullptr, nullptr, st))
        return rc;
    hipLaunchKernelGGL(k_colsum, dim3((MLP_HID + 63) / 64), dim3(1024), 0, st, w->dz2, batch, MLP_HID,
                       grads + MLP_B2);
    if (int rc = gemm(w->dz2, MLP_HID, 1, params + MLP_W2, 1, MLP_HID, w->dz1, MLP_HID, batch, MLP_HID, MLP_HID,
                      nullptr, EPI_DGELU, w->z1, nullptr, st))
        return rc;
    // layer 1
    if (int rc = gemm(x, 1, MLP_IN, w->dz1, MLP_HID, 1, grads + MLP_W1, MLP_HID, MLP_IN, MLP_HID, batch, nullptr,
                      EPI_PLAIN, nullptr, nullptr, st))
        return rc;
    hipLaunchKernelGGL(k_colsum, dim3((MLP_HID + 63) / 64), dim3(1024), 0, st, w->dz1, batch, MLP_HID,
                       grads + MLP_B1);
    return check_launch("k_colsum");
}

extern "C" int64_t am_mlp_mse_blocks(int32_t batch) { return ((int64_t)batch * am::MLP_OUT + 255) / 256; }

extern "C" int am_mlp_mse(const float *y, const float *t, int32_t batch, float *dy, double *partial, void *stream)
{
    if (!y || !t || !dy || !partial || batch <= 0) return am::set_error(AM_ERR_INVALID, "am_mlp_mse: bad argument");
    const int64_t n = (int64_t)batch * am::MLP_OUT;
    hipLaunchKernelGGL(am::k_mse, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, y, t, n, dy,
                       partial);
    return am::check_launch("k_mse");
}

extern "C" int am_mlp_adam(float *params, const float *grads, float *m, float *v, int64_t n, int32_t step, float lr,
                           float beta1, float beta2, float eps, void *stream)
{
    if (!params || !grads || !m || !v || n <= 0 || step < 1)
        return am::set_error(AM_ERR_INVALID, "am_mlp_adam: bad argument");
    const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(am::k_adam, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, params,
                       grads, m, v, n, lr, beta1, beta2, eps, c1, c2);
    return am::check_launch("k_adam");
}

extern "C" int64_t am_mlp_pairs_per_tile(const am_geometry *g)
{
    if (!g) return -1;
    const int64_t nc = (g->size_domain + g->padding_domain - 6) / 2 + 1;
    return nc * nc * nc;
}

extern "C" int am_mlp_pairs(const float *const *fine, const int32_t *tiles, int32_t n_tiles, const am_geometry *g,
                            float *x, float *t, float *mu, float *sigma, void *stream)
{
    if (!fine || !tiles || !g || !x || !t || !mu || !sigma || n_tiles < 0)
        return am::set_error(AM_ERR_INVALID, "am_mlp_pairs: null argument");
    if (n_tiles == 0) return AM_OK;
    AM_DISPATCH_TILE(*g, am::launch_pairs, fine, tiles, n_tiles, x, t, mu, sigma, (hipStream_t)stream);
}
