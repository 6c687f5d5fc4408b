/*
 * mlp.hip — the learned control variate: an MLP 2 -> 64 -> 64 -> 64 -> 3
 * (tanh; outputs u, du/dx, du/dy; 8707 parameters) trained on plain-MC node
 * estimates, and its BF16 MFMA inference.
 *
 * The reference's control-variate hook is the plugin F = -sigma^T grad u used in
 * xi += Y F(X).dW (Solvers/FeynmanKacSolver.cpp:13, BVPs/Franke_Poisson.hpp:67-73);
 * it needs grad u, unknown in practice.  The network learns u from node
 * estimates and its two gradient heads learn d u_hat / d(x, y) (forward-mode
 * targets, held fixed per step), so one forward pass gives the field F.
 *
 * Training (FP32, deterministic: every sum is sequential in a fixed order):
 *   input scaling -> dense+tanh x3 with forward-mode tangents -> output + loss
 *   gradient -> back-propagation -> weight gradients -> Adam.
 * Inference (BF16 operands, FP32 accumulate): each wave evaluates 32 points;
 * hidden layers are Act^T = W H^T on v_mfma_f32_32x32x8bf16_1k with the points
 * on the N axis, so a layer's accumulator layout is the next layer's B operand
 * and no LDS transposes are needed.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

#include "../../include/spac_mlp.h"
#include "fk_internal.h"

namespace spac {

#define MLP_HIP_CHECK(expr)                                                                    \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return set_error(SPAC_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                             __FILE__, __LINE__);                                              \
    } while (0)

constexpr int kH = 64;
/* parameter offsets */
constexpr int oW1 = 0, ob1 = oW1 + kH * 2, oW2 = ob1 + kH, ob2 = oW2 + kH * kH, oW3 = ob2 + kH,
              ob3 = oW3 + kH * kH, oW4 = ob3 + kH, ob4 = oW4 + 3 * kH, kNParams = ob4 + 3;
static_assert(kNParams == SPAC_MLP_N_PARAMS, "parameter count");

/* ------------------------------------------------------------ training -- */
__global__ void k_input(const double* x, const double* y, int n, float lox, float loy, float sx, float sy,
                        float* z) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    z[s] = 2.0f * ((float)x[s] - lox) * sx - 1.0f;
    z[n + s] = 2.0f * ((float)y[s] - loy) * sy - 1.0f;
}

/* out[o][s] = tanh(sum_i W[o][i] in[i][s] + b[o]) */
__global__ void k_dense_tanh(const float* W, const float* b, const float* in, int in_dim, float* out, int out_dim,
                             int n) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int o = blockIdx.y;
    if (s >= n || o >= out_dim) return;
    float a = 0.0f;
    for (int i = 0; i < in_dim; ++i) a = fmaf(W[o * in_dim + i], in[(size_t)i * n + s], a);
    out[(size_t)o * n + s] = tanhf(a + b[o]);
}

/* tangent: tout[o][s] = (1 - h[o][s]^2) sum_i W[o][i] tin[i][s]; layer 1: tin = scale e_d */
__global__ void k_tangent(const float* W, const float* tin, int in_dim, const float* h, float* tout, int out_dim,
                          int n, int first_dir, float first_scale) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int o = blockIdx.y;
    if (s >= n || o >= out_dim) return;
    float a;
    if (first_dir >= 0) {
        a = W[o * in_dim + first_dir] * first_scale;
    } else {
        a = 0.0f;
        for (int i = 0; i < in_dim; ++i) a = fmaf(W[o * in_dim + i], tin[(size_t)i * n + s], a);
    }
    float hv = h[(size_t)o * n + s];
    tout[(size_t)o * n + s] = (1.0f - hv * hv) * a;
}

/* outputs, forward-mode gradient of o0, residuals and dL/do */
__global__ void k_output(const float* W4, const float* b4, const float* h3, const float* t3x, const float* t3y,
                         const double* u, int n, float lam, float* out3, float* dO, float* loss_s) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    float o[3], gx = 0.0f, gy = 0.0f;
    for (int c = 0; c < 3; ++c) o[c] = 0.0f;
    for (int i = 0; i < kH; ++i) {
        float hv = h3[(size_t)i * n + s];
        for (int c = 0; c < 3; ++c) o[c] = fmaf(W4[c * kH + i], hv, o[c]);
        gx = fmaf(W4[i], t3x[(size_t)i * n + s], gx);
        gy = fmaf(W4[i], t3y[(size_t)i * n + s], gy);
    }
    for (int c = 0; c < 3; ++c) o[c] += b4[c];
    for (int c = 0; c < 3; ++c) out3[(size_t)c * n + s] = o[c];
    if (!u) return;
    float r0 = o[0] - (float)u[s], r1 = o[1] - gx, r2 = o[2] - gy;
    float inv = 1.0f / (float)n;
    dO[s] = 2.0f * r0 * inv;
    dO[(size_t)n + s] = 2.0f * lam * r1 * inv;
    dO[(size_t)2 * n + s] = 2.0f * lam * r2 * inv;
    loss_s[s] = r0 * r0 + lam * (r1 * r1 + r2 * r2);
}

/* dout[i][s] = (1 - h[i][s]^2) sum_o W[o][i] dn[o][s] */
__global__ void k_backprop(const float* W, const float* dn, int out_dim, const float* h, float* dout, int in_dim,
                           int n) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int i = blockIdx.y;
    if (s >= n || i >= in_dim) return;
    float a = 0.0f;
    for (int o = 0; o < out_dim; ++o) a = fmaf(W[o * in_dim + i], dn[(size_t)o * n + s], a);
    float hv = h[(size_t)i * n + s];
    dout[(size_t)i * n + s] = (1.0f - hv * hv) * a;
}

/* dW[o][i] = sum_s d[o][s] x[i][s]; db[o] = sum_s d[o][s] (i == in_dim) */
__global__ void k_weight_grad(const float* d, const float* xin, int out_dim, int in_dim, int n, float* dW, float* db) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    int per = in_dim + 1;
    if (t >= out_dim * per) return;
    int o = t / per, i = t % per;
    float a = 0.0f;
    if (i < in_dim) {
        for (int s = 0; s < n; ++s) a = fmaf(d[(size_t)o * n + s], xin[(size_t)i * n + s], a);
        dW[o * in_dim + i] = a;
    } else {
        for (int s = 0; s < n; ++s) a += d[(size_t)o * n + s];
        db[o] = a;
    }
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    float gk = g[k];
    float mk = b1 * m[k] + (1.0f - b1) * gk;
    float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

/* ----------------------------------------------------- MFMA inference -- */
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ short f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u); /* round to nearest even */
    return (short)(u >> 16);
}
__device__ __forceinline__ float fast_tanh(float x) {
    /* tanh x = 1 - 2 / (e^{2x} + 1) with the hardware 2^x and reciprocal */
    float e = __builtin_amdgcn_exp2f(2.8853900817779268f * x);
    return 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
}

/* Hidden layer on MFMA: acc[mb] (features 32 mb + row(j), point l%32) =
 * sum_k W[f][k] H[k][point].  h[mb][j] holds feature 32 mb + 8 (j/4) + 4 (lane/32)
 * + (j%4) of the lane's point, which is the B operand layout of the next layer. */
__device__ __forceinline__ void mfma_layer(const short* Wbf /* [64][64] bf16 */, const float* bias,
                                           const float (&hin)[2][16], float (&hout)[2][16], int lane) {
    const int half = lane >> 5, col = lane & 31;
    f32x16 acc[2];
    for (int mb = 0; mb < 2; ++mb)
        for (int j = 0; j < 16; ++j) acc[mb][j] = 0.0f;
#pragma unroll
    for (int kb = 0; kb < 8; ++kb) {
        /* B: k = 8 kb + 4 half + j, from hin[kb / 4][4 (kb % 4) + j] */
        bf16x4 bv;
        for (int j = 0; j < 4; ++j) bv[j] = f2bf(hin[kb >> 2][4 * (kb & 3) + j]);
#pragma unroll
        for (int mb = 0; mb < 2; ++mb) {
            /* A: W[m = 32 mb + col][k = 8 kb + 4 half + j] */
            const short* wr = Wbf + (32 * mb + col) * kH + 8 * kb + 4 * half;
            bf16x4 av;
            for (int j = 0; j < 4; ++j) av[j] = wr[j];
            acc[mb] = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(av, bv, acc[mb], 0, 0, 0);
        }
    }
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
        for (int j = 0; j < 16; ++j) {
            int f = 32 * mb + 8 * (j >> 2) + 4 * half + (j & 3);
            hout[mb][j] = fast_tanh(acc[mb][j] + bias[f]);
        }
}

struct InferArgs {
    const float* params;
    const short* W2bf;
    const short* W3bf;
    float lox, loy, sx, sy;
    const double* x; /* points (or nullptr: a T x T grid over the domain) */
    const double* y;
    int n;
    int grid_t;
    float* out; /* n x 4: u, gx, gy, 0 */
};

__global__ void __launch_bounds__(256) k_mlp_infer_mfma(InferArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int half = lane >> 5, col = lane & 31;
    const int pt = wave * 32 + col;
    const float* P = a.params;
    float px = 0.0f, py = 0.0f;
    if (pt < a.n) {
        if (a.x) {
            px = (float)a.x[pt];
            py = (float)a.y[pt];
        } else {
            int ix = pt % a.grid_t, iy = pt / a.grid_t;
            px = a.lox + (float)ix / ((float)(a.grid_t - 1) * a.sx);
            py = a.loy + (float)iy / ((float)(a.grid_t - 1) * a.sy);
        }
    }
    const float z0 = 2.0f * (px - a.lox) * a.sx - 1.0f, z1 = 2.0f * (py - a.loy) * a.sy - 1.0f;
    float h1[2][16], h2[2][16], h3[2][16];
    /* layer 1 (K = 2) on the VALU, directly in the B layout */
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
        for (int j = 0; j < 16; ++j) {
            int f = 32 * mb + 8 * (j >> 2) + 4 * half + (j & 3);
            h1[mb][j] = fast_tanh(fmaf(P[oW1 + 2 * f], z0, fmaf(P[oW1 + 2 * f + 1], z1, P[ob1 + f])));
        }
    mfma_layer(a.W2bf, P + ob2, h1, h2, lane);
    mfma_layer(a.W3bf, P + ob3, h2, h3, lane);
    /* output layer: each lane holds 32 of the 64 features of its point */
    float o[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int mb = 0; mb < 2; ++mb)
        for (int j = 0; j < 16; ++j) {
            int f = 32 * mb + 8 * (j >> 2) + 4 * half + (j & 3);
            for (int c = 0; c < 3; ++c) o[c] = fmaf(P[oW4 + c * kH + f], h3[mb][j], o[c]);
        }
    for (int c = 0; c < 3; ++c) o[c] += __shfl_xor(o[c], 32, 64);
    if (half == 0 && pt < a.n) {
        float4 r = make_float4(o[0] + P[ob4], o[1] + P[ob4 + 1], o[2] + P[ob4 + 2], 0.0f);
        reinterpret_cast<float4*>(a.out)[pt] = r;
    }
}

__global__ void k_to_bf16(const float* in, short* out, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = f2bf(in[k]);
}

/* --------------------------------------------------------------- handle -- */
struct Mlp {
    spac_mlp_desc desc{};
    hipStream_t stream = nullptr;
    DeviceBuffer params, grads, am, av, wbf, act, pts, tgt, table;
    int64_t cap_n = 0;
    int adam_t = 0;
    int table_t = 0;
    float lox = 0, loy = 0, sx = 1, sy = 1;
};

static spac_status mlp_prepare_batch(Mlp* m, const double* x, const double* y, const double* u, int n) {
    spac_status rc;
    size_t N = (size_t)n;
    /* act: z(2) h1 h2 h3 (3x64) t1..t3 x 2 dirs (6x64) d3 d2 d1 (3x64) out3 dO(3) loss(1) */
    size_t floats = N * (2 + 3 * kH + 6 * kH + 3 * kH + 3 + 3 + 1);
    if ((rc = m->act.ensure(floats * 4))) return rc;
    if ((rc = m->pts.ensure(N * 16))) return rc;
    if ((rc = m->tgt.ensure(N * 8))) return rc;
    MLP_HIP_CHECK(hipMemcpyAsync(m->pts.ptr, x, N * 8, hipMemcpyHostToDevice, m->stream));
    MLP_HIP_CHECK(hipMemcpyAsync((char*)m->pts.ptr + N * 8, y, N * 8, hipMemcpyHostToDevice, m->stream));
    if (u) MLP_HIP_CHECK(hipMemcpyAsync(m->tgt.ptr, u, N * 8, hipMemcpyHostToDevice, m->stream));
    return SPAC_OK;
}

/* forward + backward over the prepared batch: grads and per-sample loss on device */
static spac_status mlp_fwd_bwd(Mlp* m, int n, float lam, bool want_grad) {
    hipStream_t s = m->stream;
    const size_t N = (size_t)n;
    float* A = (float*)m->act.ptr;
    float* z = A;
    float* h1 = z + 2 * N;
    float* h2 = h1 + kH * N;
    float* h3 = h2 + kH * N;
    float* t1[2] = {h3 + kH * N, h3 + 2 * kH * N};
    float* t2[2] = {h3 + 3 * kH * N, h3 + 4 * kH * N};
    float* t3[2] = {h3 + 5 * kH * N, h3 + 6 * kH * N};
    float* d3 = h3 + 7 * kH * N;
    float* d2 = d3 + kH * N;
    float* d1 = d2 + kH * N;
    float* out3 = d1 + kH * N;
    float* dO = out3 + 3 * N;
    float* loss_s = dO + 3 * N;
    const float* P = (const float*)m->params.ptr;
    float* Gd = (float*)m->grads.ptr;
    const double* x = (const double*)m->pts.ptr;
    const double* y = x + N;
    dim3 blk(256), g1((n + 255) / 256), gH((n + 255) / 256, kH);
    hipLaunchKernelGGL(k_input, g1, blk, 0, s, x, y, n, m->lox, m->loy, m->sx, m->sy, z);
    hipLaunchKernelGGL(k_dense_tanh, gH, blk, 0, s, P + oW1, P + ob1, z, 2, h1, kH, n);
    hipLaunchKernelGGL(k_dense_tanh, gH, blk, 0, s, P + oW2, P + ob2, h1, kH, h2, kH, n);
    hipLaunchKernelGGL(k_dense_tanh, gH, blk, 0, s, P + oW3, P + ob3, h2, kH, h3, kH, n);
    const float sc[2] = {2.0f * m->sx, 2.0f * m->sy};
    for (int d = 0; d < 2; ++d) {
        hipLaunchKernelGGL(k_tangent, gH, blk, 0, s, P + oW1, (const float*)nullptr, 2, h1, t1[d], kH, n, d, sc[d]);
        hipLaunchKernelGGL(k_tangent, gH, blk, 0, s, P + oW2, t1[d], kH, h2, t2[d], kH, n, -1, 0.0f);
        hipLaunchKernelGGL(k_tangent, gH, blk, 0, s, P + oW3, t2[d], kH, h3, t3[d], kH, n, -1, 0.0f);
    }
    hipLaunchKernelGGL(k_output, g1, blk, 0, s, P + oW4, P + ob4, h3, t3[0], t3[1], (const double*)m->tgt.ptr, n, lam,
                       out3, dO, loss_s);
    if (!want_grad) return SPAC_OK;
    auto wgrad = [&](const float* d, const float* xin, int od, int id, int oW, int ob) {
        int threads = od * (id + 1);
        hipLaunchKernelGGL(k_weight_grad, dim3((threads + 127) / 128), dim3(128), 0, s, d, xin, od, id, n, Gd + oW,
                           Gd + ob);
    };
    wgrad(dO, h3, 3, kH, oW4, ob4);
    hipLaunchKernelGGL(k_backprop, gH, blk, 0, s, P + oW4, dO, 3, h3, d3, kH, n);
    wgrad(d3, h2, kH, kH, oW3, ob3);
    hipLaunchKernelGGL(k_backprop, gH, blk, 0, s, P + oW3, d3, kH, h2, d2, kH, n);
    wgrad(d2, h1, kH, kH, oW2, ob2);
    hipLaunchKernelGGL(k_backprop, gH, blk, 0, s, P + oW2, d2, kH, h1, d1, kH, n);
    wgrad(d1, z, kH, 2, oW1, ob1);
    MLP_HIP_CHECK(hipGetLastError());
    return SPAC_OK;
}

static spac_status mlp_loss_host(Mlp* m, int n, double* loss) {
    const size_t N = (size_t)n;
    std::vector<float> ls(N);
    float* A = (float*)m->act.ptr;
    float* loss_s = A + N * (2 + 3 * kH + 6 * kH + 3 * kH + 3 + 3);
    MLP_HIP_CHECK(hipMemcpyAsync(ls.data(), loss_s, N * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(m->stream));
    double acc = 0.0;
    for (float v : ls) acc += v;
    *loss = acc / (double)n;
    return SPAC_OK;
}

static spac_status mlp_refresh_bf16(Mlp* m) {
    spac_status rc;
    if ((rc = m->wbf.ensure(2 * kH * kH * sizeof(short)))) return rc;
    const float* P = (const float*)m->params.ptr;
    short* w = (short*)m->wbf.ptr;
    hipLaunchKernelGGL(k_to_bf16, dim3(16), dim3(256), 0, m->stream, P + oW2, w, kH * kH);
    hipLaunchKernelGGL(k_to_bf16, dim3(16), dim3(256), 0, m->stream, P + oW3, w + kH * kH, kH * kH);
    MLP_HIP_CHECK(hipGetLastError());
    return SPAC_OK;
}

static spac_status mlp_infer_dev(Mlp* m, const double* dx, const double* dy, int n, int grid_t, float* dout) {
    spac_status rc;
    if ((rc = mlp_refresh_bf16(m))) return rc;
    InferArgs a;
    a.params = (const float*)m->params.ptr;
    a.W2bf = (const short*)m->wbf.ptr;
    a.W3bf = a.W2bf + kH * kH;
    a.lox = m->lox; a.loy = m->loy; a.sx = m->sx; a.sy = m->sy;
    a.x = dx; a.y = dy; a.n = n; a.grid_t = grid_t; a.out = dout;
    int waves = (n + 31) / 32;
    int blocks = (waves + 3) / 4;
    hipLaunchKernelGGL(k_mlp_infer_mfma, dim3(blocks), dim3(256), 0, m->stream, a);
    MLP_HIP_CHECK(hipGetLastError());
    return SPAC_OK;
}

} // namespace spac

using namespace spac;

extern "C" {

spac_status spac_mlp_create(const spac_mlp_desc* d, spac_mlp** out) {
    if (!d || !out) return set_error(SPAC_E_INVALID_ARG, "null argument");
    *out = nullptr;
    if (!(d->domain[2] > d->domain[0]) || !(d->domain[3] > d->domain[1]))
        return set_error(SPAC_E_INVALID_ARG, "MLP domain must have xmax > xmin and ymax > ymin");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(SPAC_E_NO_DEVICE, "no HIP device visible");
    Mlp* m = new Mlp();
    m->desc = *d;
    m->lox = (float)d->domain[0];
    m->loy = (float)d->domain[1];
    m->sx = (float)(1.0 / (d->domain[2] - d->domain[0]));
    m->sy = (float)(1.0 / (d->domain[3] - d->domain[1]));
    MLP_HIP_CHECK(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    spac_status rc;
    if ((rc = m->params.ensure(kNParams * 4)) || (rc = m->grads.ensure(kNParams * 4)) ||
        (rc = m->am.ensure(kNParams * 4)) || (rc = m->av.ensure(kNParams * 4))) {
        delete m;
        return rc;
    }
    /* Glorot-uniform weights, zero biases */
    std::vector<float> p(kNParams, 0.0f);
    std::mt19937_64 g(d->seed);
    auto fill = [&](int off, int fo, int fi) {
        float lim = std::sqrt(6.0f / (float)(fo + fi));
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int k = 0; k < fo * fi; ++k) p[off + k] = U(g);
    };
    fill(oW1, kH, 2);
    fill(oW2, kH, kH);
    fill(oW3, kH, kH);
    fill(oW4, 3, kH);
    MLP_HIP_CHECK(hipMemcpy(m->params.ptr, p.data(), kNParams * 4, hipMemcpyHostToDevice));
    MLP_HIP_CHECK(hipMemset(m->am.ptr, 0, kNParams * 4));
    MLP_HIP_CHECK(hipMemset(m->av.ptr, 0, kNParams * 4));
    *out = reinterpret_cast<spac_mlp*>(m);
    return SPAC_OK;
}

spac_status spac_mlp_destroy(spac_mlp* hm) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m) return SPAC_OK;
    (void)hipStreamSynchronize(m->stream);
    for (DeviceBuffer* b : {&m->params, &m->grads, &m->am, &m->av, &m->wbf, &m->act, &m->pts, &m->tgt, &m->table})
        b->release();
    (void)hipStreamDestroy(m->stream);
    delete m;
    return SPAC_OK;
}

spac_status spac_mlp_get_params(spac_mlp* hm, float* out) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || !out) return set_error(SPAC_E_INVALID_ARG, "null argument");
    MLP_HIP_CHECK(hipMemcpy(out, m->params.ptr, kNParams * 4, hipMemcpyDeviceToHost));
    return SPAC_OK;
}

spac_status spac_mlp_set_params(spac_mlp* hm, const float* in) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || !in) return set_error(SPAC_E_INVALID_ARG, "null argument");
    MLP_HIP_CHECK(hipMemcpy(m->params.ptr, in, kNParams * 4, hipMemcpyHostToDevice));
    m->adam_t = 0;
    MLP_HIP_CHECK(hipMemset(m->am.ptr, 0, kNParams * 4));
    MLP_HIP_CHECK(hipMemset(m->av.ptr, 0, kNParams * 4));
    return SPAC_OK;
}

spac_status spac_mlp_loss_grad(spac_mlp* hm, const double* x, const double* y, const double* u, int32_t n, double lam,
                               float* grad, double* loss) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || !x || !y || !u || n < 1) return set_error(SPAC_E_INVALID_ARG, "bad arguments");
    spac_status rc;
    if ((rc = mlp_prepare_batch(m, x, y, u, n))) return rc;
    if ((rc = mlp_fwd_bwd(m, n, (float)lam, grad != nullptr))) return rc;
    if (loss && (rc = mlp_loss_host(m, n, loss))) return rc;
    if (grad) MLP_HIP_CHECK(hipMemcpy(grad, m->grads.ptr, kNParams * 4, hipMemcpyDeviceToHost));
    MLP_HIP_CHECK(hipStreamSynchronize(m->stream));
    return SPAC_OK;
}

spac_status spac_mlp_train(spac_mlp* hm, const double* x, const double* y, const double* u, int32_t n,
                           const spac_mlp_train_cfg* cfg, spac_mlp_train_report* rep) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || !x || !y || !u || n < 1 || !cfg || cfg->steps < 1) return set_error(SPAC_E_INVALID_ARG, "bad arguments");
    spac_status rc;
    if ((rc = mlp_prepare_batch(m, x, y, u, n))) return rc;
    hipEvent_t e0, e1;
    MLP_HIP_CHECK(hipEventCreate(&e0));
    MLP_HIP_CHECK(hipEventCreate(&e1));
    MLP_HIP_CHECK(hipEventRecord(e0, m->stream));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int it = 0; it < cfg->steps; ++it) {
        float lam = (float)cfg->grad_weight;
        if ((rc = mlp_fwd_bwd(m, n, lam, true))) return rc;
        m->adam_t++;
        float lr = (float)cfg->lr;
        if (cfg->lr_final > 0.0) /* geometric decay from lr to lr_final */
            lr = (float)(cfg->lr * std::pow(cfg->lr_final / cfg->lr, (double)it / std::max(1, cfg->steps - 1)));
        float c1 = 1.0f - std::pow(b1, (float)m->adam_t), c2 = 1.0f - std::pow(b2, (float)m->adam_t);
        hipLaunchKernelGGL(k_adam, dim3((kNParams + 255) / 256), dim3(256), 0, m->stream, (float*)m->params.ptr,
                           (const float*)m->grads.ptr, (float*)m->am.ptr, (float*)m->av.ptr, kNParams, lr, b1, b2,
                           eps, c1, c2);
    }
    MLP_HIP_CHECK(hipEventRecord(e1, m->stream));
    MLP_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    MLP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rep) {
        rep->train_ms = ms;
        rep->steps = cfg->steps;
        if ((rc = mlp_fwd_bwd(m, n, (float)cfg->grad_weight, false))) return rc;
        if ((rc = mlp_loss_host(m, n, &rep->final_loss))) return rc;
    }
    return SPAC_OK;
}

spac_status spac_mlp_infer(spac_mlp* hm, const double* x, const double* y, int32_t n, float* out4) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || !x || !y || !out4 || n < 1) return set_error(SPAC_E_INVALID_ARG, "bad arguments");
    spac_status rc;
    DeviceBuffer io;
    if ((rc = io.ensure((size_t)n * (16 + 16)))) return rc;
    double* dx = (double*)io.ptr;
    double* dy = dx + n;
    float* dout = (float*)(dy + n);
    MLP_HIP_CHECK(hipMemcpyAsync(dx, x, (size_t)n * 8, hipMemcpyHostToDevice, m->stream));
    MLP_HIP_CHECK(hipMemcpyAsync(dy, y, (size_t)n * 8, hipMemcpyHostToDevice, m->stream));
    if ((rc = mlp_infer_dev(m, dx, dy, n, 0, dout))) { io.release(); return rc; }
    MLP_HIP_CHECK(hipMemcpyAsync(out4, dout, (size_t)n * 16, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(m->stream));
    io.release();
    return SPAC_OK;
}

/* FP32 reference forward (the training kernels) for checks: out3 = [3][n] */
spac_status spac_mlp_forward_fp32(spac_mlp* hm, const double* x, const double* y, int32_t n, float* out3) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || !x || !y || !out3 || n < 1) return set_error(SPAC_E_INVALID_ARG, "bad arguments");
    spac_status rc;
    if ((rc = mlp_prepare_batch(m, x, y, nullptr, n))) return rc;
    hipStream_t s = m->stream;
    const size_t N = (size_t)n;
    float* A = (float*)m->act.ptr;
    float* z = A;
    float* h1 = z + 2 * N;
    float* h2 = h1 + kH * N;
    float* h3 = h2 + kH * N;
    float* t3x = h3 + 5 * kH * N; /* not computed here: k_output reads it only for the loss */
    float* out = h3 + 7 * kH * N + 3 * kH * N;
    const float* P = (const float*)m->params.ptr;
    const double* dx = (const double*)m->pts.ptr;
    dim3 blk(256), g1((n + 255) / 256), gH((n + 255) / 256, kH);
    hipLaunchKernelGGL(k_input, g1, blk, 0, s, dx, dx + N, n, m->lox, m->loy, m->sx, m->sy, z);
    hipLaunchKernelGGL(k_dense_tanh, gH, blk, 0, s, P + oW1, P + ob1, z, 2, h1, kH, n);
    hipLaunchKernelGGL(k_dense_tanh, gH, blk, 0, s, P + oW2, P + ob2, h1, kH, h2, kH, n);
    hipLaunchKernelGGL(k_dense_tanh, gH, blk, 0, s, P + oW3, P + ob3, h2, kH, h3, kH, n);
    hipLaunchKernelGGL(k_output, g1, blk, 0, s, P + oW4, P + ob4, h3, t3x, t3x, (const double*)nullptr, n, 0.0f, out,
                       (float*)nullptr, (float*)nullptr);
    MLP_HIP_CHECK(hipGetLastError());
    MLP_HIP_CHECK(hipMemcpyAsync(out3, out, N * 3 * 4, hipMemcpyDeviceToHost, s));
    MLP_HIP_CHECK(hipStreamSynchronize(s));
    return SPAC_OK;
}

spac_status spac_mlp_build_table(spac_mlp* hm, int32_t t, double* infer_ms) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || t < 2 || t > 8192) return set_error(SPAC_E_INVALID_ARG, "table size must be in [2, 8192]");
    spac_status rc;
    if ((rc = m->table.ensure((size_t)t * t * 16))) return rc;
    hipEvent_t e0, e1;
    MLP_HIP_CHECK(hipEventCreate(&e0));
    MLP_HIP_CHECK(hipEventCreate(&e1));
    MLP_HIP_CHECK(hipEventRecord(e0, m->stream));
    if ((rc = mlp_infer_dev(m, nullptr, nullptr, t * t, t, (float*)m->table.ptr))) return rc;
    MLP_HIP_CHECK(hipEventRecord(e1, m->stream));
    MLP_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    MLP_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (infer_ms) *infer_ms = ms;
    m->table_t = t;
    return SPAC_OK;
}

spac_status spac_mlp_read_table(spac_mlp* hm, float* out4) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || !out4 || !m->table_t) return set_error(SPAC_E_INVALID_ARG, "no table");
    MLP_HIP_CHECK(hipMemcpy(out4, m->table.ptr, (size_t)m->table_t * m->table_t * 16, hipMemcpyDeviceToHost));
    return SPAC_OK;
}

spac_status spac_mlp_attach(spac_mlp* hm, spac_problem* hp, int32_t fallback, double max_ratio) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || !hp) return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (!m->table_t) return set_error(SPAC_E_INVALID_ARG, "the MLP has no table (spac_mlp_build_table)");
    spac_cv_config cfg;
    std::memset(&cfg, 0, sizeof cfg);
    cfg.mode = SPAC_CV_TABLE;
    cfg.fallback = fallback;
    cfg.max_ratio = max_ratio;
    cfg.table_size = m->table_t;
    for (int i = 0; i < 4; ++i) cfg.table_domain[i] = m->desc.domain[i];
    return problem_set_cv(reinterpret_cast<Problem*>(hp), &cfg, m->table.ptr, true);
}

} /* extern "C" */
