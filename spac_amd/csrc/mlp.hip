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
 * Training (FP32, deterministic: every sum has a fixed order): one kernel per step
 *   runs, for 16 samples per block, dense+tanh x3 with forward-mode tangents, the
 *   heads and the loss gradient, back-propagation and the weight gradients, the
 *   64 x 64 products on v_mfma_f32_16x16x4f32; a second kernel sums the per-block
 *   gradients in order and takes the Adam step.
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
#include "fk_device.h"
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
static_assert(kNParams == kNetParams && oW2 == kNetW2 && ob3 == kNetB3 && oW4 == kNetW4 && ob4 == kNetB4,
              "the per-step network evaluation (fk_device.h) reads this layout");
/* row stride of the per-block gradient rows, padded to a multiple of 4 floats (16-byte rows;
 * the pad float is never written nor read) */
constexpr int kPartStride = (kNParams + 3) & ~3;

/* ------------------------------------------------------------ training -- */
constexpr int kTile = 16;        /* samples per block */
constexpr int kWs = 68;          /* LDS row stride of a transposed 64-wide weight matrix: float4
                                    aligned, and rows fall on different banks (stride 64 puts a
                                    wave's transposing writes on one bank) */
constexpr int kChunk = kTile;    /* samples per block of the forward kernels */
constexpr int kOG = 256 / kTile; /* groups of 4 features per block (16 x 4 = 64) */

/* z = 2 (x - lo) / (hi - lo) - 1, once per batch */
__global__ void k_input(const double* x, const double* y, int n, float lox, float loy, float sx, float sy,
                        float* z) {
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    z[s] = 2.0f * ((float)x[s] - lox) * sx - 1.0f;
    z[n + s] = 2.0f * ((float)y[s] - loy) * sy - 1.0f;
}

struct FwdArgs {
    const float* W;  /* [64][IN] */
    const float* b;
    const float* hin; /* [IN][n] */
    const float* txin;
    const float* tyin;
    float* hout;      /* [64][n] */
    float* txout;
    float* tyout;
    int n;
    float scx, scy;   /* layer 1: d z / d(x, y) */
    /* output heads (fused into the last hidden layer) */
    const float* W4;
    const float* b4;
    const double* u;  /* targets, or nullptr: no loss */
    float lam;
    float* out3;
    float* dO;
    float* loss;
};

/* One hidden layer with both forward-mode tangents for kTile samples:
 *   h = tanh(W hin + b), t = (1 - h^2) (W tin)   (layer 1: tin = 2 e_d / span_d).
 * Inputs and W^T are staged in LDS; thread t computes features 4 (t / kTile) .. +3 of sample
 * t % kTile.  HEADS: the block also evaluates the output layer (u, gx, gy heads; gx, gy =
 * forward-mode d u / d(x, y)), the residuals and dL/do, summing the 16 feature-group
 * partials in order; the last layer's tangents are then never written. */
template <int IN, bool HEADS>
__global__ void __launch_bounds__(256) k_fwd_tile(FwdArgs a) {
    __shared__ __align__(16) float Lw[IN * kWs]; /* W^T: [i][o] */
    __shared__ float Lh[IN * kTile], Ltx[IN * kTile], Lty[IN * kTile];
    __shared__ float Lred[kOG][5][kTile];
    const int tid = threadIdx.x, c = tid % kTile, og = tid / kTile;
    const int s0 = blockIdx.x * kTile, s = s0 + c, n = a.n;
    for (int k = tid; k < 64 * IN; k += 256) Lw[(k % IN) * kWs + k / IN] = a.W[k];
    for (int k = tid; k < IN * kTile; k += 256) {
        const int i = k / kTile, cc = k % kTile, gs = s0 + cc;
        const bool in = gs < n;
        Lh[k] = in ? a.hin[(size_t)i * n + gs] : 0.0f;
        if (IN != 2) {
            Ltx[k] = in ? a.txin[(size_t)i * n + gs] : 0.0f;
            Lty[k] = in ? a.tyin[(size_t)i * n + gs] : 0.0f;
        }
    }
    __syncthreads();
    const int f0 = 4 * og;
    float acc[4], ax[4], ay[4];
#pragma unroll
    for (int o = 0; o < 4; ++o) acc[o] = ax[o] = ay[o] = 0.0f;
#pragma unroll 4
    for (int i = 0; i < IN; ++i) {
        const float hv = Lh[i * kTile + c];
        float tx = 0.0f, ty = 0.0f;
        if (IN != 2) {
            tx = Ltx[i * kTile + c];
            ty = Lty[i * kTile + c];
        }
        const float4 w4 = *reinterpret_cast<const float4*>(Lw + i * kWs + f0);
        const float w[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
        for (int o = 0; o < 4; ++o) {
            acc[o] = fmaf(w[o], hv, acc[o]);
            if (IN != 2) {
                ax[o] = fmaf(w[o], tx, ax[o]);
                ay[o] = fmaf(w[o], ty, ay[o]);
            }
        }
    }
    float po[3] = {0.0f, 0.0f, 0.0f}, pgx = 0.0f, pgy = 0.0f;
#pragma unroll
    for (int o = 0; o < 4; ++o) {
        const int f = f0 + o;
        const float h = tanhf(acc[o] + a.b[f]);
        const float d = 1.0f - h * h;
        float tx, ty;
        if (IN == 2) {
            tx = d * (Lw[f] * a.scx);
            ty = d * (Lw[kWs + f] * a.scy);
        } else {
            tx = d * ax[o];
            ty = d * ay[o];
        }
        if (s < n) {
            const size_t k = (size_t)f * n + s;
            a.hout[k] = h;
            if (!HEADS) {
                a.txout[k] = tx;
                a.tyout[k] = ty;
            }
        }
        if (HEADS) {
            for (int q = 0; q < 3; ++q) po[q] = fmaf(a.W4[q * 64 + f], h, po[q]);
            pgx = fmaf(a.W4[f], tx, pgx);
            pgy = fmaf(a.W4[f], ty, pgy);
        }
    }
    if (!HEADS) return;
    Lred[og][0][c] = po[0];
    Lred[og][1][c] = po[1];
    Lred[og][2][c] = po[2];
    Lred[og][3][c] = pgx;
    Lred[og][4][c] = pgy;
    __syncthreads();
    if (og != 0 || s >= n) return;
    float v[5];
    for (int k = 0; k < 5; ++k) {
        float t = 0.0f;
        for (int g = 0; g < kOG; ++g) t += Lred[g][k][c];
        v[k] = t;
    }
    float o3[3];
    for (int q = 0; q < 3; ++q) {
        o3[q] = v[q] + a.b4[q];
        a.out3[(size_t)q * n + s] = o3[q];
    }
    if (!a.u) return;
    const float r0 = o3[0] - (float)a.u[s], r1 = o3[1] - v[3], r2 = o3[2] - v[4];
    const float inv = 1.0f / (float)n;
    a.dO[s] = 2.0f * r0 * inv;
    a.dO[(size_t)n + s] = 2.0f * a.lam * r1 * inv;
    a.dO[(size_t)2 * n + s] = 2.0f * a.lam * r2 * inv;
    a.loss[s] = r0 * r0 + a.lam * (r1 * r1 + r2 * r2);
}

/* One whole training step for kStepTiles tiles of kTile samples in one block, on FP32 MFMA
 * (v_mfma_f32_16x16x4f32: D[i][j] += sum_k A[i][k] B[k][j], 16 x 16 x 4; lane l supplies
 * A[l % 16][l / 16] and B[l / 16][l % 16] and holds D[4 (l / 16) + r][l % 16], r = 0..3).
 * Four waves per tile: lane l of wave w of a tile works on sample c = l % 16 of the tile and
 * lane group g = l / 16.
 *
 * Forward: every lane evaluates layer 1 (and its tangents) for features 16 g .. 16 g + 15 of
 * its sample, which is the B operand of layer 2 with the K order k = 16 g + j at step j.
 * Layers 2 and 3 are W [h | tx | ty] (M = 64, N = 3 x 16 samples, K = 64): wave w computes
 * output features 16 w .. 16 w + 15, so its A operand W[16 w + c][16 g + j] is four float4
 * loads per lane; the outputs go through LDS ([sample][feature]) to the next layer.
 * Heads and dL/do: per-lane partial sums over the lane's four features, then summed over
 * (wave, lane group) in order.  Backward: delta_{L-1} = (1 - h^2) W_L^T delta_L on the same
 * tiling.  Weight gradients dW_L = delta_L h_{L-1}^T over all the block's samples (M = 64,
 * N = 64, K = 16 kStepTiles): wave v of the block takes rows 16 (v % 4) .. +15 and
 * 4 / kStepTiles column tiles.  The block's gradient sums go to part[block][...] (non-temporal
 * stores: read once, by the reduction) for the fixed-order reduction in k_reduce_adam, so a step writes and reads n / (16 kStepTiles)
 * gradient rows.  Samples past n have z = 0 and dL/do = 0, so they add nothing. */
typedef float f32x4 __attribute__((ext_vector_type(4)));
constexpr int kStepTiles = 1;
constexpr int kStepSamples = kTile * kStepTiles;
constexpr int kStepThreads = 256 * kStepTiles;
static_assert(4 % kStepTiles == 0, "column tiles per wave");
constexpr int kXs = 68; /* LDS row stride of a [sample][feature] tile: b128 accesses of 16
                           consecutive samples fall on distinct banks, and rows 4 samples apart
                           start 16 banks apart */

struct StepArgs {
    const float* P;   /* parameters */
    const float* z;   /* [2][n] scaled inputs */
    const double* u;  /* targets */
    float lam;
    float scx, scy;
    float* part;      /* [block][kNParams] */
    float* loss;      /* [n] per-sample loss (reports) */
    int n;
};

__device__ __forceinline__ f32x4 mfma16(float a, float b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ void load16(const float* p, float* v) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 t = *reinterpret_cast<const float4*>(p + 4 * q);
        v[4 * q] = t.x; v[4 * q + 1] = t.y; v[4 * q + 2] = t.z; v[4 * q + 3] = t.w;
    }
}
__device__ __forceinline__ void store4(float* p, const float (&v)[4]) {
    *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
}

/* tanh for training: 1 - 2 / (e^{2x} + 1) with the hardware 2^x and reciprocal (a few ulp),
 * and for |x| < 1/4 the odd Taylor polynomial to x^9 (truncation below 2.2e-9), so that small
 * arguments keep their relative precision. */
__device__ __forceinline__ float train_tanh(float x) {
    const float e = __builtin_amdgcn_exp2f(2.8853900817779268f * x);
    const float big = 1.0f - 2.0f * __builtin_amdgcn_rcpf(e + 1.0f);
    const float x2 = x * x;
    const float p =
        x * fmaf(x2, fmaf(x2, fmaf(x2, fmaf(x2, 62.0f / 2835.0f, -17.0f / 315.0f), 2.0f / 15.0f), -1.0f / 3.0f), 1.0f);
    return fabsf(x) < 0.25f ? p : big;
}

/* Weight gradient of a 64 x 64 layer over the block's samples, rows 16 (v % 4) .. +15 and
 * column tiles (4 / kStepTiles) (v / 4) .. : dW[o][i] = sum_s D[s][o] X[s][i] (D, X are
 * [sample][feature] LDS tiles; K order: step j, lane group g -> sample 16 (j / 4) + 4 g + j % 4). */
__device__ __forceinline__ void wgrad64(const float* D, const float* X, float* out, int v, int g, int c) {
    constexpr int kSteps = 4 * kStepTiles, kCols = 4 / kStepTiles;
    const int o0 = 16 * (v & 3);
    float av[kSteps];
#pragma unroll
    for (int j = 0; j < kSteps; ++j) av[j] = D[(16 * (j >> 2) + 4 * g + (j & 3)) * kXs + o0 + c];
#pragma unroll
    for (int q = 0; q < kCols; ++q) {
        const int nt = kCols * (v >> 2) + q;
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < kSteps; ++j)
            acc = mfma16(av[j], X[(16 * (j >> 2) + 4 * g + (j & 3)) * kXs + 16 * nt + c], acc);
#pragma unroll
        for (int r = 0; r < 4; ++r) __builtin_nontemporal_store(acc[r], &out[(o0 + 4 * g + r) * 64 + 16 * nt + c]);
    }
}

__global__ void __launch_bounds__(kStepThreads) k_train_step_mfma(StepArgs a) {
    /* [sample][feature] tiles of all the block's samples.  X2 holds h2, t2x, t2y; after layer 3
     * has read them, X2[1] takes delta3 and X2[2] delta2, and H3 takes delta1 once dW4 is done. */
    __shared__ __align__(16) float X2[3][kStepSamples * kXs];
    __shared__ __align__(16) float H1[kStepSamples * kXs];
    __shared__ __align__(16) float H3[kStepSamples * kXs];
    __shared__ float Lz[2][kStepSamples];
    __shared__ float Lred[kStepTiles][4][4][5][kTile]; /* [tile][wave][lane group][head term][sample] */
    __shared__ float LdO[3][kStepSamples];
    float* const D3 = X2[1];
    float* const D2 = X2[2];
    float* const D1 = H3;
    const int tid = threadIdx.x, v = tid >> 6, tt = v >> 2, w = v & 3, lane = tid & 63, g = lane >> 4,
              c = lane & 15;
    const int sc = kTile * tt + c; /* the lane's sample in the block */
    const int s = blockIdx.x * kStepSamples + sc, n = a.n;
    const float* P = a.P;
    const int fo = 16 * w + 4 * g; /* first of the lane's four output features */

    /* Every global operand of the step is loaded here, in the order of first use, so that
     * no later phase waits on memory: the lane's inputs, layer 1's weights and biases for
     * features 16 g .. 16 g + 15, the MFMA A operands (forward W[16 w + c][16 g + j],
     * backward W[16 g + j][16 w + c]), the biases and head weights of the lane's four output
     * features, and the target of the sample whose residual this thread forms. */
    const float z0 = s < n ? a.z[s] : 0.0f, z1 = s < n ? a.z[(size_t)n + s] : 0.0f;
    float w1v[32], b1v[16];
    load16(P + oW1 + 32 * g, w1v);
    load16(P + oW1 + 32 * g + 16, w1v + 16);
    load16(P + ob1 + 16 * g, b1v);
    float aW2[16], aW3[16], aW2t[16], aW3t[16];
    load16(P + oW2 + (16 * w + c) * 64 + 16 * g, aW2);
    load16(P + oW3 + (16 * w + c) * 64 + 16 * g, aW3);
#pragma unroll
    for (int j = 0; j < 16; ++j) aW3t[j] = P[oW3 + (16 * g + j) * 64 + 16 * w + c];
#pragma unroll
    for (int j = 0; j < 16; ++j) aW2t[j] = P[oW2 + (16 * g + j) * 64 + 16 * w + c];
    const float4 b2v = *reinterpret_cast<const float4*>(P + ob2 + fo);
    const float4 b3v = *reinterpret_cast<const float4*>(P + ob3 + fo);
    float4 w4v[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) w4v[q] = *reinterpret_cast<const float4*>(P + oW4 + q * 64 + fo);
    const int sh = blockIdx.x * kStepSamples + tid; /* sample of the residual (tid < kStepSamples) */
    const float b40 = P[ob4], b41 = P[ob4 + 1], b42 = P[ob4 + 2];
    const float ut = (tid < kStepSamples && sh < n) ? (float)a.u[sh] : 0.0f;
    const float b2a[4] = {b2v.x, b2v.y, b2v.z, b2v.w}, b3a[4] = {b3v.x, b3v.y, b3v.z, b3v.w};
    const float w4a[3][4] = {{w4v[0].x, w4v[0].y, w4v[0].z, w4v[0].w},
                             {w4v[1].x, w4v[1].y, w4v[1].z, w4v[1].w},
                             {w4v[2].x, w4v[2].y, w4v[2].z, w4v[2].w}};

    /* layer 1 for features 16 g + j of the lane's sample */
    float xh[16], xtx[16], xty[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
        const float w0 = w1v[2 * j], w1 = w1v[2 * j + 1];
        const float h = train_tanh(fmaf(w1, z1, w0 * z0) + b1v[j]);
        const float d = 1.0f - h * h;
        xh[j] = h;
        xtx[j] = d * (w0 * a.scx);
        xty[j] = d * (w1 * a.scy);
    }
    if (w == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float hv[4] = {xh[4 * q], xh[4 * q + 1], xh[4 * q + 2], xh[4 * q + 3]};
            store4(H1 + sc * kXs + 16 * g + 4 * q, hv);
        }
        if (g == 0) {
            Lz[0][sc] = z0;
            Lz[1][sc] = z1;
        }
    }

    /* layer 2 */
    float h2r[4];
    {
        f32x4 ah = {0.0f, 0.0f, 0.0f, 0.0f}, ax = ah, ay = ah;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            ah = mfma16(aW2[j], xh[j], ah);
            ax = mfma16(aW2[j], xtx[j], ax);
            ay = mfma16(aW2[j], xty[j], ay);
        }
        float tx[4], ty[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float h = train_tanh(ah[r] + b2a[r]);
            const float d = 1.0f - h * h;
            h2r[r] = h;
            tx[r] = d * ax[r];
            ty[r] = d * ay[r];
        }
        store4(X2[0] + sc * kXs + fo, h2r);
        store4(X2[1] + sc * kXs + fo, tx);
        store4(X2[2] + sc * kXs + fo, ty);
    }
    __syncthreads();

    /* layer 3 and the heads' partial sums over the lane's features */
    float h3r[4];
    {
        load16(X2[0] + sc * kXs + 16 * g, xh);
        load16(X2[1] + sc * kXs + 16 * g, xtx);
        load16(X2[2] + sc * kXs + 16 * g, xty);
        f32x4 ah = {0.0f, 0.0f, 0.0f, 0.0f}, ax = ah, ay = ah;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            ah = mfma16(aW3[j], xh[j], ah);
            ax = mfma16(aW3[j], xtx[j], ax);
            ay = mfma16(aW3[j], xty[j], ay);
        }
        float pg[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const float h = train_tanh(ah[r] + b3a[r]);
            const float d = 1.0f - h * h;
            h3r[r] = h;
#pragma unroll
            for (int q = 0; q < 3; ++q) pg[q] = fmaf(w4a[q][r], h, pg[q]);
            pg[3] = fmaf(w4a[0][r], d * ax[r], pg[3]);
            pg[4] = fmaf(w4a[0][r], d * ay[r], pg[4]);
        }
        store4(H3 + sc * kXs + fo, h3r);
#pragma unroll
        for (int k = 0; k < 5; ++k) Lred[tt][w][g][k][c] = pg[k];
    }
    __syncthreads();

    /* heads, residuals, dL/do */
    if (tid < kStepSamples) {
        const int t2 = tid / kTile, c2 = tid % kTile;
        float hv[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            float t = 0.0f;
            for (int ww = 0; ww < 4; ++ww)
                for (int gg = 0; gg < 4; ++gg) t += Lred[t2][ww][gg][k][c2];
            hv[k] = t;
        }
        float dO0 = 0.0f, dO1 = 0.0f, dO2 = 0.0f;
        if (sh < n) {
            const float o0 = hv[0] + b40, o1 = hv[1] + b41, o2 = hv[2] + b42;
            const float r0 = o0 - ut, r1 = o1 - hv[3], r2 = o2 - hv[4];
            const float inv = 1.0f / (float)n;
            dO0 = 2.0f * r0 * inv;
            dO1 = 2.0f * a.lam * r1 * inv;
            dO2 = 2.0f * a.lam * r2 * inv;
            a.loss[sh] = r0 * r0 + a.lam * (r1 * r1 + r2 * r2);
        }
        LdO[0][tid] = dO0;
        LdO[1][tid] = dO1;
        LdO[2][tid] = dO2;
    }
    __syncthreads();

    float* out = a.part + (size_t)blockIdx.x * kPartStride;
    /* delta3 = (1 - h3^2) W4^T dO for the lane's features; dW4, db4 */
    {
        float d3[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            float t = 0.0f;
            for (int q = 0; q < 3; ++q) t = fmaf(w4a[q][r], LdO[q][sc], t);
            d3[r] = (1.0f - h3r[r] * h3r[r]) * t;
        }
        store4(D3 + sc * kXs + fo, d3);
        if (tid < 3 * 64) {
            const int q = tid >> 6, i = tid & 63;
            float gs = 0.0f;
            for (int cc = 0; cc < kStepSamples; ++cc) gs = fmaf(LdO[q][cc], H3[cc * kXs + i], gs);
            __builtin_nontemporal_store(gs, &out[oW4 + tid]);
        } else if (tid < 3 * 64 + 3) {
            const int q = tid - 3 * 64;
            float gs = 0.0f;
            for (int cc = 0; cc < kStepSamples; ++cc) gs += LdO[q][cc];
            __builtin_nontemporal_store(gs, &out[ob4 + q]);
        }
    }
    __syncthreads();

    /* delta2 = (1 - h2^2) W3^T delta3; dW3 = delta3 h2^T, db3 */
    {
        load16(D3 + sc * kXs + 16 * g, xh);
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < 16; ++j) acc = mfma16(aW3t[j], xh[j], acc);
        float d2[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) d2[r] = (1.0f - h2r[r] * h2r[r]) * acc[r];
        store4(D2 + sc * kXs + fo, d2);
        wgrad64(D3, X2[0], out + oW3, v, g, c);
        if (tid < 64) {
            float gs = 0.0f;
            for (int cc = 0; cc < kStepSamples; ++cc) gs += D3[cc * kXs + tid];
            __builtin_nontemporal_store(gs, &out[ob3 + tid]);
        }
    }
    __syncthreads();

    /* delta1 = (1 - h1^2) W2^T delta2; dW2 = delta2 h1^T, db2 */
    {
        load16(D2 + sc * kXs + 16 * g, xh);
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int j = 0; j < 16; ++j) acc = mfma16(aW2t[j], xh[j], acc);
        const float4 h1 = *reinterpret_cast<const float4*>(H1 + sc * kXs + fo);
        const float hv[4] = {h1.x, h1.y, h1.z, h1.w};
        float d1[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) d1[r] = (1.0f - hv[r] * hv[r]) * acc[r];
        store4(D1 + sc * kXs + fo, d1);
        wgrad64(D2, H1, out + oW2, v, g, c);
        if (tid < 64) {
            float gs = 0.0f;
            for (int cc = 0; cc < kStepSamples; ++cc) gs += D2[cc * kXs + tid];
            __builtin_nontemporal_store(gs, &out[ob2 + tid]);
        }
    }
    __syncthreads();

    /* dW1 = delta1 z^T, db1 */
    if (tid < 128) {
        const int o = tid >> 1, i = tid & 1;
        float gs = 0.0f;
        for (int cc = 0; cc < kStepSamples; ++cc) gs = fmaf(D1[cc * kXs + o], Lz[i][cc], gs);
        __builtin_nontemporal_store(gs, &out[oW1 + tid]);
    } else if (tid < 192) {
        const int o = tid - 128;
        float gs = 0.0f;
        for (int cc = 0; cc < kStepSamples; ++cc) gs += D1[cc * kXs + o];
        __builtin_nontemporal_store(gs, &out[ob1 + o]);
    }
}

/* grad[k] = sum over chunks of part[chunk][k] in a fixed order (32 interleaved
 * groups of chunks, each summed in index order, then the groups in order;
 * non-temporal loads: every row is read once); then, in training, one Adam
 * step with a geometric learning-rate decay.  The step counter lives in device
 * memory so that a captured graph of steps replays unchanged.  It is
 * double-buffered by step parity: a step of parity p reads ctr[p] (the steps
 * taken so far) in every block, and block 0 writes ctr[1 - p] = ctr[p] + 1 for
 * the next step, which nothing reads during this launch.  Captured graphs hold
 * an even number of steps, so a step's parity is fixed at capture. */
constexpr int kRedParams = 32, kRedGroups = 32;
struct AdamArgs {
    float* p;
    float* g;
    float* m;
    float* v;
    const float* part;  /* [chunk][kPartStride] */
    int chunks;
    int* ctr;           /* [2]: Adam steps taken so far, by step parity */
    int par;            /* this step's parity */
    int t0;             /* counter value at the first step of this schedule */
    int steps;
    float lr0, lr1;     /* lr1 <= 0: constant lr0 */
    float beta2;
    int adam;
};
__global__ void __launch_bounds__(kRedParams * kRedGroups) k_reduce_adam(AdamArgs a) {
    __shared__ float sh[kRedGroups][kRedParams];
    const int pl = threadIdx.x % kRedParams, grp = threadIdx.x / kRedParams;
    const int k = blockIdx.x * kRedParams + pl;
    float acc = 0.0f;
    if (k < kNParams) {
        /* sequential sum in chunk order; unrolled so that the loads overlap */
#pragma unroll 8
        for (int c = grp; c < a.chunks; c += kRedGroups)
            acc += __builtin_nontemporal_load(a.part + (size_t)c * kPartStride + k);
    }
    sh[grp][pl] = acc;
    __syncthreads();
    if (grp == 0 && k < kNParams) {
        float g = 0.0f;
        for (int q = 0; q < kRedGroups; ++q) g += sh[q][pl];
        a.g[k] = g;
        if (a.adam) {
            const float b1 = 0.9f, b2 = a.beta2, eps = 1e-8f;
            const int tp = a.ctr[a.par];
            const float t = (float)(tp + 1);
            float lr = a.lr0;
            if (a.lr1 > 0.0f && a.steps > 1)
                lr = a.lr0 * powf(a.lr1 / a.lr0, (float)(tp - a.t0) / (float)(a.steps - 1));
            const float c1 = 1.0f - powf(b1, t), c2 = 1.0f - powf(b2, t);
            const float mk = b1 * a.m[k] + (1.0f - b1) * g;
            const float vk = b2 * a.v[k] + (1.0f - b2) * g * g;
            a.m[k] = mk;
            a.v[k] = vk;
            a.p[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
        }
    }
    if (a.adam && blockIdx.x == 0 && threadIdx.x == 0) a.ctr[1 - a.par] = a.ctr[a.par] + 1;
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

/* u[i] = the u head of the inference output (u, gx, gy, 0) of point i */
__global__ void k_take_u(int n, const float4* __restrict__ out4, double* __restrict__ u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] = (double)out4[i].x;
}

__global__ void k_to_bf16(const float* in, short* out, int n) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) out[k] = f2bf(in[k]);
}

/* --------------------------------------------------------------- handle -- */
struct Mlp {
    spac_mlp_desc desc{};
    hipStream_t stream = nullptr;
    DeviceBuffer params, grads, am, av, wbf, act, pts, tgt, table, part, ctr, pred;
    double* pinned = nullptr; /* pinned host staging of spac_mlp_predict_system's result */
    size_t pinned_cap = 0;
    int n = 0; /* prepared batch size */
    int adam_t = 0;
    int table_t = 0;
    float beta2 = 0.999f; /* of the training call being enqueued */
    float lox = 0, loy = 0, sx = 1, sy = 1;
};

/* Activation buffer layout (floats, [feature][sample]): z(2) h1 h2 h3 tx1 ty1
 * tx2 ty2 tx3 ty3 d3 d2 d1 (64 each) out3(3) dO(3) loss(1). */
struct ActView {
    float *z, *h[3], *tx[3], *ty[3], *d3, *d2, *d1, *out3, *dO, *loss;
};
static size_t act_floats(size_t N) { return N * (2 + 13 * kH + 3 + 3 + 1); }
static ActView act_view(float* A, size_t N) {
    ActView v;
    v.z = A;
    float* p = A + 2 * N;
    for (int l = 0; l < 3; ++l) { v.h[l] = p; p += kH * N; }
    for (int l = 0; l < 3; ++l) { v.tx[l] = p; p += kH * N; v.ty[l] = p; p += kH * N; }
    v.d3 = p; p += kH * N;
    v.d2 = p; p += kH * N;
    v.d1 = p; p += kH * N;
    v.out3 = p; p += 3 * N;
    v.dO = p; p += 3 * N;
    v.loss = p;
    return v;
}
static int n_chunks(int n) { return (n + kChunk - 1) / kChunk; }
static int n_step_blocks(int n) { return (n + kStepSamples - 1) / kStepSamples; }

static spac_status mlp_prepare_batch(Mlp* m, const double* x, const double* y, const double* u, int n) {
    spac_status rc;
    const size_t N = (size_t)n;
    if ((rc = m->act.ensure(act_floats(N) * 4))) return rc;
    if ((rc = m->pts.ensure(N * 16))) return rc;
    if ((rc = m->tgt.ensure(N * 8))) return rc;
    if ((rc = m->part.ensure((size_t)n_step_blocks(n) * kPartStride * 4))) return rc;
    MLP_HIP_CHECK(hipMemcpyAsync(m->pts.ptr, x, N * 8, hipMemcpyHostToDevice, m->stream));
    MLP_HIP_CHECK(hipMemcpyAsync((char*)m->pts.ptr + N * 8, y, N * 8, hipMemcpyHostToDevice, m->stream));
    if (u) MLP_HIP_CHECK(hipMemcpyAsync(m->tgt.ptr, u, N * 8, hipMemcpyHostToDevice, m->stream));
    ActView v = act_view((float*)m->act.ptr, N);
    const double* dx = (const double*)m->pts.ptr;
    hipLaunchKernelGGL(k_input, dim3((n + 255) / 256), dim3(256), 0, m->stream, dx, dx + N, n, m->lox, m->loy, m->sx,
                       m->sy, v.z);
    MLP_HIP_CHECK(hipGetLastError());
    m->n = n;
    return SPAC_OK;
}

/* Forward (with tangents) and outputs over the prepared batch; with targets,
 * also the loss terms and dL/do. */
static void enqueue_forward(Mlp* m, bool with_targets, float lam) {
    const int n = m->n;
    const size_t N = (size_t)n;
    hipStream_t s = m->stream;
    ActView v = act_view((float*)m->act.ptr, N);
    const float* P = (const float*)m->params.ptr;
    const dim3 blk(256), grid(n_chunks(n));
    FwdArgs f;
    std::memset(&f, 0, sizeof f);
    f.n = n;
    f.W = P + oW1; f.b = P + ob1; f.hin = v.z; f.hout = v.h[0]; f.txout = v.tx[0]; f.tyout = v.ty[0];
    f.scx = 2.0f * m->sx; f.scy = 2.0f * m->sy;
    hipLaunchKernelGGL((k_fwd_tile<2, false>), grid, blk, 0, s, f);
    f.W = P + oW2; f.b = P + ob2; f.hin = v.h[0]; f.txin = v.tx[0]; f.tyin = v.ty[0];
    f.hout = v.h[1]; f.txout = v.tx[1]; f.tyout = v.ty[1];
    hipLaunchKernelGGL((k_fwd_tile<64, false>), grid, blk, 0, s, f);
    f.W = P + oW3; f.b = P + ob3; f.hin = v.h[1]; f.txin = v.tx[1]; f.tyin = v.ty[1];
    f.hout = v.h[2]; f.txout = nullptr; f.tyout = nullptr;
    f.W4 = P + oW4; f.b4 = P + ob4; f.u = with_targets ? (const double*)m->tgt.ptr : nullptr; f.lam = lam;
    f.out3 = v.out3; f.dO = v.dO; f.loss = v.loss;
    hipLaunchKernelGGL((k_fwd_tile<64, true>), grid, blk, 0, s, f);
}

/* One full step: forward, backward with per-chunk gradient partials, the fixed-order
 * reduction and (adam) the update; `par` is the parity of the step's index. */
static void enqueue_step(Mlp* m, float lam, bool adam, int t0, int steps, float lr0, float lr1, int par) {
    const int n = m->n;
    const size_t N = (size_t)n;
    hipStream_t s = m->stream;
    ActView v = act_view((float*)m->act.ptr, N);
    StepArgs st;
    st.P = (const float*)m->params.ptr;
    st.z = v.z;
    st.u = (const double*)m->tgt.ptr;
    st.lam = lam;
    st.scx = 2.0f * m->sx;
    st.scy = 2.0f * m->sy;
    st.part = (float*)m->part.ptr;
    st.loss = v.loss;
    st.n = n;
    hipLaunchKernelGGL(k_train_step_mfma, dim3(n_step_blocks(n)), dim3(kStepThreads), 0, s, st);
    AdamArgs a;
    a.p = (float*)m->params.ptr;
    a.g = (float*)m->grads.ptr;
    a.m = (float*)m->am.ptr;
    a.v = (float*)m->av.ptr;
    a.part = (const float*)m->part.ptr;
    a.chunks = n_step_blocks(n);
    a.ctr = (int*)m->ctr.ptr;
    a.par = par & 1;
    a.t0 = t0;
    a.steps = steps;
    a.lr0 = lr0;
    a.lr1 = lr1;
    a.beta2 = m->beta2;
    a.adam = adam ? 1 : 0;
    hipLaunchKernelGGL(k_reduce_adam, dim3((kNParams + kRedParams - 1) / kRedParams), dim3(kRedParams * kRedGroups),
                       0, s, a);
}

static spac_status mlp_loss_host(Mlp* m, double* loss) {
    const size_t N = (size_t)m->n;
    std::vector<float> ls(N);
    ActView v = act_view((float*)m->act.ptr, N);
    MLP_HIP_CHECK(hipMemcpyAsync(ls.data(), v.loss, N * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(m->stream));
    double acc = 0.0;
    for (float x : ls) acc += x;
    *loss = acc / (double)N;
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

/* The executable graphs and events of one training call.  They are released on every
 * return, error returns included, and only after the stream has drained, so that no graph
 * is destroyed while a launch of it may still be queued. */
struct TrainResources {
    hipStream_t stream;
    hipGraphExec_t gfull = nullptr, grest = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    explicit TrainResources(hipStream_t s) : stream(s) {}
    TrainResources(const TrainResources&) = delete;
    TrainResources& operator=(const TrainResources&) = delete;
    ~TrainResources() {
        (void)hipStreamSynchronize(stream);
        if (gfull) (void)hipGraphExecDestroy(gfull);
        if (grest) (void)hipGraphExecDestroy(grest);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
    }
};

/* Capture `k` training steps into one executable graph. */
static spac_status capture_steps(Mlp* m, int k, float lam, int t0, int steps, float lr0, float lr1,
                                 hipGraphExec_t* exec) {
    hipGraph_t g = nullptr;
    MLP_HIP_CHECK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < k; ++i) enqueue_step(m, lam, true, t0, steps, lr0, lr1, t0 + i);
    hipError_t e = hipStreamEndCapture(m->stream, &g);
    if (e != hipSuccess) return set_error(SPAC_E_HIP, "graph capture failed: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return set_error(SPAC_E_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
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
    if (!(d->first_layer_scale >= 0.0) || !(d->first_layer_scale < 1e3))
        return set_error(SPAC_E_INVALID_ARG, "first_layer_scale must be in [0, 1000)");
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
        (rc = m->am.ensure(kNParams * 4)) || (rc = m->av.ensure(kNParams * 4)) || (rc = m->ctr.ensure(64))) {
        spac_mlp_destroy(reinterpret_cast<spac_mlp*>(m));
        return rc;
    }
    /* Glorot-uniform weights, zero biases; with first_layer_scale s > 0, W1 and b1 uniform in
       (-s, s): the Glorot bound of W1 (0.30) keeps every first-layer tanh nearly linear over the
       scaled domain, and features of varied slope and offset fit the solution's bumps in fewer
       steps */
    std::vector<float> p(kNParams, 0.0f);
    std::mt19937_64 g(d->seed);
    auto fill = [&](int off, int n, float lim) {
        std::uniform_real_distribution<float> U(-lim, lim);
        for (int k = 0; k < n; ++k) p[off + k] = U(g);
    };
    auto glorot = [](int fo, int fi) { return std::sqrt(6.0f / (float)(fo + fi)); };
    const float s1 = (float)d->first_layer_scale;
    fill(oW1, kH * 2, s1 > 0.0f ? s1 : glorot(kH, 2));
    if (s1 > 0.0f) fill(ob1, kH, s1);
    fill(oW2, kH * kH, glorot(kH, kH));
    fill(oW3, kH * kH, glorot(kH, kH));
    fill(oW4, 3 * kH, glorot(3, kH));
    MLP_HIP_CHECK(hipMemcpy(m->params.ptr, p.data(), kNParams * 4, hipMemcpyHostToDevice));
    MLP_HIP_CHECK(hipMemset(m->am.ptr, 0, kNParams * 4));
    MLP_HIP_CHECK(hipMemset(m->av.ptr, 0, kNParams * 4));
    MLP_HIP_CHECK(hipMemset(m->ctr.ptr, 0, 64));
    *out = reinterpret_cast<spac_mlp*>(m);
    return SPAC_OK;
}

spac_status spac_mlp_destroy(spac_mlp* hm) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m) return SPAC_OK;
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    for (DeviceBuffer* b : {&m->params, &m->grads, &m->am, &m->av, &m->wbf, &m->act, &m->pts, &m->tgt, &m->table,
                            &m->part, &m->ctr, &m->pred})
        b->release();
    if (m->pinned) (void)hipHostFree(m->pinned);
    if (m->stream) (void)hipStreamDestroy(m->stream);
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
    MLP_HIP_CHECK(hipMemset(m->ctr.ptr, 0, 64));
    return SPAC_OK;
}

spac_status spac_mlp_loss_grad(spac_mlp* hm, const double* x, const double* y, const double* u, int32_t n, double lam,
                               float* grad, double* loss) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || !x || !y || !u || n < 1) return set_error(SPAC_E_INVALID_ARG, "bad arguments");
    spac_status rc;
    if ((rc = mlp_prepare_batch(m, x, y, u, n))) return rc;
    if (grad) enqueue_step(m, (float)lam, false, 0, 1, 0.0f, 0.0f, 0);
    else enqueue_forward(m, true, (float)lam);
    MLP_HIP_CHECK(hipGetLastError());
    if (loss && (rc = mlp_loss_host(m, loss))) return rc;
    if (grad) MLP_HIP_CHECK(hipMemcpyAsync(grad, m->grads.ptr, kNParams * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(m->stream));
    return SPAC_OK;
}

spac_status spac_mlp_train(spac_mlp* hm, const double* x, const double* y, const double* u, int32_t n,
                           const spac_mlp_train_cfg* cfg, spac_mlp_train_report* rep) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || !x || !y || !u || n < 1 || !cfg || cfg->steps < 1 || !(cfg->lr > 0.0))
        return set_error(SPAC_E_INVALID_ARG, "bad arguments (n >= 1, steps >= 1, lr > 0)");
    if (cfg->beta2 != 0.0 && !(cfg->beta2 > 0.0 && cfg->beta2 < 1.0))
        return set_error(SPAC_E_INVALID_ARG, "beta2 must be in (0, 1), or 0 for 0.999");
    spac_status rc;
    if ((rc = mlp_prepare_batch(m, x, y, u, n))) return rc;
    m->beta2 = cfg->beta2 > 0.0 ? (float)cfg->beta2 : 0.999f;
    const float lam = (float)cfg->grad_weight, lr0 = (float)cfg->lr;
    const float lr1 = cfg->lr_final > 0.0 ? (float)cfg->lr_final : 0.0f;
    const int t0 = m->adam_t;
    /* graphs of kPerGraph steps, plus one for the remainder; kPerGraph is even, so every
       replay of a graph starts at the parity it was captured with */
    constexpr int kPerGraph = 50;
    static_assert(kPerGraph % 2 == 0, "graph length must be even (step-counter parity)");
    const int full = cfg->steps / kPerGraph, rest = cfg->steps % kPerGraph;
    TrainResources r(m->stream);
    if (!cfg->eager) {
        if (full && (rc = capture_steps(m, kPerGraph, lam, t0, cfg->steps, lr0, lr1, &r.gfull))) return rc;
        if (rest && (rc = capture_steps(m, rest, lam, t0, cfg->steps, lr0, lr1, &r.grest))) return rc;
    }
    MLP_HIP_CHECK(hipEventCreate(&r.e0));
    MLP_HIP_CHECK(hipEventCreate(&r.e1));
    MLP_HIP_CHECK(hipEventRecord(r.e0, m->stream));
    if (cfg->eager) {
        for (int i = 0; i < cfg->steps; ++i) enqueue_step(m, lam, true, t0, cfg->steps, lr0, lr1, t0 + i);
        MLP_HIP_CHECK(hipGetLastError());
    } else {
        for (int i = 0; i < full; ++i) MLP_HIP_CHECK(hipGraphLaunch(r.gfull, m->stream));
        if (rest) MLP_HIP_CHECK(hipGraphLaunch(r.grest, m->stream));
    }
    MLP_HIP_CHECK(hipEventRecord(r.e1, m->stream));
    MLP_HIP_CHECK(hipEventSynchronize(r.e1));
    float ms = 0.f;
    MLP_HIP_CHECK(hipEventElapsedTime(&ms, r.e0, r.e1));
    m->adam_t += cfg->steps;
    if (rep) {
        rep->train_ms = ms;
        rep->steps = cfg->steps;
        rep->reserved = 0;
        enqueue_forward(m, true, lam);
        MLP_HIP_CHECK(hipGetLastError());
        if ((rc = mlp_loss_host(m, &rep->final_loss))) return rc;
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
    if ((rc = mlp_infer_dev(m, dx, dy, n, 0, dout))) {
        io.release();
        return rc;
    }
    MLP_HIP_CHECK(hipMemcpyAsync(out4, dout, (size_t)n * 16, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(m->stream));
    io.release();
    return SPAC_OK;
}

spac_status spac_mlp_forward_fp32(spac_mlp* hm, const double* x, const double* y, int32_t n, float* out3) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || !x || !y || !out3 || n < 1) return set_error(SPAC_E_INVALID_ARG, "bad arguments");
    spac_status rc;
    if ((rc = mlp_prepare_batch(m, x, y, nullptr, n))) return rc;
    enqueue_forward(m, false, 0.0f);
    MLP_HIP_CHECK(hipGetLastError());
    ActView v = act_view((float*)m->act.ptr, (size_t)n);
    MLP_HIP_CHECK(hipMemcpyAsync(out3, v.out3, (size_t)n * 3 * 4, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(m->stream));
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

spac_status spac_mlp_predict_system(spac_mlp* hm, const spac_system* hs, double* u) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    const DeviceSystem* sys = reinterpret_cast<const DeviceSystem*>(hs);
    if (!m || !sys || !u) return set_error(SPAC_E_INVALID_ARG, "null argument");
    int dev = -1;
    MLP_HIP_CHECK(hipGetDevice(&dev));
    if (dev != sys->device) return set_error(SPAC_E_INVALID_ARG, "the system lives on device %d; make it current first",
                                             sys->device);
    const int n = sys->n_rows;
    if (n < 1) return SPAC_OK;
    spac_status rc;
    if ((rc = m->pred.ensure((size_t)n * (16 + 8)))) return rc;
    if (m->pinned_cap < (size_t)n * 8) {
        double* old = m->pinned;
        m->pinned = nullptr;
        m->pinned_cap = 0;
        if (old) MLP_HIP_CHECK(hipHostFree(old));
        MLP_HIP_CHECK(hipHostMalloc((void**)&m->pinned, (size_t)n * 8, hipHostMallocDefault));
        m->pinned_cap = (size_t)n * 8;
    }
    const double* nx = (const double*)sys->node_xy.ptr;
    float* out4 = (float*)m->pred.ptr;
    double* ud = (double*)(out4 + 4 * (size_t)n);
    if ((rc = mlp_infer_dev(m, nx, nx + n, n, 0, out4))) return rc;
    hipLaunchKernelGGL(k_take_u, dim3((n + 255) / 256), dim3(256), 0, m->stream, n, (const float4*)out4, ud);
    MLP_HIP_CHECK(hipGetLastError());
    MLP_HIP_CHECK(hipMemcpyAsync(m->pinned, ud, (size_t)n * 8, hipMemcpyDeviceToHost, m->stream));
    MLP_HIP_CHECK(hipStreamSynchronize(m->stream));
    std::memcpy(u, m->pinned, (size_t)n * 8);
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

spac_status spac_mlp_attach_network(spac_mlp* hm, spac_problem* hp, int32_t fallback, double max_ratio) {
    Mlp* m = reinterpret_cast<Mlp*>(hm);
    if (!m || !hp) return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (!m->table_t) return set_error(SPAC_E_INVALID_ARG, "the MLP has no table (spac_mlp_build_table)");
    MLP_HIP_CHECK(hipStreamSynchronize(m->stream));
    spac_cv_config cfg;
    std::memset(&cfg, 0, sizeof cfg);
    cfg.mode = SPAC_CV_TABLE;
    cfg.fallback = fallback;
    cfg.max_ratio = max_ratio;
    cfg.table_size = m->table_t;
    for (int i = 0; i < 4; ++i) cfg.table_domain[i] = m->desc.domain[i];
    const float map[4] = {m->lox, m->loy, m->sx, m->sy};
    return problem_set_cv_network(reinterpret_cast<Problem*>(hp), &cfg, m->table.ptr, (const float*)m->params.ptr, map);
}

} /* extern "C" */
