/*
 * fk_rows.hip — interface-row (G, B) assembly on the GPU.
 *
 * One call assembles the rows of a whole batch of nodes
 * (Solve_OMP_Analytic(Nudo&, ...) + Reduce_Analytic(..., Nudo&, Interfaz) +
 * Update(Nudo&), Solvers/FeynmanKacSolver.cpp:233-292, 324-355):
 *   1. the persistent path kernel runs every node's paths, stopped by the
 *      domain rectangle or the node's interface circle;
 *   2. the estimator reduction gives B = mean of (Z + xi + Y g) over domain
 *      exits and (Z + xi) over circle exits, and the node statistics;
 *   3. boundary nodes (frontera) get B = g(x) with no paths (:346-347);
 *   4. the G kernels (one block per node) accumulate the row weights over the
 *      circle-exit paths, both in moment form.  Global IMQ interpolation:
 *      S_i = sum_p Y_p phi(theta_p, theta_i), G_j = -sum_i iPsi(i,j) S_i / N,
 *      each S_i a TwoSum-compensated sum in one fixed order (paths in chunks of
 *      256, wave w owns paths 64w..64w+63 of every chunk, four wave partials).
 *      Gaussian stencils: fixed-point moments Phi[s][i] = sum Y_p phi_i by
 *      stencil start (exact integer sums), then one S x S stencil inverse per
 *      row (spac_rbf.h).  Neither depends on the launch, and both equal the
 *      CPU restatement bit for bit.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/spac_fk.h"
#include "fk_device.h"
#include "fk_internal.h"
#include "spac_rbf.h"

namespace spac {

#define ROWS_HIP_CHECK(expr)                                                                   \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return set_error(SPAC_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                             __FILE__, __LINE__);                                              \
    } while (0)

constexpr int kRowsBlock = 256;
constexpr int kRowsWaves = kRowsBlock / 64;
constexpr int kMaxStencil = 16;
constexpr int kColGroups = 4;             /* columns per lane */
constexpr int kMaxCols = 64 * kColGroups; /* interfaces up to 256 nodes */

struct RowsArgs {
    int32_t n_jobs;
    int32_t rbf_kind;
    int32_t stencil_size;
    double gauss_eps;
    const int32_t* job_iface;
    const uint32_t* paths;
    const int64_t* rec_off;
    const int64_t* g_off;
    const double* cx;
    const double* cy;
    const int32_t* code;
    const double* c2;
    const int32_t* node_offset;
    const double* theta;
    const int64_t* ipsi_offset;
    const double* ipsi;
    const double* xt;
    const double* yt;
    const double* Y;
    const int8_t* outcome;
    double* G;
};

/* Global IMQ rows (reference mode, arcs only): moment form S_i = sum_p Y_p phi(theta_p,
 * theta_i) with lanes = columns; wave w of the block sums paths 64w..64w+63 of every 256-path
 * chunk in path order, the four wave partials merge in wave order, then G_j = -(sum_i
 * iPsi(i, j) S_i) / N. */
__global__ void __launch_bounds__(kRowsBlock) fk_rows_g_imq_kernel(RowsArgs a) {
    const int job = blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t np = a.paths[job];
    const int iface = a.job_iface[job];
    const int nb = a.node_offset[iface];
    const int M = a.node_offset[iface + 1] - nb;
    const double* th = a.theta + nb;
    const int code = a.code[iface];
    const double cx = a.cx[iface], cy = a.cy[iface], c2 = a.c2[iface];
    const double* ipsi = a.ipsi + (a.ipsi_offset[iface] >= 0 ? a.ipsi_offset[iface] : 0);
    double* G = a.G + a.g_off[job];

    __shared__ double s_th[kMaxCols];
    __shared__ double s_Y[kRowsBlock];
    __shared__ double s_t[kRowsBlock];
    __shared__ __align__(16) int s_s[kRowsBlock]; /* 0: circle exit, -1: not */
    __shared__ spac_ksum s_part[(kRowsWaves - 1) * kMaxCols];
    __shared__ double s_S[kMaxCols];

    const int lane = tid & 63, wave = tid >> 6;
    for (int i = tid; i < M; i += kRowsBlock) s_th[i] = th[i];
    __syncthreads();
    if (np == 0) {
        for (int j = tid; j < M; j += kRowsBlock) G[j] = 0.0;
        return;
    }
    spac_ksum acc[kColGroups];
#pragma unroll
    for (int q = 0; q < kColGroups; ++q) acc[q] = spac_ksum_zero();
    const int64_t off = a.rec_off[job];
    for (int64_t base = 0; base < np; base += kRowsBlock) {
        const int cnt = (np - base < (int64_t)kRowsBlock) ? (int)(np - base) : kRowsBlock;
        if (tid < cnt) {
            const int64_t g = off + base + tid;
            int s = -1;
            if (a.outcome[g] == SPAC_OUT_SUBDOMAIN) {
                double theta = 0.0;
                spac_exit_angle(code, a.xt[g] - cx, a.yt[g] - cy, s_th[0], &theta);
                s_Y[tid] = a.Y[g];
                s_t[tid] = theta;
                s = 0;
            }
            s_s[tid] = s;
        }
        __syncthreads();
        const int p0 = wave * 64, p1 = min(cnt, p0 + 64);
        for (int p = p0; p < p1; p += 4) {
            const int4 s4 = *reinterpret_cast<const int4*>(&s_s[p]);
            int sv[4] = {s4.x, s4.y, s4.z, s4.w};
            double yv[4], tv[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (p + k >= p1) sv[k] = -1;
                yv[k] = s_Y[p + k];
                tv[k] = s_t[p + k];
            }
#pragma unroll
            for (int q = 0; q < kColGroups; ++q) {
                const int j = lane + q * 64;
                if (q * 64 >= M) break; /* uniform */
                if (j < M) {
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        if (sv[k] >= 0) acc[q] = spac_ksum_add(acc[q], yv[k] * spac_imq(tv[k], s_th[j], c2));
                }
            }
        }
        __syncthreads();
    }
    if (wave > 0) {
#pragma unroll
        for (int q = 0; q < kColGroups; ++q) {
            const int j = lane + q * 64;
            if (j < M) s_part[(wave - 1) * kMaxCols + j] = acc[q];
        }
    }
    __syncthreads();
    const double N = (double)np;
    if (wave == 0) {
#pragma unroll
        for (int q = 0; q < kColGroups; ++q) {
            const int j = lane + q * 64;
            if (j >= M) continue;
            spac_ksum t = acc[q];
            for (int w = 1; w < kRowsWaves; ++w) t = spac_ksum_merge(t, s_part[(w - 1) * kMaxCols + j]);
            s_S[j] = spac_ksum_value(t);
        }
    }
    __syncthreads();
    for (int j = tid; j < M; j += kRowsBlock) {
        double w = 0.0;
        for (int i = 0; i < M; ++i) w = w + ipsi[(int64_t)i * M + j] * s_S[i];
        G[j] = -w / N;
    }
}

/* the stencil size whose column products are unrolled at compile time */
constexpr int kMomFastStencil = 15;

/* The columns of a full circle's row from its moments (both in LDS; the circle's one S x S
 * stencil inverse blk), each one FMA chain: G_c = -(sum over r, s = c - r mod M, of sum over i
 * of blk[i][r] Phi[s][i]) / N.  SS > 0 fixes S at compile time, so that every LDS address is
 * a base plus an immediate offset; SS = 0 takes the runtime S. */
template <int SS>
__device__ __forceinline__ void moment_columns_circle(int M, int S_rt, const double* __restrict__ blk,
                                                      const unsigned long long* __restrict__ mom, double N,
                                                      double* __restrict__ G) {
    const int S = SS > 0 ? SS : S_rt;
    for (int c = (int)threadIdx.x; c < M; c += (int)blockDim.x) {
        double acc = 0.0;
#pragma unroll 1
        for (int r = 0; r < S; ++r) {
            int s = c - r;
            if (s < 0) s += M;
            const unsigned long long* mrow = mom + s * S;
#pragma unroll
            for (int i = 0; i < S; ++i) acc = __builtin_fma(blk[i * S + r], spac_u64_as_f64(mrow[i]), acc);
        }
        G[c] = -acc / N;
    }
}

/* Gaussian stencil rows in moment form (spac_rbf.h), one block per row.
 * Pass 1: the largest |Y| over the row's circle exits (a max: exact in any order) fixes the
 * fixed-point scale q.  Pass 2: thread t takes the paths t, t + BLK, ...: exit angle, stencil
 * start s and the S Gaussians phi_i, then S 64-bit integer adds of Y phi_i into the moments
 * Phi[s][i] in LDS — integer sums, so they do not depend on the order of the adds.  Then the
 * moments become doubles and each column is one FMA chain in the restatement's order,
 *   G_c = -(sum over r = 0..S-1 with s = c - r, of sum over i = 0..S-1 of iPsi_s(i, r) Phi[s][i]) / N
 * (s taken mod M on a full circle; on an arc only starts 0..M-S exist).  There are no
 * per-path weight products (the S x S stencil inverse is applied once per row, not once per
 * path) and no read-modify-write chains: a path costs its exit angle, four exps and S LDS
 * atomics.  Non-finite Y values are left out (the row's estimator already reports them).
 * Dynamic LDS: M_max * S 64-bit moments. */
template <int BLK>
__global__ void __launch_bounds__(BLK) fk_rows_g_moment_kernel(RowsArgs a) {
    extern __shared__ unsigned long long s_mom[];
    __shared__ double s_th[kMaxCols];
    __shared__ double s_blk[kMaxStencil * kMaxStencil];
    __shared__ double s_red[BLK / 64];
    const int tid = threadIdx.x;
    const int job = (int)blockIdx.x;
    const int64_t np = a.paths[job];
    const int iface = a.job_iface[job];
    const int nb = a.node_offset[iface];
    const int M = a.node_offset[iface + 1] - nb;
    const int code = a.code[iface];
    const bool circle = code == 1;
    const double cx = a.cx[iface], cy = a.cy[iface];
    const int S = min(a.stencil_size, M);
    const double* ipsi = a.ipsi + (a.ipsi_offset[iface] >= 0 ? a.ipsi_offset[iface] : 0);
    double* G = a.G + a.g_off[job];
    if (np == 0) { /* uniform over the block: nobody reaches a barrier */
        for (int c = tid; c < M; c += BLK) G[c] = 0.0;
        return;
    }
    for (int i = tid; i < M; i += BLK) s_th[i] = a.theta[nb + i];
    if (circle)
        for (int i = tid; i < S * S; i += BLK) s_blk[i] = ipsi[i];
    for (int i = tid; i < M * S; i += BLK) s_mom[i] = 0ull;
    const int64_t off = a.rec_off[job];
    double ym = 0.0;
    for (int64_t p = tid; p < np; p += BLK) {
        const int64_t g = off + p;
        const double ay = SPAC_FABS(a.Y[g]);
        if (a.outcome[g] == SPAC_OUT_SUBDOMAIN && ay <= __DBL_MAX__ && ay > ym) ym = ay;
    }
    for (int o = 32; o > 0; o >>= 1) ym = fmax(ym, __shfl_xor(ym, o));
    if ((tid & 63) == 0) s_red[tid >> 6] = ym;
    __syncthreads();
    ym = s_red[0];
    for (int w = 1; w < BLK / 64; ++w) ym = fmax(ym, s_red[w]);
    const int q = spac_moment_scale((double)np, ym);
    for (int64_t p = tid; p < np; p += BLK) {
        const int64_t g = off + p;
        if (a.outcome[g] != SPAC_OUT_SUBDOMAIN) continue;
        const double Yp = a.Y[g];
        if (!(SPAC_FABS(Yp) <= __DBL_MAX__)) continue;
        double theta = 0.0;
        spac_exit_angle(code, a.xt[g] - cx, a.yt[g] - cy, s_th[0], &theta);
        /* the Gaussians one at a time (the operations of spac_circle_stencil / spac_arc_stencil) */
        if (circle) {
            double ga, gb;
            const int s = spac_circle_stencil_ab(M, S, a.gauss_eps, theta, &ga, &gb);
            unsigned long long* row = s_mom + s * S;
            spac_gauss_seq gs = spac_gauss_seq_start(ga, gb);
            for (int i = 0; i < S; ++i) {
                if (i > 0) spac_gauss_seq_next(&gs);
                atomicAdd(row + i, (unsigned long long)spac_to_fixed(Yp * gs.p, q));
            }
        } else {
            double L;
            const int s = spac_arc_stencil_start(s_th, M, S, theta, &L);
            unsigned long long* row = s_mom + s * S;
            for (int i = 0; i < S; ++i)
                atomicAdd(row + i,
                          (unsigned long long)spac_to_fixed(Yp * spac_gauss_phi((theta - s_th[s + i]) / L, a.gauss_eps), q));
        }
    }
    __syncthreads();
    for (int i = tid; i < M * S; i += BLK) s_mom[i] = spac_f64_as_u64(spac_from_fixed((int64_t)s_mom[i], q));
    __syncthreads();
    const double N = (double)np;
    if (circle && S == kMomFastStencil) {
        moment_columns_circle<kMomFastStencil>(M, S, s_blk, s_mom, N, G);
    } else if (circle) {
        moment_columns_circle<0>(M, S, s_blk, s_mom, N, G);
    } else {
        for (int c = tid; c < M; c += BLK) {
            double acc = 0.0;
            for (int r = 0; r < S; ++r) {
                const int s = c - r;
                if (s < 0 || s > M - S) continue;
                const double* blk = ipsi + (int64_t)s * S * S;
                const unsigned long long* mrow = s_mom + s * S;
                for (int i = 0; i < S; ++i) acc = __builtin_fma(blk[i * S + r], spac_u64_as_f64(mrow[i]), acc);
            }
            G[c] = -acc / N;
        }
    }
}

/* rows of at most this many paths take one wave per row */
constexpr int64_t kMomentSmallRow = 1024;

/* B for boundary nodes (g at the node) and the copy of every node's B. */
template <class Coef>
__global__ void fk_rows_b_kernel(int32_t n, const int32_t* frontera, const double* x, const double* y,
                                 const spac_point_stats* stats, double* B) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    B[k] = frontera[k] ? Coef::g(x[k], y[k]) : stats[k].B;
}
/* the same with a tabulated g where the problem's table holds one */
__global__ void fk_rows_b_tab_kernel(int32_t n, const int32_t* frontera, const double* x, const double* y,
                                     const spac_point_stats* stats, double* B, const spac_coef_table tab) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    CoefTab<CoefVarFranke> coef;
    coef.t = tab;
    B[k] = frontera[k] ? coef.g(x[k], y[k]) : stats[k].B;
}

static spac_status validate(const spac_interface_table* t, const spac_node_batch* nb) {
    if (!t || !nb) return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (t->n_interfaces < 1 || !t->cx || !t->cy || !t->radius || !t->code || !t->node_offset || !t->node_theta ||
        !t->node_global || !t->ipsi_offset || !t->ipsi || !t->c2)
        return set_error(SPAC_E_INVALID_ARG, "interface table is incomplete");
    if (t->rbf_kind != SPAC_RBF_IMQ_GLOBAL && t->rbf_kind != SPAC_RBF_GAUSS_STENCIL)
        return set_error(SPAC_E_INVALID_ARG, "unknown rbf_kind %d", t->rbf_kind);
    if (t->rbf_kind == SPAC_RBF_GAUSS_STENCIL &&
        (t->stencil_size < 1 || t->stencil_size > kMaxStencil || !(t->gauss_eps > 0.0)))
        return set_error(SPAC_E_INVALID_ARG, "stencil_size must be in [1, %d] and gauss_eps > 0", kMaxStencil);
    for (int k = 0; k < t->n_interfaces; ++k) {
        int M = t->node_offset[k + 1] - t->node_offset[k];
        if (M < 1 || M > kMaxCols)
            return set_error(SPAC_E_INVALID_ARG, "interface %d has %d nodes (1..%d supported)", k, M, kMaxCols);
        int c = t->code[k];
        bool known = c == 1 || c == -1 || c == -2 || c == -3 || c == -4 || c == -11 || c == -22 || c == -33 || c == -44;
        if (!known) return set_error(SPAC_E_BAD_INTERFACE, "interface %d has unknown code %d", k, c);
        if (!(t->radius[k] > 0.0)) return set_error(SPAC_E_INVALID_ARG, "interface %d has radius <= 0", k);
    }
    if (nb->n < 1 || !nb->x || !nb->y || !nb->iface || !nb->frontera || !nb->paths || !nb->h || !nb->stream)
        return set_error(SPAC_E_INVALID_ARG, "node batch is incomplete");
    for (int k = 0; k < nb->n; ++k) {
        int f = nb->iface[k];
        if (f < 0 || f >= t->n_interfaces) return set_error(SPAC_E_INVALID_ARG, "node %d: interface %d out of range", k, f);
        if (nb->frontera[k]) continue;
        if (nb->paths[k] < 1) return set_error(SPAC_E_INVALID_ARG, "node %d has no paths", k);
        if (t->rbf_kind == SPAC_RBF_IMQ_GLOBAL && (t->code[f] == 1 || t->ipsi_offset[f] < 0))
            return set_error(SPAC_E_BAD_INTERFACE,
                             "node %d: interface %d is a full circle; global IMQ rows exist for arcs only "
                             "(Mesh/Interfaz.cpp:248-250)", k, f);
    }
    return SPAC_OK;
}


/* Everything one row assembly needs on the device, uploaded once. */
struct RowsSetup {
    SolveSet ss;
    RowsArgs ra;
    int n = 0;
    int mmax = 1;     /* largest interface of the batch (dynamic LDS of the stencil kernel) */
    int g_block = 256; /* threads per row of the stencil kernel: 64 when every row is short */
    const int32_t* d_front = nullptr;
    double* d_B = nullptr;
    spac_point_stats* d_stats = nullptr;
    std::vector<int64_t> goff;  /* host copy of G_offset */
    std::vector<int32_t> gcol;  /* host copy of G_col */
};

static size_t stencil_lds_bytes(const RowsSetup& rs) {
    return (size_t)rs.mmax * (size_t)std::max(1, std::min(rs.ra.stencil_size, rs.mmax)) * 8;
}

} // namespace spac

/* The interface table on the device: cx, cy, c2, code in `ifc`; node_offset, then 8-aligned
 * ipsi_offset in `ifoff`; node angles in `ifnode`; stencil / global inverses in `ifipsi`.
 * Returns after the copies have finished (the host arrays may change afterwards). */
spac_status spac::iface_table_upload(const spac_interface_table* t, hipStream_t s, DeviceBuffer& ifc,
                                     DeviceBuffer& ifoff, DeviceBuffer& ifnode, DeviceBuffer& ifipsi) {
    const int K = t->n_interfaces;
    const int nodes = t->node_offset[K];
    int64_t nipsi = 0;
    for (int k = 0; k < K; ++k) {
        if (t->ipsi_offset[k] < 0) continue;
        int M = t->node_offset[k + 1] - t->node_offset[k];
        int64_t sz;
        if (t->rbf_kind == SPAC_RBF_IMQ_GLOBAL) sz = (int64_t)M * M;
        else {
            int S = std::min(t->stencil_size, M);
            sz = (t->code[k] == 1 ? 1 : (M - S + 1)) * (int64_t)S * S;
        }
        nipsi = std::max(nipsi, t->ipsi_offset[k] + sz);
    }
    const size_t szK = (size_t)K;
    const size_t noff_bytes = (((szK + 1) * 4 + 7) / 8) * 8;
    spac_status rc;
    if ((rc = ifc.ensure(szK * (3 * 8 + 4))) || (rc = ifnode.ensure((size_t)nodes * 8 + 8)) ||
        (rc = ifipsi.ensure((size_t)std::max<int64_t>(nipsi, 1) * 8)) || (rc = ifoff.ensure(noff_bytes + szK * 8)))
        return rc;
    double* d_cx = (double*)ifc.ptr;
    ROWS_HIP_CHECK(hipMemcpyAsync(d_cx, t->cx, szK * 8, hipMemcpyHostToDevice, s));
    ROWS_HIP_CHECK(hipMemcpyAsync(d_cx + K, t->cy, szK * 8, hipMemcpyHostToDevice, s));
    ROWS_HIP_CHECK(hipMemcpyAsync(d_cx + 2 * K, t->c2, szK * 8, hipMemcpyHostToDevice, s));
    ROWS_HIP_CHECK(hipMemcpyAsync(d_cx + 3 * K, t->code, szK * 4, hipMemcpyHostToDevice, s));
    ROWS_HIP_CHECK(hipMemcpyAsync(ifoff.ptr, t->node_offset, (szK + 1) * 4, hipMemcpyHostToDevice, s));
    ROWS_HIP_CHECK(hipMemcpyAsync((char*)ifoff.ptr + noff_bytes, t->ipsi_offset, szK * 8, hipMemcpyHostToDevice, s));
    ROWS_HIP_CHECK(hipMemcpyAsync(ifnode.ptr, t->node_theta, (size_t)nodes * 8, hipMemcpyHostToDevice, s));
    if (nipsi > 0) ROWS_HIP_CHECK(hipMemcpyAsync(ifipsi.ptr, t->ipsi, (size_t)nipsi * 8, hipMemcpyHostToDevice, s));
    ROWS_HIP_CHECK(hipStreamSynchronize(s));
    return SPAC_OK;
}

std::vector<char> spac::table_signature(const spac_interface_table* t) {
    const size_t K = (size_t)t->n_interfaces;
    const size_t nodes = (size_t)t->node_offset[K];
    std::vector<char> sig;
    auto put = [&sig](const void* src, size_t bytes) {
        const char* c = (const char*)src;
        sig.insert(sig.end(), c, c + bytes);
    };
    put(&t->n_interfaces, 4);
    put(&t->rbf_kind, 4);
    put(&t->stencil_size, 4);
    put(&t->gauss_eps, 8);
    put(t->cx, K * 8);
    put(t->cy, K * 8);
    put(t->radius, K * 8);
    put(t->c2, K * 8);
    put(t->code, K * 4);
    put(t->ipsi_offset, K * 8);
    put(t->node_offset, (K + 1) * 4);
    put(t->node_global, nodes * 4);
    return sig;
}

namespace spac {

static void fill_rows_args(Problem* p, RowsSetup* rs, int K, int rbf_kind, int stencil_size, double gauss_eps,
                           const DeviceBuffer& ifc, const DeviceBuffer& ifoff, const DeviceBuffer& ifnode,
                           const DeviceBuffer& ifipsi, const DeviceBuffer& jobs, const DeviceBuffer& goff);

/* Upload the node batch, interface table and work list into p's buffers and
 * allocate the records, so that rows_enqueue only launches kernels.  With sys (a device
 * system of the batch's mesh), the interface table and the job arrays are the copies the
 * system holds, and the host column list (gcol) is not built. */
static spac_status rows_prepare(Problem* p, const spac_interface_table* t, const spac_node_batch* nb, RowsSetup* rs,
                                const DeviceSystem* sys = nullptr) {
    const int n = nb->n;
    std::vector<uint32_t> paths(n);
    std::vector<double> circle(3 * (size_t)n);
    rs->n = n;
    rs->goff.assign(n + 1, 0);
    for (int k = 0; k < n; ++k) {
        int f = nb->iface[k];
        paths[k] = nb->frontera[k] ? 0u : nb->paths[k];
        circle[3 * k] = t->radius[f];
        circle[3 * k + 1] = t->cx[f];
        circle[3 * k + 2] = t->cy[f];
        int M = t->node_offset[f + 1] - t->node_offset[f];
        rs->goff[k + 1] = rs->goff[k] + M;
    }
    rs->mmax = 1;
    for (int k = 0; k < n; ++k) {
        int f = nb->iface[k];
        rs->mmax = std::max(rs->mmax, t->node_offset[f + 1] - t->node_offset[f]);
    }
    rs->ra.stencil_size = t->stencil_size;
    if (t->rbf_kind == SPAC_RBF_GAUSS_STENCIL) {
        int lds_max = 0;
        ROWS_HIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device));
        const size_t need = stencil_lds_bytes(*rs) + 5 * 1024; /* + the kernel's static LDS */
        if (need > (size_t)lds_max)
            return set_error(SPAC_E_INVALID_ARG, "interfaces of %d nodes need %zu B of LDS per block (device: %d)",
                             rs->mmax, need, lds_max);
        uint32_t pmax = 0;
        for (int k = 0; k < n; ++k) pmax = std::max(pmax, paths[k]);
        rs->g_block = (int64_t)pmax <= kMomentSmallRow ? 64 : 256;
    }
    if (!sys) {
        rs->gcol.resize((size_t)rs->goff[n]);
        for (int k = 0; k < n; ++k) {
            int f = nb->iface[k];
            int M = t->node_offset[f + 1] - t->node_offset[f];
            for (int j = 0; j < M; ++j) rs->gcol[rs->goff[k] + j] = t->node_global[t->node_offset[f] + j];
        }
    }
    spac_status rc;
    if ((rc = p->upload_solves(n, nb->x, nb->y, paths.data(), nb->h, nb->stream, circle.data(), &rs->ss))) return rc;
    const int K = t->n_interfaces;
    if ((rc = p->rows_G.ensure((size_t)std::max<int64_t>(rs->goff[n], 1) * 8))) return rc;
    if ((rc = p->rows_part.ensure((size_t)n * 8))) return rc; /* B */
    if ((rc = p->stats.ensure((size_t)n * sizeof(spac_point_stats)))) return rc;
    if ((rc = p->ensure_records(rs->ss.total_paths))) return rc;
    if ((rc = p->queue.ensure(64))) return rc;
    hipStream_t s = p->stream;
    const DeviceBuffer *ifc, *ifoff, *ifnode, *ifipsi, *jobs, *goff;
    if (sys) {
        ifc = &sys->t_ifc;
        ifoff = &sys->t_ifoff;
        ifnode = &sys->t_ifnode;
        ifipsi = &sys->t_ifipsi;
        jobs = &sys->t_jobs;
        goff = &sys->t_goff;
    } else {
        if ((rc = iface_table_upload(t, s, p->ifc, p->ifoff, p->ifnode, p->ifipsi))) return rc;
        if ((rc = p->rows_iface.ensure((size_t)n * 4 + (size_t)n * 4))) return rc; /* job_iface, frontera */
        if ((rc = p->rows_goff.ensure((size_t)(n + 1) * 8))) return rc;
        int32_t* d_iface = (int32_t*)p->rows_iface.ptr;
        ROWS_HIP_CHECK(hipMemcpyAsync(d_iface, nb->iface, (size_t)n * 4, hipMemcpyHostToDevice, s));
        ROWS_HIP_CHECK(hipMemcpyAsync(d_iface + n, nb->frontera, (size_t)n * 4, hipMemcpyHostToDevice, s));
        ROWS_HIP_CHECK(hipMemcpyAsync(p->rows_goff.ptr, rs->goff.data(), (size_t)(n + 1) * 8, hipMemcpyHostToDevice, s));
        ROWS_HIP_CHECK(hipStreamSynchronize(s));
        ifc = &p->ifc;
        ifoff = &p->ifoff;
        ifnode = &p->ifnode;
        ifipsi = &p->ifipsi;
        jobs = &p->rows_iface;
        goff = &p->rows_goff;
    }
    fill_rows_args(p, rs, K, t->rbf_kind, t->stencil_size, t->gauss_eps, *ifc, *ifoff, *ifnode, *ifipsi, *jobs, *goff);
    return SPAC_OK;
}

/* The row kernels' arguments (RowsArgs) over the interface table in (ifc, ifoff, ifnode, ifipsi)
 * of iface_table_upload's layout, the job arrays `jobs` (interface, then boundary flag; int32
 * [n] each) and G offsets `goff` (int64 [n + 1]), the solve table rs->ss and p's records. */
static void fill_rows_args(Problem* p, RowsSetup* rs, int K, int rbf_kind, int stencil_size, double gauss_eps,
                           const DeviceBuffer& ifcb, const DeviceBuffer& ifoffb, const DeviceBuffer& ifnodeb,
                           const DeviceBuffer& ifipsib, const DeviceBuffer& jobsb, const DeviceBuffer& goffb) {
    const DeviceBuffer *ifc = &ifcb, *ifoff = &ifoffb, *ifnode = &ifnodeb, *ifipsi = &ifipsib, *jobs = &jobsb,
                       *goff = &goffb;
    const int n = rs->n;
    const size_t szK = (size_t)K;
    const size_t noff_bytes = (((szK + 1) * 4 + 7) / 8) * 8;
    const double* d_cx = (const double*)ifc->ptr;
    RowsArgs& ra = rs->ra;
    std::memset(&ra, 0, sizeof ra);
    ra.n_jobs = n;
    ra.rbf_kind = rbf_kind;
    ra.stencil_size = stencil_size;
    ra.gauss_eps = gauss_eps;
    ra.job_iface = (const int32_t*)jobs->ptr;
    ra.paths = rs->ss.paths;
    ra.rec_off = rs->ss.rec_off;
    ra.g_off = (const int64_t*)goff->ptr;
    ra.cx = d_cx;
    ra.cy = d_cx + K;
    ra.c2 = d_cx + 2 * K;
    ra.code = (const int32_t*)(d_cx + 3 * K);
    ra.node_offset = (const int32_t*)ifoff->ptr;
    ra.theta = (const double*)ifnode->ptr;
    ra.ipsi_offset = (const int64_t*)((const char*)ifoff->ptr + noff_bytes);
    ra.ipsi = (const double*)ifipsi->ptr;
    ra.xt = (const double*)p->rec_xt.ptr;
    ra.yt = (const double*)p->rec_yt.ptr;
    ra.Y = (const double*)p->rec_Y.ptr;
    ra.outcome = (const int8_t*)p->rec_outcome.ptr;
    ra.G = (double*)p->rows_G.ptr;
    rs->d_front = (const int32_t*)jobs->ptr + n;
    rs->d_B = (double*)p->rows_part.ptr;
    rs->d_stats = (spac_point_stats*)p->stats.ptr;
}

/* The solve table of every job at one (paths, h), one thread per job: what upload_solves
 * builds on the host for that node batch (RNG stream = job index, boundary jobs without paths
 * or work items, live jobs' records and items in job order), from the system's resident job
 * arrays.  Job k's items are those of live job live[k]: (k, c chunk, min((c + 1) chunk, paths)). */
__global__ void k_uniform_solves(int32_t n, const double* __restrict__ jxy, const double* __restrict__ jcirc,
                                 const int32_t* __restrict__ live, const int32_t* __restrict__ front, uint32_t paths,
                                 double h, double sqrth, uint32_t chunk, int32_t cps, double* __restrict__ x0,
                                 double* __restrict__ y0, double* __restrict__ hs, uint32_t* __restrict__ strm,
                                 uint32_t* __restrict__ ps, int64_t* __restrict__ off, double* __restrict__ circ,
                                 double* __restrict__ sq, int4* __restrict__ items) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n) return;
    const int64_t lv = live[k];
    off[k] = lv * (int64_t)paths;
    if (k == n) return;
    x0[k] = jxy[k];
    y0[k] = jxy[n + k];
    hs[k] = h;
    sq[k] = sqrth;
    strm[k] = (uint32_t)k;
    circ[3 * (int64_t)k] = jcirc[3 * (int64_t)k];
    circ[3 * (int64_t)k + 1] = jcirc[3 * (int64_t)k + 1];
    circ[3 * (int64_t)k + 2] = jcirc[3 * (int64_t)k + 2];
    if (front[k]) {
        ps[k] = 0u;
        return;
    }
    ps[k] = paths;
    for (int32_t c = 0; c < cps; ++c) {
        const uint32_t p0 = (uint32_t)c * chunk;
        items[lv * cps + c] = make_int4(k, (int)p0, (int)min(p0 + chunk, paths), 0);
    }
}

spac_status Problem::device_solves_uniform(const DeviceSystem* sys, uint32_t paths, double h, SolveSet* out) {
    if (paths < 1) return set_error(SPAC_E_INVALID_ARG, "paths must be >= 1");
    if (!(h > 0.0) || !(h < 1e30)) return set_error(SPAC_E_INVALID_ARG, "bad h");
    const int32_t n = sys->n_jobs;
    const int64_t total = (int64_t)sys->n_live * (int64_t)paths;
    const uint32_t chunk = work_item_paths((uint32_t)launch.chunk_paths, total);
    const int64_t cps = ((int64_t)paths + chunk - 1) / chunk;
    const int64_t n_items = cps * (int64_t)sys->n_live;
    if (n_items > kMaxWorkItems) return set_error(SPAC_E_INVALID_ARG, "too many work items");
    const size_t nb = (size_t)std::max(n, 1);
    spac_status rc;
    if ((rc = sx.ensure(nb * 8)) || (rc = sy.ensure(nb * 8)) || (rc = sh.ensure(nb * 8)) ||
        (rc = sstream.ensure(nb * 4)) || (rc = spaths.ensure(nb * 4)) || (rc = soff.ensure((nb + 1) * 8)) ||
        (rc = scircle.ensure(nb * 24)) || (rc = sitems.ensure((size_t)std::max<int64_t>(n_items, 1) * sizeof(int4))) ||
        (rc = ssq.ensure(nb * 8)))
        return rc;
    const int32_t* jobs = (const int32_t*)sys->t_jobs.ptr;
    hipLaunchKernelGGL(k_uniform_solves, dim3((n + 1 + 255) / 256), dim3(256), 0, stream, n,
                       (const double*)sys->u_xy.ptr, (const double*)sys->u_circle.ptr, (const int32_t*)sys->u_live.ptr,
                       jobs + n, paths, h, std::sqrt(h) /* correctly rounded, as upload_solves' */, chunk, (int32_t)cps,
                       (double*)sx.ptr, (double*)sy.ptr, (double*)sh.ptr, (uint32_t*)sstream.ptr, (uint32_t*)spaths.ptr,
                       (int64_t*)soff.ptr, (double*)scircle.ptr, (double*)ssq.ptr, (int4*)sitems.ptr);
    ROWS_HIP_CHECK(hipGetLastError());
    SolveSet ss;
    ss.n = n;
    ss.x0 = (const double*)sx.ptr;
    ss.y0 = (const double*)sy.ptr;
    ss.h = (const double*)sh.ptr;
    ss.stream = (const uint32_t*)sstream.ptr;
    ss.paths = (const uint32_t*)spaths.ptr;
    ss.rec_off = (const int64_t*)soff.ptr;
    ss.circle = (const double*)scircle.ptr;
    ss.items = (const int4*)sitems.ptr;
    ss.n_items = (int32_t)n_items;
    ss.total_paths = total;
    ss.chunk = chunk;
    ss.sqrth = (const double*)ssq.ptr;
    if ((rc = ensure_solve_scratch(n, &ss))) return rc;
    *out = ss;
    return SPAC_OK;
}

/* rows_prepare for the batch of every job of sys's mesh at one (paths, h): the solve table is
 * built on the device (Problem::device_solves_uniform) and the row kernels read the system's
 * interface table and job arrays, so the host touches nothing per job. */
static spac_status rows_prepare_uniform(Problem* p, const DeviceSystem* sys, uint32_t paths, double h,
                                        RowsSetup* rs) {
    const int n = sys->n_jobs;
    rs->n = n;
    rs->mmax = sys->mmax;
    if (sys->rbf_kind == SPAC_RBF_GAUSS_STENCIL) {
        rs->ra.stencil_size = sys->stencil_size;
        int lds_max = 0;
        ROWS_HIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, p->device));
        const size_t need = stencil_lds_bytes(*rs) + 5 * 1024; /* + the kernel's static LDS */
        if (need > (size_t)lds_max)
            return set_error(SPAC_E_INVALID_ARG, "interfaces of %d nodes need %zu B of LDS per block (device: %d)",
                             rs->mmax, need, lds_max);
        rs->g_block = (int64_t)paths <= kMomentSmallRow ? 64 : 256;
    }
    spac_status rc;
    if ((rc = p->device_solves_uniform(sys, paths, h, &rs->ss))) return rc;
    if ((rc = p->rows_G.ensure((size_t)std::max<int64_t>(sys->g_total, 1) * 8))) return rc;
    if ((rc = p->rows_part.ensure((size_t)n * 8))) return rc; /* B */
    if ((rc = p->stats.ensure((size_t)n * sizeof(spac_point_stats)))) return rc;
    if ((rc = p->ensure_records(rs->ss.total_paths))) return rc;
    if ((rc = p->queue.ensure(64))) return rc;
    fill_rows_args(p, rs, sys->n_ifaces, sys->rbf_kind, sys->stencil_size, sys->gauss_eps, sys->t_ifc, sys->t_ifoff,
                   sys->t_ifnode, sys->t_ifipsi, sys->t_jobs, sys->t_goff);
    return SPAC_OK;
}

/* B (boundary nodes: g at the node; others: the estimator's B) and G, no synchronisation. */
static spac_status rows_enqueue_bg(Problem* p, const RowsSetup& rs) {
    hipStream_t s = p->stream;
    const int n = rs.n;
    const int nblk = (n + 255) / 256;
    if (p->tab_mask) {
        spac_coef_table tab;
        std::memset(&tab, 0, sizeof tab);
        tab.v = (const double*)p->tab.ptr;
        tab.t = p->tab_t;
        tab.mask = p->tab_mask;
        tab.lox = p->tab_dom[0];
        tab.loy = p->tab_dom[1];
        tab.ix = (double)(p->tab_t - 1) / (p->tab_dom[2] - p->tab_dom[0]);
        tab.iy = (double)(p->tab_t - 1) / (p->tab_dom[3] - p->tab_dom[1]);
        hipLaunchKernelGGL(fk_rows_b_tab_kernel, dim3(nblk), dim3(256), 0, s, n, rs.d_front, rs.ss.x0, rs.ss.y0,
                           (const spac_point_stats*)rs.d_stats, rs.d_B, tab);
    } else if (p->desc.coef_kind != SPAC_COEF_VARCOEF_FRANKE) /* Franke g for both Franke sets */
        hipLaunchKernelGGL(fk_rows_b_kernel<CoefFranke>, dim3(nblk), dim3(256), 0, s, n, rs.d_front, rs.ss.x0, rs.ss.y0,
                           (const spac_point_stats*)rs.d_stats, rs.d_B);
    else
        hipLaunchKernelGGL(fk_rows_b_kernel<CoefVarFranke>, dim3(nblk), dim3(256), 0, s, n, rs.d_front, rs.ss.x0,
                           rs.ss.y0, (const spac_point_stats*)rs.d_stats, rs.d_B);
    ROWS_HIP_CHECK(hipGetLastError());
    if (rs.ra.rbf_kind == SPAC_RBF_GAUSS_STENCIL) {
        const size_t lds = stencil_lds_bytes(rs);
        if (rs.g_block == 64)
            hipLaunchKernelGGL(fk_rows_g_moment_kernel<64>, dim3(n), dim3(64), lds, s, rs.ra);
        else
            hipLaunchKernelGGL(fk_rows_g_moment_kernel<256>, dim3(n), dim3(256), lds, s, rs.ra);
    } else {
        hipLaunchKernelGGL(fk_rows_g_imq_kernel, dim3(n), dim3(kRowsBlock), 0, s, rs.ra);
    }
    ROWS_HIP_CHECK(hipGetLastError());
    return SPAC_OK;
}

/* Copy the rows out and fold the per-node statuses. */
static spac_status rows_finish(Problem* p, const RowsSetup& rs, spac_row_block* out) {
    const int n = rs.n;
    for (int k = 0; k <= n; ++k) out->G_offset[k] = rs.goff[k];
    if (!rs.gcol.empty()) std::memcpy(out->G_col, rs.gcol.data(), rs.gcol.size() * 4);
    spac_status rc;
    if ((rc = p->copy_out(out->B, rs.d_B, (size_t)n * 8))) return rc;
    if ((rc = p->copy_out(out->G, p->rows_G.ptr, (size_t)rs.goff[n] * 8))) return rc;
    std::vector<spac_point_stats> st(n);
    if ((rc = p->copy_out(st.data(), rs.d_stats, (size_t)n * sizeof(spac_point_stats)))) return rc;
    if (out->stats) std::memcpy(out->stats, st.data(), (size_t)n * sizeof(spac_point_stats));
    spac_status worst = SPAC_OK;
    for (int k = 0; k < n; ++k)
        if (st[k].status < worst) worst = st[k].status;
    if (worst == SPAC_E_NAN) return set_error(SPAC_E_NAN, "NaN reached a row estimator");
    if (worst == SPAC_E_MAX_STEPS) return set_error(SPAC_E_MAX_STEPS, "paths stopped by max_steps");
    return SPAC_OK;
}

} // namespace spac

using namespace spac;

extern "C" spac_status spac_fk_assemble_rows(spac_problem* hp, const spac_interface_table* t,
                                             const spac_node_batch* nb, uint64_t seed, spac_row_block* out) {
    if (!hp || !out || !out->B || !out->G || !out->G_col || !out->G_offset)
        return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (spac_status pr = no_pending(reinterpret_cast<Problem*>(hp))) return pr;
    spac_status rc = validate(t, nb);
    if (rc) return rc;
    Problem* p = reinterpret_cast<Problem*>(hp);
    RowsSetup rs;
    if ((rc = rows_prepare(p, t, nb, &rs))) return rc;
    hipStream_t s = p->stream;
    int blocks = 0;
    ROWS_HIP_CHECK(hipEventRecord(p->ev[0], s));
    if ((rc = p->enqueue_paths(rs.ss, seed, nullptr, &blocks, p->ev[1]))) return rc;
    if ((rc = p->enqueue_reduce(rs.ss, rs.d_stats, 1))) return rc;
    ROWS_HIP_CHECK(hipEventRecord(p->ev[2], s));
    if ((rc = rows_enqueue_bg(p, rs))) return rc;
    ROWS_HIP_CHECK(hipEventRecord(p->ev[3], s));
    ROWS_HIP_CHECK(hipStreamSynchronize(s));
    float ms_path = 0.f, ms_rest = 0.f, ms_all = 0.f;
    ROWS_HIP_CHECK(hipEventElapsedTime(&ms_path, p->ev[0], p->ev[1]));
    ROWS_HIP_CHECK(hipEventElapsedTime(&ms_rest, p->ev[1], p->ev[3]));
    ROWS_HIP_CHECK(hipEventElapsedTime(&ms_all, p->ev[0], p->ev[3]));
    p->last.path_kernel_ms = ms_path;
    p->last.reduce_kernel_ms = ms_rest; /* reduction, B and G kernels */
    p->last.total_ms = ms_all;          /* all device work of the call, first kernel to last */
    p->last.path_kernel_launches = 1;
    p->last.grid_blocks = blocks;
    p->last.path_steps = 0;
    return rows_finish(p, rs, out);
}

/* ------------------------------------------------ device-resident system -- */
namespace spac {
/* G of every job into its CSR slots: the slot's initial value plus the job's value (each slot
 * receives one job, checked at system creation), Actualiza_GB's single += per slot.  Lane 0 of
 * each job also folds the job's estimator status into *flag (the smallest, i.e. worst, code). */
__global__ void __launch_bounds__(64) k_system_g(int32_t n_jobs, const int64_t* __restrict__ g_off,
                                                 const int64_t* __restrict__ slot, const double* __restrict__ G,
                                                 const double* __restrict__ init, double* __restrict__ vals,
                                                 const spac_point_stats* __restrict__ st, int32_t* flag) {
    const int k = blockIdx.x;
    if (k >= n_jobs) return;
    if (threadIdx.x == 0) {
        const int32_t c = st[k].status;
        if (c < 0) atomicMin(flag, c);
    }
    const int64_t s = slot[k];
    if (s < 0) return;
    const int64_t o = g_off[k];
    const int m = (int)(g_off[k + 1] - o);
    for (int q = threadIdx.x; q < m; q += blockDim.x) vals[s + q] = init[s + q] + G[o + q];
}
/* B of every row: 0 + the B of its jobs, in job order (Actualiza_GB's B[fila] += b).  The grid's
 * last block clears *flag (launched before k_system_g, on the same stream). */
__global__ void k_system_b(int32_t n_rows, const int32_t* __restrict__ row_job_ptr, const int32_t* __restrict__ row_jobs,
                           const double* __restrict__ Bj, double* __restrict__ B, int32_t* flag) {
    if (blockIdx.x == gridDim.x - 1) {
        if (threadIdx.x == 0) *flag = SPAC_OK;
        return;
    }
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    double b = 0.0;
    for (int i = row_job_ptr[r]; i < row_job_ptr[r + 1]; ++i) b += Bj[row_jobs[i]];
    B[r] = b;
}
} // namespace spac

namespace spac {
/* Checks and enqueues one assembly of the system's jobs (path kernel, reduction, B and G
 * kernels, accumulation into the system); the call in flight is left in p->pending. */
static spac_status enqueue_system_kernels(Problem* p, const RowsSetup& rs, uint64_t seed, DeviceSystem* sys,
                                          spac_point_stats* stats);

static spac_status enqueue_assemble_system(Problem* p, const spac_interface_table* t, const spac_node_batch* nb,
                                           uint64_t seed, DeviceSystem* sys, spac_point_stats* stats) {
    spac_status rc = validate(t, nb);
    if (rc) return rc;
    if (nb->n != sys->n_jobs)
        return set_error(SPAC_E_INVALID_ARG, "the node batch has %d jobs, the system's mesh %d", nb->n, sys->n_jobs);
    /* the batch must be the mesh's jobs: each job's G goes into the slots validated for it */
    for (int k = 0; k < nb->n; ++k) {
        const int f = nb->iface[k];
        if (f != sys->h_job_iface[k] || (nb->frontera[k] ? 1 : 0) != sys->h_job_frontera[k] ||
            t->node_offset[f + 1] - t->node_offset[f] != sys->h_job_m[k])
            return set_error(SPAC_E_INVALID_ARG, "job %d of the batch is not the system mesh's job %d", k, k);
    }
    /* and the table the mesh's: the system holds that table's device copy */
    if (table_signature(t) != sys->h_table_sig)
        return set_error(SPAC_E_INVALID_ARG, "the interface table is not the system mesh's");
    if (p->device != sys->device) return set_error(SPAC_E_INVALID_ARG, "problem and system are on different devices");
    RowsSetup rs;
    if ((rc = rows_prepare(p, t, nb, &rs, sys))) return rc;
    return enqueue_system_kernels(p, rs, seed, sys, stats);
}

/* Every job of sys's mesh at one (paths, h): the same work with the solve table built on the
 * device. */
static spac_status enqueue_assemble_system_uniform(Problem* p, DeviceSystem* sys, uint32_t paths, double h,
                                                   uint64_t seed, spac_point_stats* stats) {
    if (p->device != sys->device) return set_error(SPAC_E_INVALID_ARG, "problem and system are on different devices");
    int dev = -1;
    ROWS_HIP_CHECK(hipGetDevice(&dev));
    if (dev != sys->device) return set_error(SPAC_E_INVALID_ARG, "the system's device is not the current device");
    RowsSetup rs;
    spac_status rc;
    if ((rc = rows_prepare_uniform(p, sys, paths, h, &rs))) return rc;
    return enqueue_system_kernels(p, rs, seed, sys, stats);
}

/* The kernels of one system assembly over a prepared batch (path kernel, reduction, B and G,
 * accumulation into the system); the call in flight is left in p->pending. */
static spac_status enqueue_system_kernels(Problem* p, const RowsSetup& rs, uint64_t seed, DeviceSystem* sys,
                                          spac_point_stats* stats) {
    spac_status rc;
    hipStream_t s = p->stream;
    int blocks = 0;
    ROWS_HIP_CHECK(hipEventRecord(p->ev[0], s));
    if ((rc = p->enqueue_paths(rs.ss, seed, nullptr, &blocks, p->ev[1]))) return rc;
    if ((rc = p->enqueue_reduce(rs.ss, rs.d_stats, 1))) return rc;
    ROWS_HIP_CHECK(hipEventRecord(p->ev[2], s));
    if ((rc = rows_enqueue_bg(p, rs))) return rc;
    const int n = rs.n;
    int32_t* d_flag = (int32_t*)sys->flag.ptr;
    /* slots no job adds into keep their initial values from system creation; every job slot is
       rewritten below as initial value + G, so no reset of the values is needed */
    /* k_system_b clears the status word, k_system_g folds the job statuses into it */
    hipLaunchKernelGGL(k_system_b, dim3((sys->n_rows + 255) / 256 + 1), dim3(256), 0, s, sys->n_rows,
                       (const int32_t*)sys->row_job_ptr.ptr, (const int32_t*)sys->row_jobs.ptr, (const double*)rs.d_B,
                       (double*)sys->B.ptr, d_flag);
    hipLaunchKernelGGL(k_system_g, dim3(n), dim3(64), 0, s, n, rs.ra.g_off, (const int64_t*)sys->job_slot.ptr,
                       (const double*)rs.ra.G, (const double*)sys->init.ptr, (double*)sys->vals.ptr,
                       (const spac_point_stats*)rs.d_stats, d_flag);
    ROWS_HIP_CHECK(hipGetLastError());
    ROWS_HIP_CHECK(hipEventRecord(p->ev[3], s));
    p->pending.kind = Problem::PENDING_SYSTEM;
    p->pending.n = n;
    p->pending.blocks = blocks;
    p->pending.sys = sys;
    p->pending.d_stats = rs.d_stats;
    p->pending.host_stats = stats;
    return SPAC_OK;
}

/* Waits for an enqueued system assembly: its status word, optional estimator records and
 * timing. */
spac_status finish_assemble_system(Problem* p, const Problem::Pending& pd) {
    hipStream_t s = p->stream;
    int32_t worst = SPAC_OK;
    ROWS_HIP_CHECK(hipMemcpyAsync(&worst, pd.sys->flag.ptr, 4, hipMemcpyDeviceToHost, s));
    ROWS_HIP_CHECK(hipStreamSynchronize(s));
    spac_status rc;
    if (pd.host_stats && (rc = p->copy_out(pd.host_stats, pd.d_stats, (size_t)pd.n * sizeof(spac_point_stats))))
        return rc;
    float ms_path = 0.f, ms_rest = 0.f, ms_all = 0.f;
    ROWS_HIP_CHECK(hipEventElapsedTime(&ms_path, p->ev[0], p->ev[1]));
    ROWS_HIP_CHECK(hipEventElapsedTime(&ms_rest, p->ev[1], p->ev[3]));
    ROWS_HIP_CHECK(hipEventElapsedTime(&ms_all, p->ev[0], p->ev[3]));
    p->last.path_kernel_ms = ms_path;
    p->last.reduce_kernel_ms = ms_rest; /* reduction, B and G kernels, accumulation */
    p->last.total_ms = ms_all;
    p->last.path_kernel_launches = 1;
    p->last.grid_blocks = pd.blocks;
    p->last.path_steps = 0;
    if (worst == SPAC_E_NAN) return set_error(SPAC_E_NAN, "NaN reached a row estimator");
    if (worst == SPAC_E_MAX_STEPS) return set_error(SPAC_E_MAX_STEPS, "paths stopped by max_steps");
    return SPAC_OK;
}
} // namespace spac

extern "C" spac_status spac_fk_assemble_system(spac_problem* hp, const spac_interface_table* t,
                                               const spac_node_batch* nb, uint64_t seed, spac_system* hsys,
                                               spac_point_stats* stats) {
    if (!hp || !hsys) return set_error(SPAC_E_INVALID_ARG, "null argument");
    Problem* p = reinterpret_cast<Problem*>(hp);
    spac_status rc = no_pending(p);
    if (rc) return rc;
    if ((rc = enqueue_assemble_system(p, t, nb, seed, reinterpret_cast<DeviceSystem*>(hsys), stats))) {
        p->pending = Problem::Pending(); /* nothing is left pending by a refused call */
        (void)hipStreamSynchronize(p->stream);
        return rc;
    }
    const Problem::Pending pd = p->pending;
    p->pending = Problem::Pending();
    return finish_assemble_system(p, pd);
}

extern "C" spac_status spac_fk_assemble_system_async(spac_problem* hp, const spac_interface_table* t,
                                                     const spac_node_batch* nb, uint64_t seed, spac_system* hsys,
                                                     spac_point_stats* stats) {
    if (!hp || !hsys) return set_error(SPAC_E_INVALID_ARG, "null argument");
    Problem* p = reinterpret_cast<Problem*>(hp);
    spac_status rc = no_pending(p);
    if (rc) return rc;
    if ((rc = enqueue_assemble_system(p, t, nb, seed, reinterpret_cast<DeviceSystem*>(hsys), stats))) {
        p->pending = Problem::Pending();
        (void)hipStreamSynchronize(p->stream);
        return rc;
    }
    return SPAC_OK;
}

extern "C" spac_status spac_fk_assemble_system_uniform(spac_problem* hp, spac_system* hsys, uint32_t paths, double h,
                                                       uint64_t seed, spac_point_stats* stats) {
    if (!hp || !hsys) return set_error(SPAC_E_INVALID_ARG, "null argument");
    Problem* p = reinterpret_cast<Problem*>(hp);
    spac_status rc = no_pending(p);
    if (rc) return rc;
    if ((rc = enqueue_assemble_system_uniform(p, reinterpret_cast<DeviceSystem*>(hsys), paths, h, seed, stats))) {
        p->pending = Problem::Pending();
        (void)hipStreamSynchronize(p->stream);
        return rc;
    }
    const Problem::Pending pd = p->pending;
    p->pending = Problem::Pending();
    return finish_assemble_system(p, pd);
}

extern "C" spac_status spac_fk_assemble_system_uniform_async(spac_problem* hp, spac_system* hsys, uint32_t paths,
                                                             double h, uint64_t seed, spac_point_stats* stats) {
    if (!hp || !hsys) return set_error(SPAC_E_INVALID_ARG, "null argument");
    Problem* p = reinterpret_cast<Problem*>(hp);
    spac_status rc = no_pending(p);
    if (rc) return rc;
    if ((rc = enqueue_assemble_system_uniform(p, reinterpret_cast<DeviceSystem*>(hsys), paths, h, seed, stats))) {
        p->pending = Problem::Pending();
        (void)hipStreamSynchronize(p->stream);
        return rc;
    }
    return SPAC_OK;
}

/* ------------------------------------------------------------- plans -- */
namespace spac {
struct RowsPlan {
    Problem* p = nullptr; /* private problem: its buffers are what the graph captured */
    RowsSetup rs;
    hipGraphExec_t exec = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    uint64_t seed_host = 0;
    double last_ms = 0.0;
    ~RowsPlan() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (p) spac_fk_problem_destroy(reinterpret_cast<spac_problem*>(p));
    }
};
} // namespace spac

extern "C" spac_status spac_fk_plan_create(spac_problem* hp, const spac_interface_table* t, const spac_node_batch* nb,
                                           spac_rows_plan** out) {
    if (!hp || !out) return set_error(SPAC_E_INVALID_ARG, "null argument");
    *out = nullptr;
    spac_status rc = validate(t, nb);
    if (rc) return rc;
    Problem* src = reinterpret_cast<Problem*>(hp);
    if ((rc = no_pending(src))) return rc;
    spac_problem* hclone = nullptr;
    if ((rc = spac_fk_problem_create(&src->desc, &hclone))) return rc;
    RowsPlan* plan = new RowsPlan();
    plan->p = reinterpret_cast<Problem*>(hclone);
    Problem* p = plan->p;
    p->launch = src->launch;
    if (src->cv_mode == SPAC_CV_TABLE || src->cv_mode == SPAC_CV_NETWORK) {
        spac_cv_config cfg;
        std::memset(&cfg, 0, sizeof cfg);
        cfg.mode = SPAC_CV_TABLE;
        cfg.fallback = src->cv_fallback;
        cfg.max_ratio = src->cv_max_ratio;
        cfg.table_size = src->cv_t;
        for (int i = 0; i < 4; ++i) cfg.table_domain[i] = src->cv_dom[i];
        rc = src->cv_mode == SPAC_CV_NETWORK
                 ? problem_set_cv_network(p, &cfg, src->cv_table.ptr, (const float*)src->cv_net.ptr, src->cv_net_map)
                 : problem_set_cv(p, &cfg, src->cv_table.ptr, true);
    } else {
        p->cv_mode = src->cv_mode;
        p->cv_fallback = src->cv_fallback;
        p->cv_max_ratio = src->cv_max_ratio;
    }
    if (!rc && src->tab_mask) { /* the tabulated coefficients, copied device to device */
        const size_t n = (size_t)src->tab_t * (size_t)src->tab_t * SPAC_TAB_STRIDE;
        rc = p->tab.ensure(n * 8);
        if (!rc) rc = p->tab32.ensure(n * 4);
        if (!rc && (hipMemcpyAsync(p->tab.ptr, src->tab.ptr, n * 8, hipMemcpyDeviceToDevice, p->stream) != hipSuccess ||
                    hipMemcpyAsync(p->tab32.ptr, src->tab32.ptr, n * 4, hipMemcpyDeviceToDevice, p->stream) != hipSuccess ||
                    hipStreamSynchronize(p->stream) != hipSuccess))
            rc = set_error(SPAC_E_HIP, "copying the coefficient tables into the plan failed");
        if (!rc) {
            p->tab_mask = src->tab_mask;
            p->tab_t = src->tab_t;
            for (int i = 0; i < 4; ++i) p->tab_dom[i] = src->tab_dom[i];
        }
    }
    if (!rc) rc = rows_prepare(p, t, nb, &plan->rs);
    if (!rc) rc = p->seed_dev.ensure(64);
    if (rc) {
        delete plan;
        return rc;
    }
    hipStream_t s = p->stream;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = p->enqueue_paths(plan->rs.ss, 0, (const uint64_t*)p->seed_dev.ptr, nullptr);
        if (!rc) rc = p->enqueue_reduce(plan->rs.ss, plan->rs.d_stats, 1);
        if (!rc) rc = rows_enqueue_bg(p, plan->rs);
        hipError_t e2 = hipStreamEndCapture(s, &g);
        if (!rc && e2 != hipSuccess) rc = set_error(SPAC_E_HIP, "graph capture failed: %s", hipGetErrorString(e2));
    } else {
        rc = set_error(SPAC_E_HIP, "hipStreamBeginCapture failed: %s", hipGetErrorString(e));
    }
    if (!rc) {
        e = hipGraphInstantiate(&plan->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = set_error(SPAC_E_HIP, "graph instantiate failed: %s", hipGetErrorString(e));
    }
    if (g) (void)hipGraphDestroy(g);
    if (!rc && (hipEventCreate(&plan->e0) != hipSuccess || hipEventCreate(&plan->e1) != hipSuccess))
        rc = set_error(SPAC_E_HIP, "hipEventCreate failed");
    if (rc) {
        delete plan;
        return rc;
    }
    *out = reinterpret_cast<spac_rows_plan*>(plan);
    return SPAC_OK;
}

extern "C" spac_status spac_fk_plan_execute(spac_rows_plan* hplan, uint64_t seed, spac_row_block* out,
                                            double* device_ms) {
    RowsPlan* plan = reinterpret_cast<RowsPlan*>(hplan);
    if (!plan || !out || !out->B || !out->G || !out->G_col || !out->G_offset)
        return set_error(SPAC_E_INVALID_ARG, "null argument");
    Problem* p = plan->p;
    hipStream_t s = p->stream;
    plan->seed_host = seed;
    ROWS_HIP_CHECK(hipMemcpyAsync(p->seed_dev.ptr, &plan->seed_host, 8, hipMemcpyHostToDevice, s));
    ROWS_HIP_CHECK(hipEventRecord(plan->e0, s));
    ROWS_HIP_CHECK(hipGraphLaunch(plan->exec, s));
    ROWS_HIP_CHECK(hipEventRecord(plan->e1, s));
    ROWS_HIP_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    ROWS_HIP_CHECK(hipEventElapsedTime(&ms, plan->e0, plan->e1));
    plan->last_ms = ms;
    if (device_ms) *device_ms = ms;
    return rows_finish(p, plan->rs, out);
}

extern "C" spac_status spac_fk_plan_destroy(spac_rows_plan* hplan) {
    RowsPlan* plan = reinterpret_cast<RowsPlan*>(hplan);
    if (!plan) return SPAC_OK;
    if (plan->p) (void)hipStreamSynchronize(plan->p->stream);
    delete plan;
    return SPAC_OK;
}

/* ---------------------------------------------------------- autotuner -- */
extern "C" spac_status spac_fk_autotune_rows(spac_problem* hp, const spac_interface_table* t,
                                             const spac_node_batch* nb, double path_fraction,
                                             spac_launch_config* tried, double* tried_ms, int32_t cap,
                                             spac_autotune_report* rep) {
    if (!hp || !rep) return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (!(path_fraction > 0.0) || path_fraction > 1.0)
        return set_error(SPAC_E_INVALID_ARG, "path_fraction must be in (0, 1]");
    spac_status rc = validate(t, nb);
    if (rc) return rc;
    Problem* p = reinterpret_cast<Problem*>(hp);
    if ((rc = no_pending(p))) return rc;
    /* sample: every node, paths scaled by path_fraction (at least one) */
    const int n = nb->n;
    std::vector<uint32_t> paths(n);
    std::vector<double> circle(3 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        int f = nb->iface[k];
        paths[k] = nb->frontera[k] ? 0u
                                   : (uint32_t)std::max<double>(1.0, std::floor((double)nb->paths[k] * path_fraction));
        circle[3 * k] = t->radius[f];
        circle[3 * k + 1] = t->cx[f];
        circle[3 * k + 2] = t->cy[f];
    }
    const spac_launch_config start = p->launch;
    int count = 0;
    auto measure = [&](const spac_launch_config& c, double* ms) -> spac_status {
        p->launch = c;
        SolveSet ss;
        spac_status r = p->upload_solves(n, nb->x, nb->y, paths.data(), nb->h, nb->stream, circle.data(), &ss);
        if (r) return r;
        double best = 1e300;
        for (int rep_i = 0; rep_i < 2; ++rep_i) {
            if ((r = p->run(ss, 0x5eedULL + (uint64_t)rep_i, nullptr, 1))) return r;
            best = std::min(best, p->last.path_kernel_ms);
        }
        if (tried && tried_ms && count < cap) {
            tried[count] = c;
            tried_ms[count] = best;
        }
        ++count;
        *ms = best;
        return SPAC_OK;
    };
    spac_launch_config best = start;
    double best_ms = 0.0;
    if ((rc = measure(best, &best_ms))) { p->launch = start; return rc; } /* warm-up */
    count = 0;
    if ((rc = measure(best, &best_ms))) { p->launch = start; return rc; }
    const double default_ms = best_ms;
    const int blocks[] = {64, 128, 256};
    const int bpc[] = {1, 2, 4, 8, 16};
    const int chunks[] = {32, 64, 128, 256, 512, 1024};
    const int orders[] = {SPAC_ORDER_NODE, SPAC_ORDER_COST};
    for (int pass = 0; pass < 2; ++pass) {
        for (int field = 0; field < 4; ++field) {
            const int* vals = field == 0 ? blocks : field == 1 ? bpc : field == 2 ? chunks : orders;
            const int nv = field == 0 ? 3 : field == 1 ? 5 : field == 2 ? 6 : 2;
            for (int i = 0; i < nv; ++i) {
                spac_launch_config c = best;
                int* slot = field == 0 ? &c.block_size : field == 1 ? &c.blocks_per_cu
                                                       : field == 2 ? &c.chunk_paths : &c.item_order;
                if (*slot == vals[i]) continue;
                *slot = vals[i];
                double ms = 0.0;
                if ((rc = measure(c, &ms))) { p->launch = start; return rc; }
                if (ms < best_ms) {
                    best_ms = ms;
                    best = c;
                }
            }
        }
    }
    p->launch = best;
    rep->n_candidates = count;
    rep->reserved = 0;
    rep->chosen = best;
    rep->best_ms = best_ms;
    rep->default_ms = default_ms;
    return SPAC_OK;
}
