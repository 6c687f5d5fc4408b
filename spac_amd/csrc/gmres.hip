/*
 * gmres.hip — restarted GMRES(m) with right Jacobi preconditioning on the GPU,
 * for the assembled interface system G u = B (the reference solves it with
 * SistemaLinealSolverGMRES, Solvers/SistemaLinealSolverGMRES_OMP.hpp:519-577).
 *
 * Device work per Arnoldi step: the CSR SpMV w = A z_k of z_k = D^-1 v_k (written
 * together with v_k; a few lanes per row) into the slot of v_{k+1}, classical
 * Gram–Schmidt (block dot products of w against all basis vectors, which also
 * give ||w||^2, then one fused update), the norm and the scaling of the new
 * basis vector.  Steps start with one Gram–Schmidt pass; the host watches a
 * DGKS-type test of every step (Daniel, Gragg, Kaufman, Stewart, Math. Comp. 30
 * (1976) 772-795: how much of ||w|| the pass keeps) and, from the first step
 * that keeps less than kKeepMin of it, queues every later step with the second
 * pass (CGS2).  The classical bound 1/sqrt(2) asks for the second pass on most
 * steps of the interface systems, where one pass already gives the two-pass
 * iteration counts and residuals; kKeepMin = 0.01 switches only on heavy
 * cancellation.  Convergence is only declared on the true residual, so a step
 * with a less orthogonal basis can cost iterations but not correctness.  The host
 * keeps the (m+1) x m Hessenberg matrix, applies the Givens rotations and
 * tests convergence, as the host solver does, a few steps behind the device:
 * the columns reach it through mapped pinned memory.  Every reduction has a
 * fixed order (per-block tree, blocks summed in index order), so a solve is
 * bit-reproducible.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <vector>

#include "../../include/spac_mesh.h"
#include "fk_internal.h"

namespace spac {

#define GM_CHECK(expr)                                                                         \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess) {                                                                \
            st.release();                                                                      \
            return set_error(SPAC_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                             __FILE__, __LINE__);                                              \
        }                                                                                      \
    } while (0)

constexpr int kDotBlock = 256;
constexpr int kDotPerThread = 8; /* rows per thread in the dot kernels */
constexpr int kDotChunk = kDotBlock * kDotPerThread;

/* inverse diagonal (1 where the diagonal is missing or zero) */
__global__ void k_diag_inv(int n, const int32_t* Ip, const int32_t* J, const double* V, double* dinv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 1.0;
    for (int k = Ip[i]; k < Ip[i + 1]; ++k)
        if (J[k] == i && V[k] != 0.0) d = 1.0 / V[k];
    dinv[i] = d;
}

/* y = A x with L lanes per row (64 / L rows per wavefront), each row summed in a fixed lane
 * order: lane `sub` adds entries sub, sub + L, sub + 2L, ... in that order, then the lanes are
 * combined by a fixed shuffle tree.  L is chosen from the mean row length so that a row takes
 * about three passes.  Four passes are issued together (column indices and values first, then
 * the gathers of x), so a row costs one dependent load round trip instead of one per pass. */
template <int L>
__global__ void __launch_bounds__(256) k_spmv(int n, const int32_t* __restrict__ Ip, const int32_t* __restrict__ J,
                                              const double* __restrict__ V, const double* __restrict__ x,
                                              double* __restrict__ y) {
    constexpr int U = 4;
    const int row = (int)((blockIdx.x * blockDim.x + threadIdx.x) / L);
    const int sub = (int)(threadIdx.x % L);
    if (row >= n) return; /* whole L-lane groups leave together */
    const int e = Ip[row + 1];
    double acc = 0.0;
    for (int k0 = Ip[row] + sub; k0 < e; k0 += U * L) {
        int jj[U];
        double vv[U], xx[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u * L;
            jj[u] = (k < e) ? J[k] : -1;
            vv[u] = (k < e) ? V[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) xx[u] = (jj[u] >= 0) ? x[jj[u]] : 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u)
            if (k0 + u * L < e) acc = fma(vv[u], xx[u], acc);
    }
    for (int off = L / 2; off >= 1; off >>= 1) acc += __shfl_down(acc, off, L);
    if (sub == 0) y[row] = acc;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
    return t;
}

/* part[j][chunk] = sum over the chunk's rows of Vb[j][i] w[i], j = 0..nv-1 (grid: chunks x nv) */
__global__ void __launch_bounds__(kDotBlock) k_multidot(int n, const double* __restrict__ Vb, size_t ld,
                                                        const double* __restrict__ w, double* __restrict__ part,
                                                        int nchunks) {
    __shared__ double sh[kDotBlock / 64];
    const int chunk = blockIdx.x, j = blockIdx.y;
    const double* v = Vb + (size_t)j * ld;
    double acc = 0.0;
    const int base = chunk * kDotChunk + threadIdx.x;
#pragma unroll
    for (int r = 0; r < kDotPerThread; ++r) {
        const int i = base + r * kDotBlock;
        if (i < n) acc = fma(v[i], w[i], acc);
    }
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) part[(size_t)j * nchunks + chunk] = t;
}

/* h[j] = sum over chunks in order of part[j][chunk] (root: its square root, a norm);
 * one wavefront per j.  mirror (host-mapped pinned memory, or nullptr) receives a copy
 * for the host's Hessenberg column. */
__global__ void k_finish_dots(int nv, int nchunks, const double* __restrict__ part, double* __restrict__ h,
                              double* __restrict__ mirror, int root) {
    const int j = blockIdx.x;
    const int lane = threadIdx.x;
    if (j >= nv) return;
    double acc = 0.0;
    for (int c = lane; c < nchunks; c += 64) acc += part[(size_t)j * nchunks + c];
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_down(acc, off, 64);
    if (lane == 0) {
        const double v = root ? sqrt(acc) : acc;
        h[j] = v;
        if (mirror) mirror[j] = v;
    }
}

/* w[i] -= sum_j h[j] Vb[j][i] (j = 0..nv-1), h staged in LDS */
__global__ void __launch_bounds__(256) k_update(int n, const double* __restrict__ Vb, size_t ld,
                                                const double* __restrict__ hdev, int nv, double* __restrict__ w) {
    __shared__ double hs[1024];
    for (int j = threadIdx.x; j < nv; j += blockDim.x) hs[j] = hdev[j];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < nv; ++j) s = fma(hs[j], Vb[(size_t)j * ld + i], s);
    w[i] -= s;
}

/* out = x / *den (the next Krylov vector) and z = dinv .* out (the SpMV input of the next
 * Arnoldi step, right preconditioning) */
__global__ void k_scale(int n, const double* __restrict__ x, const double* __restrict__ den,
                        const double* __restrict__ dinv, double* __restrict__ out, double* __restrict__ z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double v = x[i] / den[0];
    out[i] = v;
    z[i] = dinv[i] * v;
}

/* ------------------------------------------------ coarse space (aggregates) -- */
constexpr double kFix = 1099511627776.0; /* 2^40: fixed-point scale of the Galerkin sums */

__global__ void k_fill_u64(size_t n, unsigned long long* x) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = 0ull;
}

/* efix[agg(i)][agg(j)] += A_ij in fixed point.  Integer sums are exact, so the result does not
 * depend on the order of the atomics; a row's consecutive columns of one aggregate are summed
 * in a register first. */
__global__ void k_coarse_galerkin(int n, const int32_t* __restrict__ Ip, const int32_t* __restrict__ J,
                                  const double* __restrict__ V, const int32_t* __restrict__ agg, int k,
                                  unsigned long long* efix) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long* row = efix + (size_t)agg[i] * k;
    int cur = -1;
    long long acc = 0;
    for (int j = Ip[i]; j < Ip[i + 1]; ++j) {
        const int b = agg[J[j]];
        if (b != cur) {
            if (cur >= 0) atomicAdd(row + cur, (unsigned long long)acc);
            cur = b;
            acc = 0;
        }
        acc += __double2ll_rn(V[j] * kFix);
    }
    if (cur >= 0) atomicAdd(row + cur, (unsigned long long)acc);
}

/* cs[a] = sum of x over the rows of aggregate a (rows ascending, lane-strided, fixed tree) */
__global__ void __launch_bounds__(256) k_coarse_restrict(const int32_t* __restrict__ ptr,
                                                         const int32_t* __restrict__ rows,
                                                         const double* __restrict__ x, double* __restrict__ cs) {
    __shared__ double sh[4];
    const int a = blockIdx.x;
    double acc = 0.0;
    for (int q = ptr[a] + (int)threadIdx.x; q < ptr[a + 1]; q += 256) acc += x[rows[q]];
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) cs[a] = t;
}

/* cc = E^-1 cs (einv row-major) */
__global__ void k_coarse_solve(int k, const double* __restrict__ einv, const double* __restrict__ cs,
                               double* __restrict__ cc) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= k) return;
    double t = 0.0;
    for (int b = 0; b < k; ++b) t = fma(einv[(size_t)a * k + b], cs[b], t);
    cc[a] = t;
}

/* z[i] += (E^-1 cs)[agg(i)], the coarse value recomputed by every row of the aggregate with
 * k_coarse_solve's FMA order (rows of a wave share aggregates, so the loads of a row of E^-1
 * are contiguous broadcasts) */
__global__ void k_coarse_prolong(int n, const int32_t* __restrict__ agg, int k, const double* __restrict__ einv,
                                 const double* __restrict__ cs, double* __restrict__ z) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* e = einv + (size_t)agg[i] * k;
    double t = 0.0;
    for (int b = 0; b < k; ++b) t = fma(e[b], cs[b], t);
    z[i] += t;
}

/* t[i] = sum_j y[j] Vb[j][i] (the Krylov combination, before preconditioning) */
__global__ void __launch_bounds__(256) k_combine(int n, const double* __restrict__ Vb, size_t ld,
                                                 const double* __restrict__ y, int k, double* __restrict__ t) {
    __shared__ double ys[1024];
    for (int j = threadIdx.x; j < k; j += blockDim.x) ys[j] = y[j];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < k; ++j) s = fma(ys[j], Vb[(size_t)j * ld + i], s);
    t[i] = s;
}

/* u[i] += dinv[i] t[i] + cc[agg(i)]: u += M^-1 t with the two-level M^-1 */
__global__ void k_correct_coarse(int n, const double* __restrict__ t, const double* __restrict__ dinv,
                                 const int32_t* __restrict__ agg, const double* __restrict__ cc,
                                 double* __restrict__ u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] += dinv[i] * t[i] + cc[agg[i]];
}

/* In-place inverse of the k x k row-major matrix a by Gauss-Jordan elimination with partial
 * pivoting, written row-major into out.  False if a pivot vanishes. */
static bool invert_matrix(int k, std::vector<double>& a, std::vector<double>& out) {
    std::vector<double> inv((size_t)k * k, 0.0);
    for (int i = 0; i < k; ++i) inv[(size_t)i * k + i] = 1.0;
    for (int c = 0; c < k; ++c) {
        int piv = c;
        for (int r = c + 1; r < k; ++r)
            if (std::fabs(a[(size_t)r * k + c]) > std::fabs(a[(size_t)piv * k + c])) piv = r;
        if (!(std::fabs(a[(size_t)piv * k + c]) > 0.0)) return false;
        if (piv != c)
            for (int j = 0; j < k; ++j) {
                std::swap(a[(size_t)c * k + j], a[(size_t)piv * k + j]);
                std::swap(inv[(size_t)c * k + j], inv[(size_t)piv * k + j]);
            }
        const double d = 1.0 / a[(size_t)c * k + c];
        for (int j = 0; j < k; ++j) {
            a[(size_t)c * k + j] *= d;
            inv[(size_t)c * k + j] *= d;
        }
        for (int r = 0; r < k; ++r) {
            if (r == c) continue;
            const double f = a[(size_t)r * k + c];
            if (f == 0.0) continue;
            double* ar = &a[(size_t)r * k];
            double* ir = &inv[(size_t)r * k];
            const double* ac = &a[(size_t)c * k];
            const double* ic = &inv[(size_t)c * k];
            for (int j = 0; j < k; ++j) {
                ar[j] -= f * ac[j];
                ir[j] -= f * ic[j];
            }
        }
    }
    out.swap(inv);
    return true;
}

/* r = B - r */
__global__ void k_residual(int n, const double* __restrict__ B, double* __restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = B[i] - r[i];
}

/* u[i] += dinv[i] sum_j y[j] Vb[j][i] */
__global__ void __launch_bounds__(256) k_correct(int n, const double* __restrict__ Vb, size_t ld,
                                                 const double* __restrict__ y, int k, const double* __restrict__ dinv,
                                                 double* __restrict__ u) {
    __shared__ double ys[1024];
    for (int j = threadIdx.x; j < k; j += blockDim.x) ys[j] = y[j];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int j = 0; j < k; ++j) s = fma(ys[j], Vb[(size_t)j * ld + i], s);
    u[i] += dinv[i] * s;
}

/* Arnoldi steps queued ahead of the host's Givens processing: the host waits for step k's
 * Hessenberg column while steps k + 1 .. k + kAhead run, so the GPU never idles on a host
 * round trip.  Steps queued past convergence only add basis vectors that are not used. */
constexpr int kAhead = 3;
constexpr double kKeepMin = 0.01; /* one-pass steps must keep this share of ||w|| */
constexpr int kEvents = kAhead + 2;

struct GmresState {
    DeviceBuffer Ip, J, V, B, u, dinv, Vb, r, part, h, nrm, z;
    double* col = nullptr;     /* pinned, host-mapped: one Hessenberg column block per Arnoldi step */
    double* col_dev = nullptr; /* its device address */
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipEvent_t step_done[kEvents] = {};
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (DeviceBuffer* b : {&Ip, &J, &V, &B, &u, &dinv, &Vb, &r, &part, &h, &nrm, &z}) b->release();
        if (col) (void)hipHostFree(col);
        col = col_dev = nullptr;
        for (hipEvent_t& e : step_done) {
            if (e) (void)hipEventDestroy(e);
            e = nullptr;
        }
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        if (stream) (void)hipStreamDestroy(stream);
        e0 = e1 = nullptr;
        stream = nullptr;
    }
    ~GmresState() { release(); } /* error returns leave nothing behind */
};

} // namespace spac

using namespace spac;

/* The solve on either a host CSR (validated, uploaded into the state's buffers) or a
 * device-resident system (sys: its arrays are used in place; its pattern came from a mesh). */
static spac_status gmres_impl(int32_t n, const int32_t* Ip, const int32_t* J, const double* V, const double* B,
                              const DeviceSystem* sys, double* u, double tol, int32_t max_iter, int32_t restart,
                              spac_gmres_info* info) {
    if (sys) n = sys->n_rows;
    if (n < 1 || !u || (!sys && (!Ip || !J || !V || !B))) return set_error(SPAC_E_INVALID_ARG, "bad GMRES arguments");
    if (restart < 1) restart = 50;
    if (restart > 1000) return set_error(SPAC_E_INVALID_ARG, "restart must be <= 1000");
    if (max_iter < 1) max_iter = 1000;
    int64_t nnz = 0;
    if (sys) {
        nnz = sys->nnz;
    } else {
        if (Ip[0] != 0 || Ip[n] < 0) return set_error(SPAC_E_INVALID_ARG, "CSR row pointer must start at 0");
        for (int32_t i = 0; i < n; ++i)
            if (Ip[i + 1] < Ip[i]) return set_error(SPAC_E_INVALID_ARG, "CSR row pointer decreases at row %d", i);
        nnz = Ip[n];
        /* column range: a branch-free min / max pass (vectorised), the offending index only on error */
        int32_t jmin = 0, jmax = 0;
        if (nnz > 0) {
            jmin = J[0];
            jmax = J[0];
            for (int64_t k = 1; k < nnz; ++k) {
                jmin = std::min(jmin, J[k]);
                jmax = std::max(jmax, J[k]);
            }
        }
        if (jmin < 0 || jmax >= n)
            for (int64_t k = 0; k < nnz; ++k)
                if (J[k] < 0 || J[k] >= n)
                    return set_error(SPAC_E_INVALID_ARG, "CSR column %d out of range at %lld", J[k], (long long)k);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_error(SPAC_E_NO_DEVICE, "no HIP device visible");
    const auto t_start = std::chrono::steady_clock::now();
    const int m = restart;
    const size_t N = (size_t)n;
    const int nchunks = (int)((N + kDotChunk - 1) / kDotChunk);
    /* host-visible column block: h1[m+1] | h2[m+1] | ||w|| */
    const size_t hlen = 2 * (size_t)(m + 1) + 1;
    GmresState st;
    spac_status rc;
    if ((!sys && ((rc = st.Ip.ensure((N + 1) * 4)) || (rc = st.J.ensure((size_t)nnz * 4 + 4)) ||
                  (rc = st.V.ensure((size_t)nnz * 8 + 8)) || (rc = st.B.ensure(N * 8)))) ||
        (rc = st.u.ensure(N * 8)) || (rc = st.dinv.ensure(N * 8)) || (rc = st.Vb.ensure((size_t)(m + 1) * N * 8)) ||
        (rc = st.r.ensure(N * 8)) || (rc = st.part.ensure((size_t)(m + 1) * nchunks * 8)) ||
        (rc = st.h.ensure(hlen * 8)) || (rc = st.nrm.ensure(64)) || (rc = st.z.ensure(N * 8))) {
        st.release();
        return rc;
    }
    GM_CHECK(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
    GM_CHECK(hipEventCreate(&st.e0));
    GM_CHECK(hipEventCreate(&st.e1));
    for (hipEvent_t& e : st.step_done) GM_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    GM_CHECK(hipHostMalloc((void**)&st.col, (size_t)m * hlen * 8, hipHostMallocMapped));
    GM_CHECK(hipHostGetDevicePointer((void**)&st.col_dev, st.col, 0));
    hipStream_t s = st.stream;
    if (!sys) {
        GM_CHECK(hipMemcpyAsync(st.Ip.ptr, Ip, (N + 1) * 4, hipMemcpyHostToDevice, s));
        GM_CHECK(hipMemcpyAsync(st.J.ptr, J, (size_t)nnz * 4, hipMemcpyHostToDevice, s));
        GM_CHECK(hipMemcpyAsync(st.V.ptr, V, (size_t)nnz * 8, hipMemcpyHostToDevice, s));
        GM_CHECK(hipMemcpyAsync(st.B.ptr, B, N * 8, hipMemcpyHostToDevice, s));
    }
    GM_CHECK(hipMemcpyAsync(st.u.ptr, u, N * 8, hipMemcpyHostToDevice, s));
    GM_CHECK(hipEventRecord(st.e0, s));

    const int32_t* dIp = (const int32_t*)(sys ? sys->Ip.ptr : st.Ip.ptr);
    const int32_t* dJ = (const int32_t*)(sys ? sys->J.ptr : st.J.ptr);
    const double* dV = (const double*)(sys ? sys->vals.ptr : st.V.ptr);
    const double* dB = (const double*)(sys ? sys->B.ptr : st.B.ptr);
    double* du = (double*)st.u.ptr;
    double* dinv = (double*)st.dinv.ptr;
    double* Vb = (double*)st.Vb.ptr;
    double* r = (double*)st.r.ptr;
    double* part = (double*)st.part.ptr;
    double* h1 = (double*)st.h.ptr;
    double* h2 = h1 + (m + 1);
    double* hw = h2 + (m + 1);
    double* dn = (double*)st.nrm.ptr;
    double* dz = (double*)st.z.ptr;
    /* lanes per SpMV row: about three passes over the mean row */
    const double mean_row = (double)nnz / (double)n;
    const int row_lanes = mean_row > 96.0 ? 64 : mean_row > 48.0 ? 32 : mean_row > 24.0 ? 16 : 8;
    auto spmv = [&](const double* x, double* y) {
        const dim3 grid((unsigned)(((size_t)n * row_lanes + 255) / 256));
        if (row_lanes == 64) hipLaunchKernelGGL(k_spmv<64>, grid, dim3(256), 0, s, n, dIp, dJ, dV, x, y);
        else if (row_lanes == 32) hipLaunchKernelGGL(k_spmv<32>, grid, dim3(256), 0, s, n, dIp, dJ, dV, x, y);
        else if (row_lanes == 16) hipLaunchKernelGGL(k_spmv<16>, grid, dim3(256), 0, s, n, dIp, dJ, dV, x, y);
        else hipLaunchKernelGGL(k_spmv<8>, grid, dim3(256), 0, s, n, dIp, dJ, dV, x, y);
    };
    const dim3 g1((n + 255) / 256), b1(256);

    auto dots = [&](const double* vecs, int nv, const double* x, double* out, double* mirror, int root) {
        hipLaunchKernelGGL(k_multidot, dim3(nchunks, nv), dim3(kDotBlock), 0, s, n, vecs, N, x, part, nchunks);
        hipLaunchKernelGGL(k_finish_dots, dim3(nv), dim3(64), 0, s, nv, nchunks, (const double*)part, out, mirror,
                           root);
    };
    auto norm = [&](const double* x, double* out) { dots(x, 1, x, out, (double*)nullptr, 1); };

    hipLaunchKernelGGL(k_diag_inv, g1, b1, 0, s, n, dIp, dJ, dV, dinv);
    /* two-level preconditioner of a device system: Galerkin matrix on the device, inverse on the host */
    const int ck = sys ? sys->coarse_k : 0;
    const int32_t* cagg = ck ? (const int32_t*)sys->c_agg.ptr : nullptr;
    double* ccs = ck ? (double*)sys->c_vec.ptr : nullptr;
    double* ccc = ccs + ck;
    if (ck > 0) {
        const size_t kk = (size_t)ck * ck;
        unsigned long long* efix = (unsigned long long*)sys->c_efix.ptr;
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((kk + 255) / 256)), b1, 0, s, kk, efix);
        hipLaunchKernelGGL(k_coarse_galerkin, g1, b1, 0, s, n, dIp, dJ, dV, cagg, ck, efix);
        std::vector<long long> ef(kk);
        GM_CHECK(hipMemcpyAsync(ef.data(), efix, kk * 8, hipMemcpyDeviceToHost, s));
        GM_CHECK(hipStreamSynchronize(s));
        std::vector<double> e(kk), einv;
        for (size_t q = 0; q < kk; ++q) e[q] = (double)ef[q] / kFix;
        if (!invert_matrix(ck, e, einv)) {
            st.release();
            return set_error(SPAC_E_INVALID_ARG, "the coarse Galerkin matrix is singular");
        }
        GM_CHECK(hipMemcpyAsync(sys->c_einv.ptr, einv.data(), kk * 8, hipMemcpyHostToDevice, s));
        GM_CHECK(hipStreamSynchronize(s));
    }
    /* z += Z E^-1 Z^T x (after z = D^-1 x) */
    auto coarse = [&](const double* x, double* z) {
        hipLaunchKernelGGL(k_coarse_restrict, dim3(ck), b1, 0, s, (const int32_t*)sys->c_ptr.ptr,
                           (const int32_t*)sys->c_rows.ptr, x, ccs);
        hipLaunchKernelGGL(k_coarse_prolong, g1, b1, 0, s, n, cagg, ck, (const double*)sys->c_einv.ptr,
                           (const double*)ccs, z);
    };
    norm(dB, dn);
    double bnorm = 0.0;
    GM_CHECK(hipMemcpyAsync(&bnorm, dn, 8, hipMemcpyDeviceToHost, s));
    GM_CHECK(hipStreamSynchronize(s));
    if (bnorm == 0.0) bnorm = 1.0;

    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), yv(m);
    std::vector<char> pass2(m, 0); /* whether step k of the cycle took the second Gram–Schmidt pass */
    bool two_pass = false;
    int it = 0, restarts = 0;
    double res = 0.0;
    bool conv = false;
    auto true_residual = [&](double* out) -> spac_status {
        spmv((const double*)du, r);
        hipLaunchKernelGGL(k_residual, g1, b1, 0, s, n, dB, r);
        norm(r, dn);
        GM_CHECK(hipMemcpyAsync(out, dn, 8, hipMemcpyDeviceToHost, s));
        GM_CHECK(hipStreamSynchronize(s));
        return SPAC_OK;
    };
    /* Arnoldi step k: w = A D^-1 v_k in the slot of v_{k+1}, Gram–Schmidt against v_0..v_k
       (once, or twice when two_pass), v_{k+1} = w / ||w||.  The coefficient vectors, ||w||^2
       before the first pass (h1[k + 1]) and the final norm also go to the step's pinned
       column block. */
    auto enqueue_step = [&](int k, bool two_pass) -> spac_status {
        double* colk = st.col_dev + (size_t)k * hlen;
        double* w = Vb + (size_t)(k + 1) * N;
        spmv((const double*)dz, w);
        const int nv = k + 1;
        dots(Vb, nv + 1, w, h1, colk, 0);
        hipLaunchKernelGGL(k_update, g1, b1, 0, s, n, (const double*)Vb, N, (const double*)h1, nv, w);
        if (two_pass) {
            dots(Vb, nv, w, h2, colk + (m + 1), 0);
            hipLaunchKernelGGL(k_update, g1, b1, 0, s, n, (const double*)Vb, N, (const double*)h2, nv, w);
        }
        dots(w, 1, w, hw, colk + 2 * (m + 1), 1);
        hipLaunchKernelGGL(k_scale, g1, b1, 0, s, n, (const double*)w, (const double*)hw, (const double*)dinv, w, dz);
        if (ck > 0) coarse((const double*)w, dz);
        GM_CHECK(hipGetLastError());
        GM_CHECK(hipEventRecord(st.step_done[k % kEvents], s));
        return SPAC_OK;
    };
    while (it < max_iter) {
        double beta = 0.0;
        if ((rc = true_residual(&beta))) return rc;
        res = beta / bnorm;
        if (res < tol) { conv = true; break; }
        hipLaunchKernelGGL(k_scale, g1, b1, 0, s, n, (const double*)r, (const double*)dn, (const double*)dinv, Vb, dz);
        if (ck > 0) coarse((const double*)Vb, dz);
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        const int steps = std::min(m, max_iter - it); /* Arnoldi steps this cycle may take */
        int k = 0, queued = 0;
        while (k < steps) {
            while (queued < steps && queued - k <= kAhead) {
                pass2[queued] = two_pass ? 1 : 0;
                if ((rc = enqueue_step(queued, two_pass))) return rc;
                ++queued;
            }
            GM_CHECK(hipEventSynchronize(st.step_done[k % kEvents]));
            const double* hc = st.col + (size_t)k * hlen;
            for (int j = 0; j <= k; ++j) H[(size_t)j * m + k] = hc[j] + (pass2[k] ? hc[(size_t)(m + 1) + j] : 0.0);
            const double hn = hc[2 * (m + 1)];
            /* a one-pass step that kept too little of ||w||: from now on, steps take two passes */
            if (!pass2[k] && hn * hn < kKeepMin * kKeepMin * hc[k + 1]) two_pass = true;
            H[(size_t)(k + 1) * m + k] = hn;
            for (int j = 0; j < k; ++j) {
                double a = H[(size_t)j * m + k], b = H[(size_t)(j + 1) * m + k];
                H[(size_t)j * m + k] = cs[j] * a + sn[j] * b;
                H[(size_t)(j + 1) * m + k] = -sn[j] * a + cs[j] * b;
            }
            double a = H[(size_t)k * m + k], b = H[(size_t)(k + 1) * m + k];
            double rr = std::hypot(a, b);
            cs[k] = rr > 0 ? a / rr : 1.0;
            sn[k] = rr > 0 ? b / rr : 0.0;
            H[(size_t)k * m + k] = rr;
            H[(size_t)(k + 1) * m + k] = 0.0;
            g[k + 1] = -sn[k] * g[k];
            g[k] = cs[k] * g[k];
            res = std::fabs(g[k + 1]) / bnorm;
            ++k;
            ++it;
            if (res < tol || hn == 0.0) break;
        }
        /* y from the triangular system, u += D^-1 V y */
        for (int i = k - 1; i >= 0; --i) {
            double sum = g[i];
            for (int j = i + 1; j < k; ++j) sum -= H[(size_t)i * m + j] * yv[j];
            yv[i] = sum / H[(size_t)i * m + i];
        }
        if (k > 0) {
            GM_CHECK(hipMemcpyAsync(h1, yv.data(), (size_t)k * 8, hipMemcpyHostToDevice, s));
            if (ck > 0) {
                hipLaunchKernelGGL(k_combine, g1, b1, 0, s, n, (const double*)Vb, N, (const double*)h1, k, r);
                hipLaunchKernelGGL(k_coarse_restrict, dim3(ck), b1, 0, s, (const int32_t*)sys->c_ptr.ptr,
                                   (const int32_t*)sys->c_rows.ptr, (const double*)r, ccs);
                hipLaunchKernelGGL(k_coarse_solve, dim3((ck + 255) / 256), b1, 0, s, ck,
                                   (const double*)sys->c_einv.ptr, (const double*)ccs, ccc);
                hipLaunchKernelGGL(k_correct_coarse, g1, b1, 0, s, n, (const double*)r, (const double*)dinv, cagg,
                                   (const double*)ccc, du);
            } else {
                hipLaunchKernelGGL(k_correct, g1, b1, 0, s, n, (const double*)Vb, N, (const double*)h1, k,
                                   (const double*)dinv, du);
            }
        }
        restarts++;
        if (res < tol) {
            double rn = 0.0;
            if ((rc = true_residual(&rn))) return rc;
            res = rn / bnorm;
            conv = res < tol;
            if (conv) break;
        }
    }
    /* max_iter ran out: report the true residual of the last iterate, not the Givens estimate */
    if (!conv) {
        double rn = 0.0;
        if ((rc = true_residual(&rn))) return rc;
        res = rn / bnorm;
        conv = res < tol;
    }
    GM_CHECK(hipEventRecord(st.e1, s));
    GM_CHECK(hipMemcpyAsync(u, du, N * 8, hipMemcpyDeviceToHost, s));
    GM_CHECK(hipStreamSynchronize(s));
    float ms = 0.f;
    GM_CHECK(hipEventElapsedTime(&ms, st.e0, st.e1));
    if (info) {
        info->iterations = it;
        info->restarts = restarts;
        info->residual = res;
        info->converged = conv ? 1 : 0;
        info->reserved = 0;
        info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        info->device_ms = ms;
    }
    st.release();
    return SPAC_OK;
}

extern "C" spac_status spac_gmres_solve_gpu(int32_t n, const int32_t* Ip, const int32_t* J, const double* V,
                                            const double* B, double* u, double tol, int32_t max_iter,
                                            int32_t restart, spac_gmres_info* info) {
    return gmres_impl(n, Ip, J, V, B, nullptr, u, tol, max_iter, restart, info);
}

extern "C" spac_status spac_gmres_solve_system(spac_system* hs, double* u, double tol, int32_t max_iter,
                                               int32_t restart, spac_gmres_info* info) {
    if (!hs) return set_error(SPAC_E_INVALID_ARG, "null system");
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != reinterpret_cast<const DeviceSystem*>(hs)->device)
        return set_error(SPAC_E_INVALID_ARG, "the system lives on device %d; make it current first",
                         reinterpret_cast<const DeviceSystem*>(hs)->device);
    return gmres_impl(0, nullptr, nullptr, nullptr, nullptr, reinterpret_cast<const DeviceSystem*>(hs), u, tol,
                      max_iter, restart, info);
}
