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
#include <cstring>
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

/* the coarse-space block lists are over the 256-row blocks of the vector kernels (g1) */
static_assert(kCoarseRowBlock == 256, "coarse row blocks are the vector kernels' blocks");

constexpr int kDotBlock = 256;
constexpr int kDotPerThread = 8; /* rows per thread in the dot kernels */
constexpr int kDotChunk = kDotBlock * kDotPerThread;

/* inverse diagonal: 1 / the row's last nonzero diagonal entry, 1 where there is none; one
 * wavefront per row (four per block), lanes strided over the row */
__global__ void __launch_bounds__(256) k_diag_inv(int n, const int32_t* __restrict__ Ip,
                                                  const int32_t* __restrict__ J, const double* __restrict__ V,
                                                  double* __restrict__ dinv) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    int last = -1;
    for (int k = Ip[i] + lane; k < Ip[i + 1]; k += 64)
        if (J[k] == i && V[k] != 0.0) last = k;
    for (int off = 32; off >= 1; off >>= 1) last = max(last, __shfl_xor(last, off, 64));
    if (lane == 0) dinv[i] = last >= 0 ? 1.0 / V[last] : 1.0;
}

/* the same from each row's diagonal slot (-1: none), for a pattern with at most one diagonal
 * entry per row (a device system's) */
__global__ void k_diag_inv_slot(int n, const int32_t* __restrict__ dslot, const double* __restrict__ V,
                                double* __restrict__ dinv) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int k = dslot[i];
    dinv[i] = (k >= 0 && V[k] != 0.0) ? 1.0 / V[k] : 1.0;
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

/* w[i] -= sum_j h[j] Vb[j][i] (j = 0..nv-1), h staged in LDS; with npart, also the block's sum
 * of the updated w[i]^2 (256 rows per block, fixed tree) in npart[block]: k_scale forms ||w||
 * from these partials, so the norm needs no pass of its own */
__global__ void __launch_bounds__(256) k_update(int n, const double* __restrict__ Vb, size_t ld,
                                                const double* __restrict__ hdev, int nv, double* __restrict__ w,
                                                double* __restrict__ npart) {
    __shared__ double hs[1024];
    __shared__ double sh[4];
    for (int j = threadIdx.x; j < nv; j += blockDim.x) hs[j] = hdev[j];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double sq = 0.0;
    if (i < n) {
        double s = 0.0;
        for (int j = 0; j < nv; ++j) s = fma(hs[j], Vb[(size_t)j * ld + i], s);
        const double wi = w[i] - s;
        w[i] = wi;
        sq = wi * wi;
    }
    if (npart) {
        const double t = block_sum(sq, sh);
        if (threadIdx.x == 0) npart[blockIdx.x] = t;
    }
}

/* The sums over the rows of this 256-row block of x[i] per aggregate of the block (slot rslot[i]
 * of its list bagg[bptr[b] ..]), into cpart[bptr[b] + slot]: wavefront sums (fixed tree), then
 * the four wavefronts in order.  The coarse restriction Z^T x is then, per aggregate, the sum
 * of its blocks' entries (coarse_sums). */
__device__ __forceinline__ void coarse_partials(double x, int i, int n, const int32_t* __restrict__ bptr,
                                                const uint8_t* __restrict__ rslot, double* __restrict__ cpart) {
    __shared__ double sh[kCoarseBlockAggs][4];
    const int b = blockIdx.x, a0 = bptr[b], na = bptr[b + 1] - a0;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int slot = i < n ? (int)rslot[i] : -1;
    for (int q = 0; q < na; ++q) {
        double v = slot == q ? x : 0.0;
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_down(v, off, 64);
        if (lane == 0) sh[q][w] = v;
    }
    __syncthreads();
    if ((int)threadIdx.x < na) {
        const int q = threadIdx.x;
        cpart[a0 + q] = ((sh[q][0] + sh[q][1]) + sh[q][2]) + sh[q][3];
    }
}

/* out = x / den (the next Krylov vector) and z = dinv .* out (the SpMV input of the next Arnoldi
 * step, right preconditioning).  den = *den_ptr, or (den_ptr null) the square root of the sum of
 * k_update's nparts norm partials in block order, which block 0 also writes to *mirror.  With a
 * coarse space (cpart), also the block's per-aggregate sums of out. */
__global__ void __launch_bounds__(256) k_scale(int n, const double* __restrict__ x, const double* __restrict__ den_ptr,
                                               const double* __restrict__ npart, int nparts,
                                               double* __restrict__ mirror, const double* __restrict__ dinv,
                                               double* __restrict__ out, double* __restrict__ z,
                                               const int32_t* __restrict__ bptr, const uint8_t* __restrict__ rslot,
                                               double* __restrict__ cpart) {
    __shared__ double sh[4];
    __shared__ double den_s;
    if (!den_ptr) {
        double acc = 0.0;
        for (int q = threadIdx.x; q < nparts; q += blockDim.x) acc += npart[q];
        const double t = block_sum(acc, sh);
        if (threadIdx.x == 0) {
            den_s = sqrt(t);
            if (blockIdx.x == 0 && mirror) *mirror = den_s;
        }
        __syncthreads();
    }
    const double den = den_ptr ? den_ptr[0] : den_s;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double v = 0.0;
    if (i < n) {
        v = x[i] / den;
        out[i] = v;
        z[i] = dinv[i] * v;
    }
    if (cpart) coarse_partials(v, i, n, bptr, rslot, cpart);
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

/* k_coarse_galerkin for the kCoarseRowBlock rows of a block: one wavefront per row (lanes strided
 * over its entries) adds each entry into the block's LDS copy of the rows of E its rows' aggregates
 * own (slot rslot[i] of the block's aggregate list), then the block adds its nonzero sums into
 * efix: one global atomic per (aggregate, column aggregate) pair of the block instead of one per
 * run of every row.  Integer sums: the result equals k_coarse_galerkin's. */
__global__ void __launch_bounds__(256) k_coarse_galerkin_blk(int n, const int32_t* __restrict__ Ip,
                                                             const int32_t* __restrict__ J,
                                                             const double* __restrict__ V,
                                                             const int32_t* __restrict__ agg,
                                                             const int32_t* __restrict__ bptr,
                                                             const int32_t* __restrict__ bagg,
                                                             const uint8_t* __restrict__ rslot, int k,
                                                             unsigned long long* efix) {
    extern __shared__ unsigned long long eacc[]; /* [na][k] */
    const int b = blockIdx.x, a0 = bptr[b], na = bptr[b + 1] - a0;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int q = threadIdx.x; q < na * k; q += blockDim.x) eacc[q] = 0ull;
    __syncthreads();
    const int i1 = min(n, (b + 1) * kCoarseRowBlock);
    for (int i = b * kCoarseRowBlock + w; i < i1; i += 4) {
        unsigned long long* row = eacc + (size_t)rslot[i] * k;
        for (int q = Ip[i] + lane; q < Ip[i + 1]; q += 64)
            atomicAdd(row + agg[J[q]], (unsigned long long)__double2ll_rn(V[q] * kFix));
    }
    __syncthreads();
    for (int q = threadIdx.x; q < na * k; q += blockDim.x) {
        const unsigned long long v = eacc[q];
        if (v != 0ull) atomicAdd(efix + (size_t)bagg[a0 + q / k] * k + q % k, v);
    }
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

/* (E^-1 cs)[a] for the row e = E^-1[a] (row-major), by one wavefront: lane l sums entries
 * l, l + 64, ... in order, then a fixed butterfly; every lane returns the same value.  Both the
 * per-step prolongation and the final correction form coarse values with it, so the two apply
 * the same M^-1. */
__device__ __forceinline__ double coarse_dot(const double* __restrict__ e, const double* __restrict__ cs, int k,
                                             int lane) {
    double t = 0.0;
    for (int b = lane; b < k; b += 64) t = fma(e[b], cs[b], t);
    for (int off = 32; off >= 1; off >>= 1) t += __shfl_xor(t, off, 64);
    return t;
}

/* cc = E^-1 cs: one wavefront per aggregate (four per block) */
__global__ void __launch_bounds__(256) k_coarse_solve(int k, const double* __restrict__ einv,
                                                      const double* __restrict__ cs, double* __restrict__ cc) {
    const int a = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (a >= k) return;
    const double t = coarse_dot(einv + (size_t)a * k, cs, k, lane);
    if (lane == 0) cc[a] = t;
}

/* z[i] += (E^-1 cs)[agg(i)] row by row, one wavefront per row: the prolongation of coarse spaces
 * whose 256-row blocks span more than kCoarseBlockAggs aggregates. */
__global__ void __launch_bounds__(256) k_coarse_prolong_rows(int n, const int32_t* __restrict__ agg, int k,
                                                             const double* __restrict__ einv,
                                                             const double* __restrict__ cs, double* __restrict__ z) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    const double t = coarse_dot(einv + (size_t)agg[i] * k, cs, k, lane);
    if (lane == 0) z[i] += t;
}

/* E^-1 by Gauss-Jordan elimination with partial pivoting on W = [E | I] (k x 2k, row-major,
 * global memory), one workgroup: E from the fixed-point Galerkin sums, then per column c the
 * first row of largest |W[r][c]| (r >= c) as pivot, the row swap, the pivot row scaled by
 * d = 1 / pivot, and every other row r minus (its W[r][c]) times the scaled row (rows with
 * W[r][c] = 0 skipped): the operations of the textbook host routine, entry by entry.  Column c
 * and the scaled pivot row are staged in LDS; the elimination runs over all (row, column)
 * entries, kInvBatch loads in flight per thread.  einv = the right half; *singular = 1 (einv
 * unset) when a pivot is 0 or not finite. */
constexpr int kInvThreads = 1024, kInvBatch = 8;
__global__ void __launch_bounds__(kInvThreads) k_coarse_invert(int k, const unsigned long long* __restrict__ efix,
                                                               double* __restrict__ W, double* __restrict__ einv,
                                                               int* __restrict__ singular) {
    __shared__ double col[kMaxCoarse];
    __shared__ double prow[2 * kMaxCoarse];
    __shared__ double rv[kInvThreads / 64];
    __shared__ int ri[kInvThreads / 64];
    __shared__ int piv_s, bad_s;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int w2 = 2 * k, total = k * w2;
    for (int q = t; q < total; q += kInvThreads) {
        const int r = q / w2, j = q - r * w2;
        W[q] = j < k ? (double)(long long)efix[(size_t)r * k + j] / kFix : (j - k == r ? 1.0 : 0.0);
    }
    if (t == 0) bad_s = 0;
    __syncthreads();
    for (int c = 0; c < k; ++c) {
        for (int r = t; r < k; r += kInvThreads) col[r] = W[(size_t)r * w2 + c];
        __syncthreads();
        /* pivot: largest |col[r]|, r >= c, first index on ties */
        double best = -1.0;
        int bi = k;
        for (int r = c + t; r < k; r += kInvThreads) {
            const double v = fabs(col[r]);
            if (v > best) { best = v; bi = r; }
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const double ov = __shfl_xor(best, off, 64);
            const int oi = __shfl_xor(bi, off, 64);
            if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
        }
        if (lane == 0) { rv[wv] = best; ri[wv] = bi; }
        __syncthreads();
        if (t == 0) {
            double bv = rv[0];
            int b = ri[0];
            for (int q = 1; q < kInvThreads / 64; ++q)
                if (rv[q] > bv || (rv[q] == bv && ri[q] < b)) { bv = rv[q]; b = ri[q]; }
            if (b >= k || !(fabs(col[b]) > 0.0) || !(fabs(col[b]) < INFINITY)) bad_s = 1;
            piv_s = b;
        }
        __syncthreads();
        if (bad_s) break;
        const int piv = piv_s;
        const double d = 1.0 / col[piv];
        /* swap rows c and piv; the new row c scaled, to LDS */
        for (int j = t; j < w2; j += kInvThreads) {
            const double vc = W[(size_t)c * w2 + j], vp = W[(size_t)piv * w2 + j];
            const double nc = (piv != c ? vp : vc) * d;
            if (piv != c) W[(size_t)piv * w2 + j] = vc;
            W[(size_t)c * w2 + j] = nc;
            prow[j] = nc;
        }
        __syncthreads();
        /* every other row minus its multiplier (W[r][c] before the step, rows c and piv
           swapped) times the scaled pivot row */
        for (int q0 = t; q0 < total; q0 += kInvThreads * kInvBatch) {
            double v[kInvBatch], f[kInvBatch];
            int jj[kInvBatch];
            bool on[kInvBatch];
#pragma unroll
            for (int u = 0; u < kInvBatch; ++u) {
                const int q = q0 + u * kInvThreads;
                const int r = q / w2;
                jj[u] = q - r * w2;
                f[u] = (q < total && r != c) ? (r == piv ? col[c] : col[r]) : 0.0;
                on[u] = f[u] != 0.0;
                v[u] = on[u] ? W[q] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kInvBatch; ++u)
                if (on[u]) W[q0 + u * kInvThreads] = v[u] - f[u] * prow[jj[u]];
        }
        __syncthreads();
    }
    if (bad_s) {
        if (t == 0) *singular = 1;
        return;
    }
    if (t == 0) *singular = 0;
    for (int q = t; q < k * k; q += kInvThreads) einv[q] = W[(size_t)(q / k) * w2 + k + q % k];
}

/* E^-1 by in-place Gauss-Jordan elimination with partial pivoting, E held in registers: row r
 * belongs to the tpr = kInvThreads / k threads t = r tpr + sub, thread sub holding the nv =
 * ceil(k / tpr) (at most VMAX) columns sub nv .. sub nv + nv - 1.  Per column c: the first row of
 * largest |E[r][c]| (r >= c) becomes row c (the swap recorded); row c gets E[c][c] = 1 and is
 * scaled by d = 1 / pivot; every other row r gets E[r][c] = 0 and subtracts f = E[r][c] (before
 * the step) times row c, rows with f = 0 untouched.  Afterwards the recorded swaps, undone in
 * reverse order on the columns, place the inverse.  A step takes three barriers: wavefront 0
 * finds the pivot in column c (staged in LDS by the previous step); the pivot row's owners write
 * it scaled, and row c's owners write row c, to LDS; every row updates its registers and the
 * owners of column c + 1 stage it.  *singular = 1 (einv unset) when a pivot is 0 or not
 * finite. */
template <int VMAX>
__global__ void __launch_bounds__(kInvThreads) k_coarse_invert_reg(int k, const unsigned long long* __restrict__ efix,
                                                                   double* __restrict__ einv,
                                                                   int* __restrict__ singular) {
    __shared__ double col[2][kMaxCoarse];
    __shared__ double rowc[kMaxCoarse], prow[kMaxCoarse];
    __shared__ int perm[kMaxCoarse], pos[kMaxCoarse];
    __shared__ int piv_s;
    const int t = threadIdx.x, lane = t & 63;
    bool bad = false;
    const int tpr = kInvThreads / k, nv = (k + tpr - 1) / tpr;
    const int r = t / tpr, sub = t - r * tpr;
    const bool owner = r < k;
    const int j0 = sub * nv;                              /* first of the thread's columns */
    const int nown = owner ? max(0, min(nv, k - j0)) : 0; /* its columns inside the matrix */
    double w[VMAX];
#pragma unroll
    for (int v = 0; v < VMAX; ++v)
        w[v] = v < nown ? (double)(long long)efix[(size_t)r * k + j0 + v] / kFix : 0.0;
    if (owner && sub == 0) col[0][r] = w[0];
    __syncthreads();
    for (int c = 0; c < k; ++c) {
        const double* cc = col[c & 1];
        /* pivot: largest |E[r][c]|, r >= c, first index on ties (wavefront 0) */
        if (t < 64) {
            double best = -1.0;
            int bi = k;
            for (int q = c + lane; q < k; q += 64) {
                const double v = fabs(cc[q]);
                if (v > best) { best = v; bi = q; }
            }
            for (int off = 32; off >= 1; off >>= 1) {
                const double ov = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
            }
            if (lane == 0) {
                piv_s = (bi >= k || !(fabs(cc[bi]) > 0.0) || !(fabs(cc[bi]) < INFINITY)) ? -1 : bi;
                perm[c] = bi;
            }
        }
        __syncthreads();
        const int piv = piv_s;
        if (piv < 0) {
            bad = true;
            break;
        }
        const double d = 1.0 / cc[piv];
        /* the new row c (row piv, its column c set to 1, scaled) and the old row c (the new row
           piv) to LDS */
        if (owner && r == piv) {
#pragma unroll
            for (int v = 0; v < VMAX; ++v)
                if (v < nown) prow[j0 + v] = (j0 + v == c ? 1.0 : w[v]) * d;
        }
        if (owner && r == c && piv != c) {
#pragma unroll
            for (int v = 0; v < VMAX; ++v)
                if (v < nown) rowc[j0 + v] = w[v];
        }
        __syncthreads();
        if (owner) {
            if (r == c) {
#pragma unroll
                for (int v = 0; v < VMAX; ++v)
                    if (v < nown) w[v] = prow[j0 + v];
            } else {
                const bool swapped = r == piv;
                const double f = swapped ? cc[c] : cc[r];
                if (swapped) {
#pragma unroll
                    for (int v = 0; v < VMAX; ++v)
                        if (v < nown) w[v] = rowc[j0 + v];
                }
                if (f != 0.0) {
#pragma unroll
                    for (int v = 0; v < VMAX; ++v)
                        if (v < nown) w[v] = (j0 + v == c ? 0.0 : w[v]) - f * prow[j0 + v];
                }
            }
            /* column c + 1 for the next step */
            const int cn = c + 1;
            if (cn < k && sub == cn / nv) {
                const int vn = cn - sub * nv;
                double x = 0.0;
#pragma unroll
                for (int v = 0; v < VMAX; ++v)
                    if (v == vn) x = w[v];
                col[cn & 1][r] = x;
            }
        }
        __syncthreads();
    }
    if (bad) { /* every thread read the same piv_s, so all of them left the loop here */
        if (t == 0) *singular = 1;
        return;
    }
    /* the row swaps undone on the columns in reverse order: after swapping columns c and perm[c]
       for c = k - 1 .. 0, position j holds reduced column lab[j]; so reduced column m is column
       pos[m] of the inverse, lab[pos[m]] = m (lab is built in col[0]'s storage) */
    if (t == 0) {
        *singular = 0;
        int* lab = reinterpret_cast<int*>(col[0]);
        for (int m = 0; m < k; ++m) lab[m] = m;
        for (int c = k - 1; c >= 0; --c) {
            const int a = perm[c];
            if (a != c) {
                const int x = lab[c];
                lab[c] = lab[a];
                lab[a] = x;
            }
        }
        for (int m = 0; m < k; ++m) pos[lab[m]] = m;
    }
    __syncthreads();
    if (owner) {
#pragma unroll
        for (int v = 0; v < VMAX; ++v)
            if (v < nown) einv[(size_t)r * k + pos[j0 + v]] = w[v];
    }
}
/* columns per thread of the register-resident inversion (two instantiations): k fits one when
 * ceil(k / (kInvThreads / k)) is at most its count */
constexpr int kInvRegColsSmall = 24, kInvRegCols = 48;
static int inversion_columns(int k) {
    const int tpr = kInvThreads / k;
    return tpr >= 1 ? (k + tpr - 1) / tpr : kMaxCoarse;
}

/* t[i] = sum_j y[j] Vb[j][i] (the Krylov combination, before preconditioning); with cpart, also
 * the block's per-aggregate sums of t (coarse_partials) */
__global__ void __launch_bounds__(256) k_combine(int n, const double* __restrict__ Vb, size_t ld,
                                                 const double* __restrict__ y, int k, double* __restrict__ t,
                                                 const int32_t* __restrict__ bptr, const uint8_t* __restrict__ rslot,
                                                 double* __restrict__ cpart) {
    __shared__ double ys[1024];
    for (int j = threadIdx.x; j < k; j += blockDim.x) ys[j] = y[j];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    double s = 0.0;
    if (i < n) {
        for (int j = 0; j < k; ++j) s = fma(ys[j], Vb[(size_t)j * ld + i], s);
        t[i] = s;
    }
    if (cpart) coarse_partials(s, i, n, bptr, rslot, cpart);
}

/* cs[a] = sum of aggregate a's block partials cpart[aent[aptr[a] ..]] in block order (Z^T x),
 * into LDS, by the whole block */
__device__ __forceinline__ void coarse_sums(int k, const int32_t* __restrict__ aptr, const int32_t* __restrict__ aent,
                                            const double* __restrict__ cpart, double* cs) {
    for (int a = threadIdx.x; a < k; a += blockDim.x) {
        double t = 0.0;
        for (int e = aptr[a]; e < aptr[a + 1]; ++e) t += cpart[aent[e]];
        cs[a] = t;
    }
}

/* z[i] += (E^-1 Z^T x)[agg(i)] for the 256 rows of a block, from the block partials of x: the
 * block forms Z^T x (coarse_sums), its wavefronts the coarse values of the aggregates its rows
 * span, then every row adds its own */
__global__ void __launch_bounds__(kCoarseRowBlock) k_coarse_prolong_part(
    int n, int k, const int32_t* __restrict__ aptr, const int32_t* __restrict__ aent,
    const double* __restrict__ cpart, const int32_t* __restrict__ bptr, const int32_t* __restrict__ bagg,
    const uint8_t* __restrict__ rslot, const double* __restrict__ einv, double* __restrict__ z) {
    __shared__ double cs[kMaxCoarse];
    __shared__ double cv[kCoarseBlockAggs];
    coarse_sums(k, aptr, aent, cpart, cs);
    __syncthreads();
    const int b = blockIdx.x, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int a0 = bptr[b], na = bptr[b + 1] - a0;
    for (int q = w; q < na; q += kCoarseRowBlock / 64) {
        const double t = coarse_dot(einv + (size_t)bagg[a0 + q] * k, cs, k, lane);
        if (lane == 0) cv[q] = t;
    }
    __syncthreads();
    const int i = b * kCoarseRowBlock + threadIdx.x;
    if (i < n) z[i] += cv[rslot[i]];
}

/* cc = E^-1 Z^T x from the block partials of x (one block: the sums, then one wavefront per
 * aggregate), the coarse values k_coarse_prolong_part forms */
__global__ void __launch_bounds__(1024) k_coarse_solve_part(int k, const int32_t* __restrict__ aptr,
                                                            const int32_t* __restrict__ aent,
                                                            const double* __restrict__ cpart,
                                                            const double* __restrict__ einv, double* __restrict__ cc) {
    __shared__ double cs[kMaxCoarse];
    coarse_sums(k, aptr, aent, cpart, cs);
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int a = w; a < k; a += 16) {
        const double t = coarse_dot(einv + (size_t)a * k, cs, k, lane);
        if (lane == 0) cc[a] = t;
    }
}

/* u[i] += dinv[i] t[i] + cc[agg(i)]: u += M^-1 t with the two-level M^-1 */
__global__ void k_correct_coarse(int n, const double* __restrict__ t, const double* __restrict__ dinv,
                                 const int32_t* __restrict__ agg, const double* __restrict__ cc,
                                 double* __restrict__ u) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] += dinv[i] * t[i] + cc[agg[i]];
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
    DeviceBuffer Ip, J, V, B, u, dinv, Vb, r, part, h, nrm, z, npart;
    double* col = nullptr;     /* pinned, host-mapped: one Hessenberg column block per Arnoldi step */
    double* col_dev = nullptr; /* its device address */
    size_t col_bytes = 0;
    double* hu = nullptr;      /* pinned staging of the iterate (a device system's solves) */
    size_t hu_bytes = 0;
    hipStream_t stream = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipEvent_t step_done[kEvents] = {};
    void release() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (DeviceBuffer* b : {&Ip, &J, &V, &B, &u, &dinv, &Vb, &r, &part, &h, &nrm, &z, &npart}) b->release();
        if (col) (void)hipHostFree(col);
        col = col_dev = nullptr;
        col_bytes = 0;
        if (hu) (void)hipHostFree(hu);
        hu = nullptr;
        hu_bytes = 0;
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

/* the work state a device system keeps between solves (DeviceSystem::gm) */
static void free_gmres_state(void* p) { delete static_cast<GmresState*>(p); }

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
    /* a device system keeps its work state (buffers, stream, events, pinned columns) between
       solves; a host CSR solve has its own, released on return */
    GmresState local;
    if (sys && !sys->gm) {
        sys->gm = new GmresState();
        sys->gm_free = free_gmres_state;
    }
    GmresState& st = sys ? *static_cast<GmresState*>(sys->gm) : local;
    spac_status rc;
    if ((!sys && ((rc = st.Ip.ensure((N + 1) * 4)) || (rc = st.J.ensure((size_t)nnz * 4 + 4)) ||
                  (rc = st.V.ensure((size_t)nnz * 8 + 8)) || (rc = st.B.ensure(N * 8)))) ||
        (rc = st.u.ensure(N * 8)) || (rc = st.dinv.ensure(N * 8)) || (rc = st.Vb.ensure((size_t)(m + 1) * N * 8)) ||
        (rc = st.r.ensure(N * 8)) || (rc = st.part.ensure((size_t)(m + 1) * nchunks * 8)) ||
        (rc = st.h.ensure(hlen * 8)) || (rc = st.nrm.ensure(64)) || (rc = st.z.ensure(N * 8)) ||
        (rc = st.npart.ensure((N + 255) / 256 * 8 + 8))) {
        st.release();
        return rc;
    }
    if (!st.stream) GM_CHECK(hipStreamCreateWithFlags(&st.stream, hipStreamNonBlocking));
    if (!st.e0) GM_CHECK(hipEventCreate(&st.e0));
    if (!st.e1) GM_CHECK(hipEventCreate(&st.e1));
    for (hipEvent_t& e : st.step_done)
        if (!e) GM_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    if (st.col_bytes < (size_t)m * hlen * 8) {
        double* old = st.col;
        st.col = st.col_dev = nullptr;
        st.col_bytes = 0;
        if (old) GM_CHECK(hipHostFree(old));
        GM_CHECK(hipHostMalloc((void**)&st.col, (size_t)m * hlen * 8, hipHostMallocMapped));
        GM_CHECK(hipHostGetDevicePointer((void**)&st.col_dev, st.col, 0));
        st.col_bytes = (size_t)m * hlen * 8;
    }
    hipStream_t s = st.stream;
    if (!sys) {
        GM_CHECK(hipMemcpyAsync(st.Ip.ptr, Ip, (N + 1) * 4, hipMemcpyHostToDevice, s));
        GM_CHECK(hipMemcpyAsync(st.J.ptr, J, (size_t)nnz * 4, hipMemcpyHostToDevice, s));
        GM_CHECK(hipMemcpyAsync(st.V.ptr, V, (size_t)nnz * 8, hipMemcpyHostToDevice, s));
        GM_CHECK(hipMemcpyAsync(st.B.ptr, B, N * 8, hipMemcpyHostToDevice, s));
    }
    /* the initial guess goes up through pinned staging when the state is kept (system solves) */
    if (sys && st.hu_bytes < N * 8) {
        double* old = st.hu;
        st.hu = nullptr;
        st.hu_bytes = 0;
        if (old) GM_CHECK(hipHostFree(old));
        GM_CHECK(hipHostMalloc((void**)&st.hu, N * 8, hipHostMallocDefault));
        st.hu_bytes = N * 8;
    }
    if (sys) {
        GM_CHECK(hipStreamSynchronize(s)); /* the staging block may still feed the last solve's copy */
        std::memcpy(st.hu, u, N * 8);
        GM_CHECK(hipMemcpyAsync(st.u.ptr, st.hu, N * 8, hipMemcpyHostToDevice, s));
    } else {
        GM_CHECK(hipMemcpyAsync(st.u.ptr, u, N * 8, hipMemcpyHostToDevice, s));
    }
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
    double* dn = (double*)st.nrm.ptr;
    double* npart = (double*)st.npart.ptr;
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

    if (sys && sys->dslot.ptr)
        hipLaunchKernelGGL(k_diag_inv_slot, g1, b1, 0, s, n, (const int32_t*)sys->dslot.ptr, dV, dinv);
    else
        hipLaunchKernelGGL(k_diag_inv, dim3((n + 3) / 4), b1, 0, s, n, dIp, dJ, dV, dinv);
    /* two-level preconditioner of a device system: Galerkin matrix and its inverse on the device */
    const int ck = sys ? sys->coarse_k : 0;
    const int32_t* cagg = ck ? (const int32_t*)sys->c_agg.ptr : nullptr;
    double* ccs = ck ? (double*)sys->c_vec.ptr : nullptr;
    double* ccc = ccs + ck;
    int* csing = ck ? (int*)((char*)sys->c_work.ptr + (size_t)2 * ck * ck * 8) : nullptr;
    if (ck > 0) {
        const size_t kk = (size_t)ck * ck;
        unsigned long long* efix = (unsigned long long*)sys->c_efix.ptr;
        hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)((kk + 255) / 256)), b1, 0, s, kk, efix);
        const size_t glds = (size_t)sys->c_blk_max * ck * 8;
        if (sys->c_blk_max <= kCoarseBlockAggs && glds <= 65536)
            hipLaunchKernelGGL(k_coarse_galerkin_blk, dim3((n + kCoarseRowBlock - 1) / kCoarseRowBlock), b1, glds, s,
                               n, dIp, dJ, dV, cagg, (const int32_t*)sys->c_bptr.ptr, (const int32_t*)sys->c_bagg.ptr,
                               (const uint8_t*)sys->c_rslot.ptr, ck, efix);
        else
            hipLaunchKernelGGL(k_coarse_galerkin, g1, b1, 0, s, n, dIp, dJ, dV, cagg, ck, efix);
        const int icols = inversion_columns(ck);
        if (icols <= kInvRegColsSmall)
            hipLaunchKernelGGL(k_coarse_invert_reg<kInvRegColsSmall>, dim3(1), dim3(kInvThreads), 0, s, ck,
                               (const unsigned long long*)efix, (double*)sys->c_einv.ptr, csing);
        else if (icols <= kInvRegCols)
            hipLaunchKernelGGL(k_coarse_invert_reg<kInvRegCols>, dim3(1), dim3(kInvThreads), 0, s, ck,
                               (const unsigned long long*)efix, (double*)sys->c_einv.ptr, csing);
        else
            hipLaunchKernelGGL(k_coarse_invert, dim3(1), dim3(kInvThreads), 0, s, ck, (const unsigned long long*)efix,
                               (double*)sys->c_work.ptr, (double*)sys->c_einv.ptr, csing);
        GM_CHECK(hipGetLastError());
    }
    /* z += Z E^-1 Z^T x (after z = D^-1 x).  When every 256-row block spans at most
       kCoarseBlockAggs aggregates, k_scale leaves Z^T x as per-block partials and one kernel
       prolongs from them (coarse_blocks); otherwise a restriction and a row-wise prolongation. */
    const bool coarse_blocks = ck > 0 && sys->c_blk_max <= kCoarseBlockAggs;
    const int32_t* cbptr = ck ? (const int32_t*)sys->c_bptr.ptr : nullptr;
    const uint8_t* crslot = ck ? (const uint8_t*)sys->c_rslot.ptr : nullptr;
    double* cpart = coarse_blocks ? (double*)sys->c_part.ptr : nullptr;
    auto coarse = [&](const double* x, double* z) {
        if (coarse_blocks) {
            hipLaunchKernelGGL(k_coarse_prolong_part, dim3((n + kCoarseRowBlock - 1) / kCoarseRowBlock),
                               dim3(kCoarseRowBlock), 0, s, n, ck, (const int32_t*)sys->c_aptr.ptr,
                               (const int32_t*)sys->c_aent.ptr, (const double*)cpart, cbptr,
                               (const int32_t*)sys->c_bagg.ptr, crslot, (const double*)sys->c_einv.ptr, z);
        } else {
            hipLaunchKernelGGL(k_coarse_restrict, dim3(ck), b1, 0, s, (const int32_t*)sys->c_ptr.ptr,
                               (const int32_t*)sys->c_rows.ptr, x, ccs);
            hipLaunchKernelGGL(k_coarse_prolong_rows, dim3((n + 3) / 4), b1, 0, s, n, cagg, ck,
                               (const double*)sys->c_einv.ptr, (const double*)ccs, z);
        }
    };
    norm(dB, dn);
    double bnorm = 0.0;
    int singular = 0;
    GM_CHECK(hipMemcpyAsync(&bnorm, dn, 8, hipMemcpyDeviceToHost, s));
    if (ck > 0) GM_CHECK(hipMemcpyAsync(&singular, csing, 4, hipMemcpyDeviceToHost, s));
    GM_CHECK(hipStreamSynchronize(s));
    if (singular) return set_error(SPAC_E_INVALID_ARG, "the coarse Galerkin matrix is singular");
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
        hipLaunchKernelGGL(k_update, g1, b1, 0, s, n, (const double*)Vb, N, (const double*)h1, nv, w,
                           two_pass ? (double*)nullptr : npart);
        if (two_pass) {
            dots(Vb, nv, w, h2, colk + (m + 1), 0);
            hipLaunchKernelGGL(k_update, g1, b1, 0, s, n, (const double*)Vb, N, (const double*)h2, nv, w, npart);
        }
        /* v_{k+1} = w / ||w|| (the norm from k_update's partials, mirrored for the host) */
        hipLaunchKernelGGL(k_scale, g1, b1, 0, s, n, (const double*)w, (const double*)nullptr, (const double*)npart,
                           (int)g1.x, colk + 2 * (m + 1), (const double*)dinv, w, dz, cbptr, crslot, cpart);
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
        hipLaunchKernelGGL(k_scale, g1, b1, 0, s, n, (const double*)r, (const double*)dn, (const double*)nullptr, 0,
                           (double*)nullptr, (const double*)dinv, Vb, dz, cbptr, crslot, cpart);
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
                hipLaunchKernelGGL(k_combine, g1, b1, 0, s, n, (const double*)Vb, N, (const double*)h1, k, r, cbptr,
                                   crslot, cpart);
                if (coarse_blocks) {
                    hipLaunchKernelGGL(k_coarse_solve_part, dim3(1), dim3(1024), 0, s, ck,
                                       (const int32_t*)sys->c_aptr.ptr, (const int32_t*)sys->c_aent.ptr,
                                       (const double*)cpart, (const double*)sys->c_einv.ptr, ccc);
                } else {
                    hipLaunchKernelGGL(k_coarse_restrict, dim3(ck), b1, 0, s, (const int32_t*)sys->c_ptr.ptr,
                                       (const int32_t*)sys->c_rows.ptr, (const double*)r, ccs);
                    hipLaunchKernelGGL(k_coarse_solve, dim3((ck + 3) / 4), b1, 0, s, ck,
                                       (const double*)sys->c_einv.ptr, (const double*)ccs, ccc);
                }
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
    GM_CHECK(hipMemcpyAsync(sys ? (void*)st.hu : (void*)u, du, N * 8, hipMemcpyDeviceToHost, s));
    GM_CHECK(hipStreamSynchronize(s));
    if (sys) std::memcpy(u, st.hu, N * 8);
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
    return SPAC_OK; /* a host CSR solve's state is released here (local), a system's is kept */
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
