/*
 * fk_kernels.hip — MI355X (gfx950) Feynman–Kac path kernels and the C ABI of
 * include/spac_fk.h.
 *
 * Path kernel (Simulacion_OMP, Solvers/FeynmanKacSolver.cpp:137-197):
 *   - one wavefront lane per path, all path state in VGPRs;
 *   - persistent grid: each wave takes work items (solve, path range) from a
 *     global queue and refills lanes individually as soon as their path
 *     stops, so heavy-tailed exit times do not idle the other 63 lanes;
 *   - counter-based Philox stream keyed by (seed, stream) and counted by
 *     (step, path): a path's bits do not depend on which lane, wave or launch
 *     shape ran it;
 *   - per-path exit data written to fixed SoA slots (solve offset + path).
 * Reduction kernel (Reduce_Analytic + Update, :198-334): one wave per solve,
 * lane l sums paths l, l+64, ... with TwoSum error terms, then the 64 lane
 * sums are folded pairwise — a fixed order, so estimators are bit-stable
 * across launches and equal to the CPU restatement's.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <cstddef>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/spac_fk.h"
#include "fk_device.h"
#include "fk_internal.h"

namespace spac {


#define SPAC_HIP_CHECK(expr)                                                                   \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return set_error(SPAC_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                             __FILE__, __LINE__);                                              \
    } while (0)

/* ------------------------------------------------------ kernel args -- */
enum { SUB_NONE = 0, SUB_CIRCLE = 2 };
enum { CV_ANALYTIC = SPAC_CV_ANALYTIC, CV_NONE = SPAC_CV_NONE, CV_TABLE = SPAC_CV_TABLE, CV_NETWORK = SPAC_CV_NETWORK };

struct PathArgs {
    double dom[4];
    double gobet;
    uint64_t seed;
    const uint64_t* seed_dev; /* non-null: the seed is read here (captured graphs) */
    int32_t sigma_fresh;      /* 1: re-evaluate sigma(X) each step (a stop rule that can keep a path
                                 alive outside the boundary, gobet_constant >= 0) */
    int32_t max_steps;
    int32_t n_items;
    int32_t reflecting;     /* SPAC_REFLECT_* sides (REFL kernels) */
    double rho;             /* Lepingle band */
    /* solves */
    const double* x0;
    const double* y0;
    const double* h;        /* per solve, or nullptr: h_uniform */
    double h_uniform;
    const double* sqrth;    /* per solve sqrt(h), or nullptr: sqrth_uniform */
    double sqrth_uniform;
    const double* sig0;     /* per solve sigma(X0) (coefficient sets with a variable sigma) */
    const float* sig0f;
    const uint32_t* stream; /* per solve, or nullptr: solve index */
    const double* circle;   /* 3 per solve (R, cx, cy), or nullptr: circle_uniform */
    double circle_uniform[3];
    const int64_t* rec_off; /* per solve, or nullptr: solve * paths_uniform */
    uint32_t paths_uniform;
    uint32_t chunk;
    /* work queue: items[i] = (solve, p0, p1, 0) or nullptr: uniform chunks */
    const int4* items;
    int32_t chunks_per_solve;
    int32_t* queue;
    /* records */
    double* xt;
    double* yt;
    double* Y;
    double* Z;
    double* xi;
    double* tau;
    double* gval;
    int32_t* steps;
    int8_t* outcome;
    /* tabulated control variate (CV_TABLE): FP64 and FP32 views of one table */
    spac_cv_table cv;
    CvTable32 cv32;
    /* the network itself (CV_NETWORK, FP32 kernel; cv and cv32 hold its table for the exit scores) */
    CvNet32 cvnet;
    /* tabulated coefficients (CoefTab kernels): FP64 table and its float copy */
    spac_coef_table tab;
    CoefTable32 tab32;
};

/* A coefficient-set object for a kernel: the compiled sets are stateless, a tabulated set
 * carries its table. */
template <class C>
struct CoefMaker {
    __device__ __forceinline__ static C make(const spac_coef_table&, const CoefTable32&) { return C(); }
};
template <class B>
struct CoefMaker<CoefTab<B>> {
    __device__ __forceinline__ static CoefTab<B> make(const spac_coef_table& t, const CoefTable32&) {
        CoefTab<B> c;
        c.t = t;
        return c;
    }
};
template <class B>
struct CoefMaker<CoefTab32<B>> {
    __device__ __forceinline__ static CoefTab32<B> make(const spac_coef_table&, const CoefTable32& t) {
        CoefTab32<B> c;
        c.t = t;
        return c;
    }
};

/* Lanes of a wave whose exits (and refills) are handled together.  The FP32 kernel's exit
 * branch projects the exit point in FP64 (tens of instructions under a divergent branch that
 * would otherwise run on almost every wave-step), so it waits for a batch; the FP64 kernel's
 * exit branch is only the record stores, and waiting would only idle lanes. */
#ifndef SPAC_FP32_LANE_BATCH
#define SPAC_FP32_LANE_BATCH 8
#endif
constexpr int kLaneBatch32 = SPAC_FP32_LANE_BATCH;
constexpr int kLaneBatch64 = 1;

/* A wave claims a run of work items with one atomic: several while many items remain (fewer
 * atomics on the one queue counter), one near the end so that the last items stay balanced
 * across waves.  Returns the first item (>= n_items: the queue is drained) and sets *end. */
constexpr int32_t kMaxClaim = 4;
__device__ __forceinline__ int32_t claim_items(const PathArgs& a, uint32_t lane, int32_t* end) {
    const int32_t waves = (int32_t)(gridDim.x * (blockDim.x >> 6));
    const int32_t left = a.n_items - *end; /* from the wave's last claim: a lower bound of what was left */
    const int32_t want = max(1, min(kMaxClaim, left / (4 * waves)));
    int32_t item = 0;
    if (lane == 0) item = atomicAdd(a.queue, want);
    item = __shfl(item, 0);
    *end = min(item + want, a.n_items);
    return item;
}

/* Paths that have taken many steps are the likely last ones of the launch (exit times have an
 * exponential tail): a wave holding one issues ahead of the others on its SIMD, so the longest
 * paths advance at close to their own latency while the bulk of the work fills the rest.  Used
 * by the FP32 kernel, whose launches at a few tens of paths per row end on such paths; the FP64
 * kernel measured no difference with it. */
constexpr int32_t kOldPathSteps = 384;
__device__ __forceinline__ void wave_priority_by_age(bool active, int32_t nstep) {
    if (__ballot(active && nstep >= kOldPathSteps) != 0ull) __builtin_amdgcn_s_setprio(3);
    else __builtin_amdgcn_s_setprio(0);
}

/* The Brownian normals of path-step nstep: on an even step the pair's Philox block, whose
 * odd-step words go to *keep0, *keep1; on an odd step those kept words (the same path's previous
 * step on this lane).  A path starts at step 0, so an odd step always follows its even one. */
__device__ __forceinline__ void step_normals(uint64_t seed, uint32_t stream, uint32_t path, int32_t nstep,
                                             uint32_t* keep0, uint32_t* keep1, double* n0, double* n1) {
    uint32_t w0, w1;
    if ((nstep & 1) == 0) {
        const spac_u32x4 r = spac_brownian_block(seed, stream, path, (uint32_t)nstep >> 1);
        w0 = r.v[0]; w1 = r.v[1]; *keep0 = r.v[2]; *keep1 = r.v[3];
    } else {
        w0 = *keep0; w1 = *keep1;
    }
    spac_normal_words(w0, w1, n0, n1);
}

/* Idle lanes take new paths only while the wave's active paths are at an even step (or none is
 * active), so that all active lanes of a wave share their step parity and the Philox block of a
 * step pair is drawn under a wave-uniform branch; an idle lane waits at most one step. */
__device__ __forceinline__ bool odd_step_phase(bool active, int32_t nstep) {
    return __ballot(active && (nstep & 1)) != 0ull;
}

/* A kernel argument read where it is used.  The kernarg segment pointer passes through an empty
 * asm, so the scalar load stays at the use and the value does not occupy scalar registers across
 * the path loop (its scalar registers are full: a value held there is spilled to VGPR lanes and
 * costs two VALU instructions per use).  PathArgs is the kernels' only argument, at offset 0. */
template <class T>
__device__ __forceinline__ T late_arg(size_t offset) {
    const __attribute__((address_space(4))) char* k =
        (const __attribute__((address_space(4))) char*)__builtin_amdgcn_kernarg_segment_ptr();
    asm volatile("" : "+s"(k));
    return *(const __attribute__((address_space(4))) T*)(k + offset);
}
#define SPAC_LATE_ARG(field) late_arg<decltype(PathArgs::field)>(offsetof(PathArgs, field))

__device__ __forceinline__ uint32_t lane_rank(unsigned long long mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__global__ void zero_counter_kernel(int32_t* q) {
    if (threadIdx.x == 0) q[0] = 0;
}

/* Tail-first work order (SPAC_ORDER_TAIL).  The exit time of a path started at X0 in a region of
 * size R has the tail P(tau > t) ~ exp(-lambda t), 1/lambda proportional to R^2 / sigma^2: the
 * last paths to finish a launch come from the solves with the largest R^2 / (sigma(X0)^2 h) (in
 * steps), and the launch ends sooner when those start first.  A counting sort of the items by
 * that key, bucketed in eighth octaves, in three kernels without global atomics: histograms of
 * contiguous item slices (one block each), bucket offsets (largest key first, then slice
 * order), and the scatter through LDS cursors.  Only the queue order changes: records sit at
 * fixed slots and the random stream is keyed by (solve, path), so results do not depend on it. */
struct TailKey {
    const double* sig0;    /* sigma(X0) per solve, or nullptr: 1 */
    const double* h;       /* per solve, or nullptr: h_uniform */
    double h_uniform;
    const double* circle;  /* (R, cx, cy) per solve, or nullptr: R_uniform */
    double R_uniform;
};

__device__ __forceinline__ int tail_bucket(const TailKey& k, int32_t solve) {
    const double s = k.sig0 ? k.sig0[solve] : 1.0;
    const double h = k.h ? k.h[solve] : k.h_uniform;
    const double R = k.circle ? k.circle[3 * (int64_t)solve] : k.R_uniform;
    const float key = (float)((R * R) / (s * s * h));
    const float b = (key > 1.0f) ? __builtin_amdgcn_logf(key) * 8.0f : 0.0f; /* log2 */
    return (int)__builtin_fminf(b, (float)(kTailBuckets - 1));
}

__device__ __forceinline__ void tail_slice(int32_t n, int32_t* i0, int32_t* i1) {
    const int32_t per = (n + kTailBlocks - 1) / kTailBlocks;
    *i0 = min(n, (int32_t)blockIdx.x * per);
    *i1 = min(n, *i0 + per);
}

__global__ void __launch_bounds__(kTailBuckets) k_tail_hist(int32_t n, const int4* __restrict__ items, TailKey k,
                                                             int32_t* __restrict__ hist) {
    __shared__ int32_t cnt[kTailBuckets];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    int32_t i0, i1;
    tail_slice(n, &i0, &i1);
    for (int32_t i = i0 + (int32_t)threadIdx.x; i < i1; i += kTailBuckets) atomicAdd(&cnt[tail_bucket(k, items[i].x)], 1);
    __syncthreads();
    hist[blockIdx.x * kTailBuckets + threadIdx.x] = cnt[threadIdx.x];
}

/* one block: hist[slice][bucket] becomes the first output slot of that slice's items of that
 * bucket; buckets from the largest key down, slices in order within a bucket */
__global__ void __launch_bounds__(kTailBuckets) k_tail_offsets(int32_t* __restrict__ hist) {
    __shared__ int32_t tot[kTailBuckets];
    const int b = threadIdx.x;
    int32_t t = 0;
    for (int s = 0; s < kTailBlocks; ++s) t += hist[s * kTailBuckets + b];
    tot[b] = t;
    __syncthreads();
    int32_t base = 0;
    for (int c = kTailBuckets - 1; c > b; --c) base += tot[c];
    for (int s = 0; s < kTailBlocks; ++s) {
        const int32_t v = hist[s * kTailBuckets + b];
        hist[s * kTailBuckets + b] = base;
        base += v;
    }
}

__global__ void __launch_bounds__(kTailBuckets) k_tail_scatter(int32_t n, const int4* __restrict__ items, TailKey k,
                                                                const int32_t* __restrict__ offs,
                                                                int4* __restrict__ out) {
    __shared__ int32_t cur[kTailBuckets];
    cur[threadIdx.x] = offs[blockIdx.x * kTailBuckets + threadIdx.x];
    __syncthreads();
    int32_t i0, i1;
    tail_slice(n, &i0, &i1);
    for (int32_t i = i0 + (int32_t)threadIdx.x; i < i1; i += kTailBuckets) {
        const int4 it = items[i];
        out[atomicAdd(&cur[tail_bucket(k, it.x)], 1)] = it;
    }
}

/* u at a subdomain exit for the statistics' score (Solvers/FeynmanKacSolver.cpp:254):
 * the coefficient set's u, or u_hat from the control-variate table. */
template <class Coef, int CV>
__device__ __forceinline__ double exit_u(const PathArgs& a, const Coef& coef, double x, double y) {
    if (CV == CV_TABLE) {
        int32_t k00;
        double ax, ay;
        spac_cv_cell(&a.cv, x, y, &k00, &ax, &ay);
        return spac_cv_eval(&a.cv, k00, ax, ay, 0);
    }
    return coef.u(x, y);
}

/* The frontier Dentro tests (Solvers/FeynmanKacSolver.cpp:72-80): the domain
 * rectangle or, when nearer, the solve's circle; *dom says which. */
template <int SUB>
__device__ __forceinline__ BoundaryHit nearest_boundary(const PathArgs& a, double R, double cx, double cy, double X,
                                                        double Yc, bool* dom) {
    BoundaryHit hit = rectangle_distance(a.dom, X, Yc);
    bool use_dom = true;
    if (SUB == SUB_CIRCLE) {
        BoundaryHit sub = circle_distance(R, cx, cy, X, Yc);
        use_dom = hit.d >= sub.d;
        if (!use_dom) hit = sub;
    } else {
        if (!(hit.d >= -__builtin_inf())) {
            use_dom = false;
            hit.d = -__builtin_inf(); hit.ex = X; hit.ey = Yc; hit.nx = 0.0; hit.ny = 0.0;
        }
    }
    *dom = use_dom;
    return hit;
}

/* REFL = true adds reflecting rectangle sides: the Lepingle step
 * (Paso_Reflejante, Solvers/FeynmanKacSolver.cpp:19-67, restated without its
 * defects) for paths whose nearest frontier is a reflecting side, and the
 * projection of points found beyond one (reflejada, :88-96). */
template <class Coef, int SUB, int CV, bool REFL>
__global__ void __launch_bounds__(256) fk_path_kernel_fp64(PathArgs a) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t seed = a.seed_dev ? *a.seed_dev : a.seed;
    const Coef coef = CoefMaker<Coef>::make(a.tab, a.tab32);
    bool active = false;
    /* lane path state */
    int32_t solve = 0;
    uint32_t path = 0;
    int32_t nstep = 0;
    uint32_t keep0 = 0u, keep1 = 0u; /* the odd step's Philox words of the current step pair */
    double X = 0.0, Yc = 0.0, Yw = 0.0, Z = 0.0, xi = 0.0, t = 0.0;
    double h = 0.0, sqrth = 0.0, R = 0.0, cx = 0.0, cy = 0.0;
    double sig = 0.0; /* sigma at the current X */
    /* reflecting sides: the last Dentro's distance and normal, whether the next
     * step reflects, and local time carried from a projection */
    double rd = 0.0, rnx = 0.0, rny = 0.0, chi_p = 0.0;
    bool rrefl = false;
    bool done = false;   /* stopped, its record not yet stored (batched exits) */
    int8_t done_out = 0;
    double done_ex = 0.0, done_ey = 0.0;
    uint32_t stream = 0;
    int64_t rec = 0;
    /* wave-uniform pool */
    int32_t pool_solve = 0;
    uint32_t pool_next = 0, pool_end = 0;
    int32_t pool_item = 0, pool_item_end = 0; /* the wave's claimed run of items */
    bool drained = false;

    while (true) {
        /* Exits and refills in batches of kLaneBatch64 lanes or when no lane steps any more (the
           structure of the FP32 kernel); with batches of one a record is stored at its stop and
           none of this is compiled. */
        if constexpr (kLaneBatch64 > 1) {
            const unsigned long long stopped = __ballot(done);
            if (stopped != 0ull && (__popcll(stopped) >= kLaneBatch64 || __ballot(active) == 0ull)) {
                if (done) {
                    int64_t g = rec + (int64_t)path;
                    a.xt[g] = done_ex;
                    a.yt[g] = done_ey;
                    a.Y[g] = Yw;
                    a.Z[g] = Z;
                    a.xi[g] = xi;
                    a.tau[g] = t;
                    a.steps[g] = nstep; /* gval: fk_exit_values_kernel, lane-dense after the launch */
                    a.outcome[g] = done_out;
                    done = false;
                }
            }
        }
        unsigned long long idle = __ballot(!active && !done);
        if constexpr (kLaneBatch64 > 1)
            if (__popcll(idle) < kLaneBatch64 && __ballot(active) != 0ull) idle = 0ull; /* refill later, together */
        if (odd_step_phase(active, nstep)) idle = 0ull;
        while (idle != 0ull && !drained) {
            if (pool_next >= pool_end) {
                if (pool_item >= pool_item_end) {
                    const int32_t item0 = claim_items(a, lane, &pool_item_end);
                    if (item0 >= a.n_items) { drained = true; break; }
                    pool_item = item0;
                }
                const int32_t item = pool_item++;
                if (a.items) {
                    int4 it = a.items[item];
                    pool_solve = it.x;
                    pool_next = (uint32_t)it.y;
                    pool_end = (uint32_t)it.z;
                } else {
                    pool_solve = item / a.chunks_per_solve;
                    uint32_t c = (uint32_t)(item % a.chunks_per_solve);
                    pool_next = c * a.chunk;
                    pool_end = min(pool_next + a.chunk, a.paths_uniform);
                }
            }
            uint32_t avail = pool_end - pool_next;
            uint32_t nidle = (uint32_t)__popcll(idle);
            uint32_t rank = lane_rank(idle);
            if (!active && !done && rank < avail) {
                solve = pool_solve;
                path = pool_next + rank;
                const double* ah = SPAC_LATE_ARG(h);
                const double* asq = SPAC_LATE_ARG(sqrth);
                const uint32_t* ast = SPAC_LATE_ARG(stream);
                const int64_t* aoff = SPAC_LATE_ARG(rec_off);
                h = ah ? ah[solve] : SPAC_LATE_ARG(h_uniform);
                sqrth = asq ? asq[solve] : SPAC_LATE_ARG(sqrth_uniform);
                stream = ast ? ast[solve] : (uint32_t)solve;
                rec = aoff ? aoff[solve] : (int64_t)solve * (int64_t)SPAC_LATE_ARG(paths_uniform);
                if (SUB == SUB_CIRCLE) {
                    const double* ac = SPAC_LATE_ARG(circle);
                    if (ac) {
                        R = ac[3 * solve]; cx = ac[3 * solve + 1]; cy = ac[3 * solve + 2];
                    } else {
                        R = a.circle_uniform[0]; cx = a.circle_uniform[1]; cy = a.circle_uniform[2];
                    }
                }
                X = SPAC_LATE_ARG(x0)[solve];
                Yc = SPAC_LATE_ARG(y0)[solve];
                /* sigma(X0), evaluated once per solve before the launch */
                sig = Coef::kConstSigma ? coef.sigma(X, Yc) : SPAC_LATE_ARG(sig0)[solve];
                if constexpr (REFL) {
                    /* initial Dentro (:166): a start beyond a reflecting side is projected */
                    bool dom0;
                    BoundaryHit h0 = nearest_boundary<SUB>(a, R, cx, cy, X, Yc, &dom0);
                    rrefl = dom0 && side_reflects(a.reflecting, h0.nx, h0.ny);
                    rnx = h0.nx; rny = h0.ny; rd = h0.d; chi_p = 0.0;
                    if (rrefl && !(h0.d <= 0.0)) {
                        X = h0.ex; Yc = h0.ey; chi_p = h0.d; rd = 0.0;
                        sig = coef.sigma(X, Yc);
                    }
                }
                Yw = 1.0; Z = 0.0; xi = 0.0; t = 0.0;
                nstep = 0;
                active = true;
            }
            uint32_t take = min(avail, nidle);
            pool_next += take;
            idle = __ballot(!active && !done);
        }
        if (__ballot(active) == 0ull && __ballot(done) == 0ull) break;
        bool reflect_step = false;
        if constexpr (REFL) reflect_step = rrefl;
        if (active && reflect_step) {
            /* Paso_Reflejante: X' = X + b h + sigma dW; within rho of the side the
             * step's largest outward bridge excursion nu (exponential draw, mean
             * 2h) gives the local time chi = max(0, nu + d) and X' -= chi N; a point
             * still beyond a reflecting side is projected (chi += overshoot); Robin
             * data at the foot W = X - d N of X on the side. */
            StepCoef cf = coef.step_s(X, Yc, sig);
            double n0, n1;
            step_normals(seed, stream, path, nstep, &keep0, &keep1, &n0, &n1);
            double dWx = sqrth * n0, dWy = sqrth * n1;
            double xp, yp;
            if (Coef::kHasDrift) {
                xp = X + (cf.bx * h + cf.s * dWx);
                yp = Yc + (cf.by * h + cf.s * dWy);
            } else {
                xp = X + cf.s * dWx;
                yp = Yc + cf.s * dWy;
            }
            double chi = chi_p;
            if (rd > -a.rho) {
                double omega = (2.0 * h) * spac_exponential(seed, stream, path, (uint32_t)nstep);
                double r0 = rnx * cf.s, r1 = rny * cf.s;
                double uc = (r0 * dWx + r1 * dWy) + (rnx * cf.bx + rny * cf.by) * h;
                double nu = 0.5 * (uc + __builtin_sqrt((r0 * r0 + r1 * r1) * omega + uc * uc));
                double v = nu + rd;
                double chil = (v > 0.0) ? v : 0.0;
                xp = xp - chil * rnx;
                yp = yp - chil * rny;
                chi = chi + chil;
            }
            BoundaryHit ph = rectangle_distance(a.dom, xp, yp);
            if (ph.d > 0.0 && side_reflects(a.reflecting, ph.nx, ph.ny)) {
                xp = ph.ex;
                yp = ph.ey;
                chi = chi + ph.d;
            }
            double psi = 0.0, phi = 0.0;
            if (Coef::kHasRobin) coef.robin(X - rd * rnx, Yc - rd * rny, rnx, rny, &psi, &phi);
            if (Coef::kHasRobin) Z = Z + ((cf.f * Yw) * h + (psi * Yw) * chi);
            else Z = Z + (cf.f * Yw) * h;
            if (CV == CV_ANALYTIC) {
                xi = xi + Yw * (cf.Fx * dWx + cf.Fy * dWy);
            } else if (CV == CV_TABLE) {
                int32_t k00;
                double ax, ay;
                spac_cv_cell(&a.cv, X, Yc, &k00, &ax, &ay);
                double Fx = -(cf.s * spac_cv_eval(&a.cv, k00, ax, ay, 1));
                double Fy = -(cf.s * spac_cv_eval(&a.cv, k00, ax, ay, 2));
                xi = xi + Yw * (Fx * dWx + Fy * dWy);
            }
            if (Coef::kHasKilling && Coef::kHasRobin) Yw = Yw + ((cf.c * Yw) * h + (phi * Yw) * chi);
            else if (Coef::kHasKilling) Yw = Yw + (cf.c * Yw) * h;
            else if (Coef::kHasRobin) Yw = Yw + (phi * Yw) * chi;
            X = xp;
            Yc = yp;
            t = t + h;
            nstep++;
        } else if (active) {
            /* Paso_Absorbente (Solvers/FeynmanKacSolver.cpp:7-18) */
            StepCoef cf = coef.step_s(X, Yc, sig);
            double n0, n1;
            step_normals(seed, stream, path, nstep, &keep0, &keep1, &n0, &n1);
            double dWx = sqrth * n0, dWy = sqrth * n1;
            Z = Z + (cf.f * Yw) * h;
            if (CV == CV_ANALYTIC) {
                xi = xi + Yw * (cf.Fx * dWx + cf.Fy * dWy);
            } else if (CV == CV_TABLE) {
                /* F = -sigma grad u_hat with sigma = s I */
                int32_t k00;
                double ax, ay;
                spac_cv_cell(&a.cv, X, Yc, &k00, &ax, &ay);
                double Fx = -(cf.s * spac_cv_eval(&a.cv, k00, ax, ay, 1));
                double Fy = -(cf.s * spac_cv_eval(&a.cv, k00, ax, ay, 2));
                xi = xi + Yw * (Fx * dWx + Fy * dWy);
            }
            if (Coef::kHasKilling) Yw = Yw + (cf.c * Yw) * h;
            if (Coef::kHasDrift) {
                X = X + (cf.bx * h + cf.s * dWx);
                Yc = Yc + (cf.by * h + cf.s * dWy);
            } else {
                X = X + cf.s * dWx;
                Yc = Yc + cf.s * dWy;
            }
            t = t + h;
            nstep++;
        }
        if (active) {
            /* Dentro (Solvers/FeynmanKacSolver.cpp:68-106) */
            bool use_dom;
            BoundaryHit hit = nearest_boundary<SUB>(a, R, cx, cy, X, Yc, &use_dom);
            /* sigma at X inside the boundary, at the projection outside (one evaluation) */
            const double px = (hit.d < 0.0) ? X : hit.ex, py = (hit.d < 0.0) ? Yc : hit.ey;
            double s = coef.sigma(px, py);
            /* a path that goes on is inside (d < gobet |N sigma| sqrt h <= 0): s = sigma(X) */
            sig = a.sigma_fresh ? coef.sigma(X, Yc) : s;
            double r0 = hit.nx * s, r1 = hit.ny * s;
            bool inside = spac_gobet_inside(hit.d, a.gobet * sqrth, r0 * r0 + r1 * r1);
            int8_t out = inside ? (int8_t)SPAC_OUT_INSIDE : (use_dom ? (int8_t)SPAC_OUT_DOMAIN : (int8_t)SPAC_OUT_SUBDOMAIN);
            if constexpr (REFL) {
                /* a reflecting side keeps the path; beyond it, reflejada: project */
                bool rside = use_dom && side_reflects(a.reflecting, hit.nx, hit.ny);
                if (rside) out = (hit.d <= 0.0) ? (int8_t)SPAC_OUT_INSIDE : (int8_t)SPAC_OUT_REFLECTED;
                if ((out == SPAC_OUT_INSIDE || out == SPAC_OUT_REFLECTED) && nstep >= a.max_steps)
                    out = SPAC_OUT_MAX_STEPS;
                rnx = hit.nx; rny = hit.ny; rrefl = rside;
                rd = hit.d; chi_p = 0.0;
                if (out == SPAC_OUT_REFLECTED) {
                    X = hit.ex; Yc = hit.ey; chi_p = hit.d; rd = 0.0;
                    out = SPAC_OUT_INSIDE;
                }
                /* the next step's sigma at the current (possibly projected) X */
                sig = (px == X && py == Yc) ? s : coef.sigma(X, Yc);
            } else {
                if (inside && nstep >= a.max_steps) out = SPAC_OUT_MAX_STEPS;
            }
            if (out != SPAC_OUT_INSIDE) {
                if constexpr (kLaneBatch64 > 1) { /* stored with the wave's next batch of exits */
                    done = true;
                    done_out = out;
                    done_ex = hit.ex;
                    done_ey = hit.ey;
                } else {
                    int64_t g = rec + (int64_t)path;
                    SPAC_LATE_ARG(xt)[g] = hit.ex;
                    SPAC_LATE_ARG(yt)[g] = hit.ey;
                    SPAC_LATE_ARG(Y)[g] = Yw;
                    SPAC_LATE_ARG(Z)[g] = Z;
                    SPAC_LATE_ARG(xi)[g] = xi;
                    SPAC_LATE_ARG(tau)[g] = t;
                    SPAC_LATE_ARG(steps)[g] = nstep; /* gval: fk_exit_values_kernel, lane-dense after the launch */
                    SPAC_LATE_ARG(outcome)[g] = out;
                }
                active = false;
            }
        }
    }
}

/* Mixed-precision path kernel: FP32 stepping and stopping test (hardware
 * transcendentals), FP64 weight Y and sums Z, xi, FP64 exit projection.  Same
 * work queue, lane refill and record layout as the FP64 kernel. */
template <class Coef32, int SUB, int CV>
__global__ void __launch_bounds__(256) fk_path_kernel_fp32(PathArgs a) {
#pragma clang fp contract(fast) /* FP32 path: tolerance-checked, not bit-pinned */
    const Coef32 coef32 = CoefMaker<Coef32>::make(a.tab, a.tab32);
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t seed = a.seed_dev ? *a.seed_dev : a.seed;
    bool active = false;
    bool done = false;   /* stopped, its record not yet stored (batched exits) */
    int8_t done_out = 0; /* the stopped path's outcome */
    int32_t solve = 0;
    uint32_t path = 0;
    int32_t nstep = 0;
    uint32_t keep0 = 0u, keep1 = 0u; /* the odd step's Philox words of the current step pair */
    float xf = 0.f, yf = 0.f, hf = 0.f, sqrthf = 0.f, Rf = 0.f, cxf = 0.f, cyf = 0.f;
    float sigf = 0.f; /* sigma at the current X */
    double Yw = 0.0, Z = 0.0, xi = 0.0, h = 0.0, R = 0.0, cx = 0.0, cy = 0.0;
    uint32_t stream = 0;
    int64_t rec = 0;
    int32_t pool_solve = 0;
    uint32_t pool_next = 0, pool_end = 0;
    int32_t pool_item = 0, pool_item_end = 0; /* the wave's claimed run of items */
    bool drained = false;
    float domf[4] = {(float)a.dom[0], (float)a.dom[1], (float)a.dom[2], (float)a.dom[3]};
    const float gobetf = (float)a.gobet;

    while (true) {
        /* Exits and refills in batches: a stopped lane keeps its final state in registers until
           kLaneBatch32 lanes of the wave have stopped (or no lane steps any more); then all of them
           project and store their records in one pass of the divergent branch, and the idle
           lanes take new paths together.  A path's record depends only on its own state, so
           batching changes no result. */
        {
            const unsigned long long stopped = __ballot(done);
            if (stopped != 0ull && (__popcll(stopped) >= kLaneBatch32 || __ballot(active) == 0ull)) {
                if (done) {
                    /* exit projection in FP64 onto the boundary the FP32 test chose */
                    BoundaryHit hit = (done_out == SPAC_OUT_SUBDOMAIN) ? circle_distance(R, cx, cy, (double)xf, (double)yf)
                                                                       : rectangle_distance(a.dom, (double)xf, (double)yf);
                    int64_t g = rec + (int64_t)path;
                    SPAC_LATE_ARG(xt)[g] = hit.ex;
                    SPAC_LATE_ARG(yt)[g] = hit.ey;
                    SPAC_LATE_ARG(Y)[g] = Yw;
                    SPAC_LATE_ARG(Z)[g] = Z;
                    SPAC_LATE_ARG(xi)[g] = xi;
                    SPAC_LATE_ARG(tau)[g] = (double)nstep * h;
                    SPAC_LATE_ARG(steps)[g] = nstep;
                    SPAC_LATE_ARG(outcome)[g] = done_out;
                    done = false;
                }
            }
        }
        unsigned long long idle = __ballot(!active && !done);
        if (__popcll(idle) < kLaneBatch32 && __ballot(active) != 0ull) idle = 0ull; /* refill later, together */
        if (odd_step_phase(active, nstep)) idle = 0ull;
        while (idle != 0ull && !drained) {
            if (pool_next >= pool_end) {
                if (pool_item >= pool_item_end) {
                    const int32_t item0 = claim_items(a, lane, &pool_item_end);
                    if (item0 >= a.n_items) { drained = true; break; }
                    pool_item = item0;
                }
                const int32_t item = pool_item++;
                if (a.items) {
                    int4 it = a.items[item];
                    pool_solve = it.x;
                    pool_next = (uint32_t)it.y;
                    pool_end = (uint32_t)it.z;
                } else {
                    pool_solve = item / a.chunks_per_solve;
                    uint32_t c = (uint32_t)(item % a.chunks_per_solve);
                    pool_next = c * a.chunk;
                    pool_end = min(pool_next + a.chunk, a.paths_uniform);
                }
            }
            uint32_t avail = pool_end - pool_next;
            uint32_t nidle = (uint32_t)__popcll(idle);
            uint32_t rank = lane_rank(idle);
            if (!active && !done && rank < avail) {
                solve = pool_solve;
                path = pool_next + rank;
                const double* ah = SPAC_LATE_ARG(h);
                const uint32_t* ast = SPAC_LATE_ARG(stream);
                const int64_t* aoff = SPAC_LATE_ARG(rec_off);
                h = ah ? ah[solve] : SPAC_LATE_ARG(h_uniform);
                hf = (float)h;
                sqrthf = __builtin_sqrtf(hf);
                stream = ast ? ast[solve] : (uint32_t)solve;
                rec = aoff ? aoff[solve] : (int64_t)solve * (int64_t)SPAC_LATE_ARG(paths_uniform);
                if (SUB == SUB_CIRCLE) {
                    const double* ac = SPAC_LATE_ARG(circle);
                    if (ac) {
                        R = ac[3 * solve]; cx = ac[3 * solve + 1]; cy = ac[3 * solve + 2];
                    } else {
                        R = a.circle_uniform[0]; cx = a.circle_uniform[1]; cy = a.circle_uniform[2];
                    }
                    Rf = (float)R; cxf = (float)cx; cyf = (float)cy;
                }
                xf = (float)SPAC_LATE_ARG(x0)[solve];
                yf = (float)SPAC_LATE_ARG(y0)[solve];
                sigf = Coef32::kConstSigma ? coef32.sigma(xf, yf) : SPAC_LATE_ARG(sig0f)[solve];
                Yw = 1.0; Z = 0.0; xi = 0.0;
                nstep = 0;
                active = true;
            }
            uint32_t take = min(avail, nidle);
            pool_next += take;
            idle = __ballot(!active && !done);
        }
        if (__ballot(active) == 0ull && __ballot(done) == 0ull) break;
        wave_priority_by_age(active, nstep);
        if (active) {
            StepCoef32 cf = coef32.step_s(xf, yf, sigf);
            float n0, n1;
            step_normals32(seed, stream, path, nstep, &keep0, &keep1, &n0, &n1);
            const float dWx = sqrthf * n0, dWy = sqrthf * n1;
            const float Ywf = (float)Yw;
            Z += (double)((cf.f * Ywf) * hf);
            if (CV == CV_ANALYTIC) {
                xi += (double)(Ywf * (cf.Fx * dWx + cf.Fy * dWy));
            } else if (CV == CV_TABLE || CV == CV_NETWORK) {
                float gx, gy;
                if (CV == CV_NETWORK) net_grad32(a.cvnet, xf, yf, &gx, &gy);
                else cv_grad32(a.cv32, xf, yf, &gx, &gy);
                xi += (double)(Ywf * (-cf.s) * (gx * dWx + gy * dWy));
            }
            if (Coef32::kHasKilling) Yw += (double)(cf.c * hf) * Yw;
            if (Coef32::kHasDrift) {
                xf += cf.bx * hf + cf.s * dWx;
                yf += cf.by * hf + cf.s * dWy;
            } else {
                xf += cf.s * dWx;
                yf += cf.s * dWy;
            }
            nstep++;
            float nx, ny, ex, ey;
            float d = rect_distance32(domf, xf, yf, &nx, &ny, &ex, &ey);
            bool use_dom = true;
            if (SUB == SUB_CIRCLE) {
                float dx = xf - cxf, dy = yf - cyf;
                float rr = __builtin_sqrtf(dx * dx + dy * dy);
                rr = (5E-4f < rr) ? rr : 5E-4f;
                float ir = __builtin_amdgcn_rcpf(rr);
                float ds = rr - Rf;
                use_dom = d >= ds;
                if (!use_dom) { d = ds; nx = dx * ir; ny = dy * ir; ex = nx * Rf + cxf; ey = ny * Rf + cyf; }
            }
            const float pxf = (d < 0.0f) ? xf : ex, pyf = (d < 0.0f) ? yf : ey;
            float sg = coef32.sigma(pxf, pyf);
            sigf = a.sigma_fresh ? coef32.sigma(xf, yf) : sg;
            float r0 = nx * sg, r1 = ny * sg;
            /* d < g |N sigma| sqrt h without the square root when g <= 0 (both sides <= 0) */
            const float kf = gobetf * sqrthf, n2 = r0 * r0 + r1 * r1;
            bool inside = (kf <= 0.0f) ? (d < 0.0f && d * d > (kf * kf) * n2) : (d < kf * __builtin_sqrtf(n2));
            int8_t out = inside ? (int8_t)SPAC_OUT_INSIDE : (use_dom ? (int8_t)SPAC_OUT_DOMAIN : (int8_t)SPAC_OUT_SUBDOMAIN);
            if (inside && nstep >= a.max_steps) out = SPAC_OUT_MAX_STEPS;
            if (out != SPAC_OUT_INSIDE) { /* the record is stored with the wave's next batch of exits */
                done = true;
                done_out = out;
                active = false;
            }
        }
    }
}

/* ------------------------------------------- per-solve set-up, exits -- */
/* sigma(X0) of every solve in FP64 and in FP32 (the FP32 kernel's rounding of X0). */
template <class Coef, class Coef32>
__global__ void __launch_bounds__(256) fk_solve_sigma_kernel(int32_t n, const double* x0, const double* y0, double* sig0,
                                                             float* sig0f, const spac_coef_table tab,
                                                             const CoefTable32 tab32) {
    const int32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const Coef coef = CoefMaker<Coef>::make(tab, tab32);
    const Coef32 coef32 = CoefMaker<Coef32>::make(tab, tab32);
    const double x = x0[k], y = y0[k];
    sig0[k] = coef.sigma(x, y);
    sig0f[k] = coef32.sigma((float)x, (float)y);
}

/* g(X_tau) for domain exits and capped paths, u(X_tau) (or u_hat) for subdomain
 * exits (Solvers/FeynmanKacSolver.cpp:249-254): one lane per record, so the four
 * Franke exponentials of an exit no longer run under a divergent branch in the
 * path kernel. */
template <class Coef, int CV>
__global__ void __launch_bounds__(256) fk_exit_values_kernel(PathArgs a, int64_t total) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    const Coef coef = CoefMaker<Coef>::make(a.tab, a.tab32);
    for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += stride) {
        const double x = a.xt[g], y = a.yt[g];
        a.gval[g] = (a.outcome[g] == SPAC_OUT_SUBDOMAIN) ? exit_u<Coef, CV>(a, coef, x, y) : coef.g(x, y);
    }
}

/* ---------------------------------------------------------- reduction -- */
struct ReduceArgs {
    int32_t n_solves;
    int32_t node_mode;
    const uint32_t* paths; /* per solve or nullptr: paths_uniform */
    uint32_t paths_uniform;
    const int64_t* rec_off;
    const double* Y;
    const double* Z;
    const double* xi;
    const double* tau;
    const double* gval;
    const int32_t* steps;
    const int8_t* outcome;
    spac_point_stats* stats;
    int32_t cv_fallback;
    double cv_max_ratio;
};

#define NSUMS 10
/* paths whose records a lane of the reduction loads together */
constexpr int kReducePaths = 4;
enum { S_SCORE, S_SCORE2, S_XI, S_XI2, S_VR, S_VR2, S_XISCORE, S_RNG, S_TAU, S_B };

__global__ void __launch_bounds__(64) fk_reduce_kernel(ReduceArgs a) {
    const int solve = blockIdx.x;
    const int lane = threadIdx.x;
    if (solve >= a.n_solves) return;
    int64_t np = a.paths ? (int64_t)a.paths[solve] : (int64_t)a.paths_uniform;
    int64_t off = a.rec_off ? a.rec_off[solve] : (int64_t)solve * (int64_t)a.paths_uniform;
    /* lane l sums paths l, l + 64, l + 128, ... in order; the 64 lane sums are
     * then folded pairwise (lane i += lane i + off, off = 32, 16, ..., 1) */
    spac_ksum tot[NSUMS];
#pragma unroll
    for (int k = 0; k < NSUMS; ++k) tot[k] = spac_ksum_zero();
    int64_t my_steps = 0, my_max = 0, my_dom = 0;
    /* kReducePaths of the lane's paths are loaded before any of them is summed (one round
       trip to memory per group); past the solve's end the lane reloads its first path */
    for (int64_t p0 = lane; p0 < np; p0 += 64 * kReducePaths) {
        double Yv[kReducePaths], Zv[kReducePaths], xv[kReducePaths], gvv[kReducePaths], tv[kReducePaths];
        int32_t sv[kReducePaths];
        int8_t ov[kReducePaths];
#pragma unroll
        for (int u = 0; u < kReducePaths; ++u) {
            const int64_t p = p0 + 64 * u;
            const int64_t g = off + (p < np ? p : p0);
            Yv[u] = a.Y[g]; Zv[u] = a.Z[g]; xv[u] = a.xi[g]; gvv[u] = a.gval[g]; tv[u] = a.tau[g];
            sv[u] = a.steps[g];
            ov[u] = a.outcome[g];
        }
#pragma unroll
        for (int u = 0; u < kReducePaths; ++u) {
            if (p0 + 64 * u >= np) break;
            const double Y = Yv[u], Z = Zv[u], xi = xv[u], gv = gvv[u];
            const int32_t st = sv[u];
            const int8_t oc = ov[u];
            const double score = Z + Y * gv;
            const double vr = score + xi;
            double v[NSUMS];
            v[S_SCORE] = score;
            v[S_SCORE2] = score * score;
            v[S_XI] = xi;
            v[S_XI2] = xi * xi;
            v[S_VR] = vr;
            v[S_VR2] = vr * vr;
            v[S_XISCORE] = xi * score;
            v[S_RNG] = 2.0 * (double)st;
            v[S_TAU] = tv[u];
            v[S_B] = a.node_mode ? ((oc == SPAC_OUT_SUBDOMAIN) ? (Z + xi) : vr) : 0.0;
#pragma unroll
            for (int k = 0; k < NSUMS; ++k) tot[k] = spac_ksum_add(tot[k], v[k]);
            my_steps += st;
            my_max += (oc == SPAC_OUT_MAX_STEPS);
            my_dom += (oc != SPAC_OUT_SUBDOMAIN);
        }
    }
    /* The fold (lane l += lane l + off, off = 32, ..., 1) of every sum, through LDS: a level's
       NSUMS x off merges are spread over the wave's lanes instead of every lane merging every
       sum at every level.  The same pairs are merged in the same order, so the totals are those
       of a shuffle tree; a merge reads slots l and l + off and writes slot l, and no other merge
       of its level touches either. */
    __shared__ double s_fold[2][NSUMS][64];
#pragma unroll
    for (int k = 0; k < NSUMS; ++k) {
        s_fold[0][k][lane] = tot[k].s;
        s_fold[1][k][lane] = tot[k].c;
    }
    __syncthreads();
#pragma unroll
    for (int lg = 5; lg >= 0; --lg) {
        const int off = 1 << lg;
        for (int m = lane; m < NSUMS * off; m += 64) {
            const int k = m >> lg, l = m & (off - 1);
            spac_ksum x, y;
            x.s = s_fold[0][k][l]; x.c = s_fold[1][k][l];
            y.s = s_fold[0][k][l + off]; y.c = s_fold[1][k][l + off];
            const spac_ksum r = spac_ksum_merge(x, y);
            s_fold[0][k][l] = r.s;
            s_fold[1][k][l] = r.c;
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < NSUMS; ++k) {
        tot[k].s = s_fold[0][k][0];
        tot[k].c = s_fold[1][k][0];
    }
    /* exact integer totals */
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        my_steps += __shfl_xor(my_steps, o, 64);
        my_max += __shfl_xor(my_max, o, 64);
        my_dom += __shfl_xor(my_dom, o, 64);
    }
    if (lane == 0 && np == 0) {
        spac_point_stats st;
        memset(&st, 0, sizeof st);
        a.stats[solve] = st;
        return;
    }
    if (lane == 0) {
        double N = (double)np;
        double s[NSUMS];
        bool has_nan = false;
#pragma unroll
        for (int k = 0; k < NSUMS; ++k) {
            s[k] = spac_ksum_value(tot[k]);
            if (s[k] != s[k]) has_nan = true;
        }
        spac_point_stats st;
        st.phi = s[S_SCORE] / N;
        st.phiphi = s[S_SCORE2] / N;
        st.xi = s[S_XI] / N;
        st.xixi = s[S_XI2] / N;
        st.phi_VR = s[S_VR] / N;
        st.phiphi_VR = s[S_VR2] / N;
        st.xiphi = s[S_XISCORE] / N;
        st.RNGC = s[S_RNG];
        st.APL = s[S_TAU] / N;
        st.B = a.node_mode ? s[S_B] / N : 0.0;
        st.paths = np;
        st.path_steps = my_steps;
        st.n_max_steps = my_max;
        st.n_domain_exits = (int32_t)my_dom;
        st.status = has_nan ? SPAC_E_NAN : (my_max ? SPAC_E_MAX_STEPS : SPAC_OK);
        st.cv_fallback = 0;
        st.reserved = 0;
        if (a.cv_fallback) {
            /* keep the control variate only where it does not add variance */
            double var = st.phiphi - st.phi * st.phi;
            double var_vr = st.phiphi_VR - st.phi_VR * st.phi_VR;
            if (var_vr > a.cv_max_ratio * var) {
                st.cv_fallback = 1;
                if (a.node_mode) st.B = (s[S_B] - s[S_XI]) / N;
            }
        }
        a.stats[solve] = st;
    }
}

/* ------------------------------------------------------ host side -- */
template <class Coef, int CV, bool REFL>
static void launch_fp64(int sub, dim3 grid, dim3 block, hipStream_t s, const PathArgs& a) {
    if (sub == SUB_CIRCLE) hipLaunchKernelGGL((fk_path_kernel_fp64<Coef, SUB_CIRCLE, CV, REFL>), grid, block, 0, s, a);
    else hipLaunchKernelGGL((fk_path_kernel_fp64<Coef, SUB_NONE, CV, REFL>), grid, block, 0, s, a);
}

template <int CV>
static void launch_paths_cv(int coef, int precision, int sub, bool refl, dim3 grid, dim3 block, hipStream_t s,
                            const PathArgs& a) {
    if (coef == kCoefTabulated) { /* absorbing sides only (checked when the table is installed) */
        if (precision == SPAC_FP32) {
            if (sub == SUB_CIRCLE)
                hipLaunchKernelGGL((fk_path_kernel_fp32<CoefTab32<CoefVarFranke32>, SUB_CIRCLE, CV>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((fk_path_kernel_fp32<CoefTab32<CoefVarFranke32>, SUB_NONE, CV>), grid, block, 0, s, a);
        } else {
            launch_fp64<CoefTab<CoefVarFranke>, CV, false>(sub, grid, block, s, a);
        }
        return;
    }
    if (precision == SPAC_FP32) {
        /* FRANKE_ROBIN differs from FRANKE_POISSON only on reflecting sides (FP64 only) */
        if (coef != SPAC_COEF_VARCOEF_FRANKE) {
            if (sub == SUB_CIRCLE) hipLaunchKernelGGL((fk_path_kernel_fp32<CoefFranke32, SUB_CIRCLE, CV>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((fk_path_kernel_fp32<CoefFranke32, SUB_NONE, CV>), grid, block, 0, s, a);
        } else {
            if (sub == SUB_CIRCLE) hipLaunchKernelGGL((fk_path_kernel_fp32<CoefVarFranke32, SUB_CIRCLE, CV>), grid, block, 0, s, a);
            else hipLaunchKernelGGL((fk_path_kernel_fp32<CoefVarFranke32, SUB_NONE, CV>), grid, block, 0, s, a);
        }
        return;
    }
    if (refl) {
        if (coef == SPAC_COEF_FRANKE_POISSON) launch_fp64<CoefFranke, CV, true>(sub, grid, block, s, a);
        else if (coef == SPAC_COEF_FRANKE_ROBIN) launch_fp64<CoefFrankeRobin, CV, true>(sub, grid, block, s, a);
        else launch_fp64<CoefVarFranke, CV, true>(sub, grid, block, s, a);
        return;
    }
    if (coef != SPAC_COEF_VARCOEF_FRANKE) launch_fp64<CoefFranke, CV, false>(sub, grid, block, s, a);
    else launch_fp64<CoefVarFranke, CV, false>(sub, grid, block, s, a);
}

static void launch_paths(int coef, int precision, int sub, int cv, bool refl, dim3 grid, dim3 block, hipStream_t s,
                         const PathArgs& a) {
    if (cv == CV_NETWORK) { /* FP32, the compiled variable-coefficient set (checked at installation) */
        if (sub == SUB_CIRCLE)
            hipLaunchKernelGGL((fk_path_kernel_fp32<CoefVarFranke32, SUB_CIRCLE, CV_NETWORK>), grid, block, 0, s, a);
        else hipLaunchKernelGGL((fk_path_kernel_fp32<CoefVarFranke32, SUB_NONE, CV_NETWORK>), grid, block, 0, s, a);
        return;
    }
    if (cv == CV_NONE) launch_paths_cv<CV_NONE>(coef, precision, sub, refl, grid, block, s, a);
    else if (cv == CV_TABLE) launch_paths_cv<CV_TABLE>(coef, precision, sub, refl, grid, block, s, a);
    else launch_paths_cv<CV_ANALYTIC>(coef, precision, sub, refl, grid, block, s, a);
}

spac_status DeviceBuffer::ensure(size_t bytes) {
    if (bytes <= cap) return SPAC_OK;
    if (ptr) { (void)hipFree(ptr); ptr = nullptr; cap = 0; }
    size_t want = std::max(bytes, (size_t)256);
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        return set_error(SPAC_E_OUT_OF_MEMORY, "hipMalloc(%zu) failed: %s", want, hipGetErrorString(e));
    }
    cap = want;
    return SPAC_OK;
}
void DeviceBuffer::release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    cap = 0;
}

spac_status Problem::ensure_records(int64_t total_paths) {
    spac_status rc;
    size_t n = (size_t)std::max<int64_t>(total_paths, 1);
    if ((rc = rec_xt.ensure(n * 8))) return rc;
    if ((rc = rec_yt.ensure(n * 8))) return rc;
    if ((rc = rec_Y.ensure(n * 8))) return rc;
    if ((rc = rec_Z.ensure(n * 8))) return rc;
    if ((rc = rec_xi.ensure(n * 8))) return rc;
    if ((rc = rec_tau.ensure(n * 8))) return rc;
    if ((rc = rec_gval.ensure(n * 8))) return rc;
    if ((rc = rec_steps.ensure(n * 4))) return rc;
    if ((rc = rec_outcome.ensure(n))) return rc;
    return SPAC_OK;
}

/* Paths per work item: the launch's chunk size, enlarged for very large batches so
 * that the queue holds at most ~2^20 items (every item costs one atomic on a single
 * counter; 1.6e7 of them cost ~120 ms).  Results do not depend on it. */
uint32_t work_item_paths(uint32_t chunk, int64_t total_paths) {
    const int64_t kMaxItems = (int64_t)1 << 20;
    const int64_t c = (total_paths + kMaxItems - 1) / kMaxItems;
    return (uint32_t)std::max<int64_t>((int64_t)chunk, c);
}

spac_status Problem::ensure_stage(size_t bytes) {
    if (bytes <= stage_cap) return SPAC_OK;
    if (stage) (void)hipHostFree(stage);
    stage = nullptr;
    stage_cap = 0;
    const size_t cap = bytes + bytes / 4;
    SPAC_HIP_CHECK(hipHostMalloc(&stage, cap, hipHostMallocDefault));
    stage_cap = cap;
    return SPAC_OK;
}

spac_status Problem::copy_out(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return SPAC_OK;
    if (bytes < ((size_t)1 << 20)) {
        SPAC_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
        SPAC_HIP_CHECK(hipStreamSynchronize(stream));
        return SPAC_OK;
    }
    spac_status rc;
    if ((rc = ensure_stage(bytes))) return rc;
    SPAC_HIP_CHECK(hipMemcpyAsync(stage, src, bytes, hipMemcpyDeviceToHost, stream));
    SPAC_HIP_CHECK(hipStreamSynchronize(stream));
    std::memcpy(dst, stage, bytes);
    return SPAC_OK;
}

/* Upload a host batch of solves into the handle's solve table and build the
 * work queue (solves with zero paths get no work items). */
spac_status Problem::upload_solves(int32_t n, const double* x0, const double* y0, const uint32_t* paths,
                                   const double* h, const uint32_t* strm, const double* circle, SolveSet* out) {
    int64_t total = 0;
    size_t n_items = 0;
    for (int k = 0; k < n; ++k) total += paths[k];
    const uint32_t chunk = work_item_paths((uint32_t)launch.chunk_paths, total);
    for (int k = 0; k < n; ++k) {
        if (!(h[k] > 0.0) || !(h[k] < 1e30)) return set_error(SPAC_E_INVALID_ARG, "solve %d has bad h", k);
        if (circle && !(circle[3 * k] > 0.0))
            return set_error(SPAC_E_INVALID_ARG, "solve %d has a circle of radius <= 0", k);
        n_items += (paths[k] + chunk - 1) / chunk;
    }
    if (n_items > (size_t)kMaxWorkItems) return set_error(SPAC_E_INVALID_ARG, "too many work items");
    /* every array is built in (or copied to) one pinned staging block and goes up by DMA */
    const size_t nb = (size_t)n;
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    const size_t o_x = 0, o_y = o_x + al(nb * 8), o_h = o_y + al(nb * 8), o_st = o_h + al(nb * 8),
                 o_p = o_st + al(nb * 4), o_off = o_p + al(nb * 4), o_c = o_off + al((nb + 1) * 8),
                 o_it = o_c + al(nb * 24), o_sq = o_it + al(std::max<size_t>(n_items, 1) * sizeof(int4)),
                 o_end = o_sq + al(nb * 8);
    spac_status rc;
    if ((rc = ensure_stage(o_end))) return rc;
    char* sb = (char*)stage;
    std::memcpy(sb + o_x, x0, nb * 8);
    std::memcpy(sb + o_y, y0, nb * 8);
    std::memcpy(sb + o_h, h, nb * 8);
    std::memcpy(sb + o_st, strm, nb * 4);
    std::memcpy(sb + o_p, paths, nb * 4);
    if (circle) std::memcpy(sb + o_c, circle, nb * 24);
    int64_t* off = (int64_t*)(sb + o_off);
    int4* items = (int4*)(sb + o_it);
    double* sq = (double*)(sb + o_sq);
    off[0] = 0;
    size_t q = 0;
    for (int k = 0; k < n; ++k) {
        off[k + 1] = off[k] + paths[k];
        for (uint32_t p0 = 0; p0 < paths[k]; p0 += chunk) items[q++] = make_int4(k, (int)p0, (int)std::min(p0 + chunk, paths[k]), 0);
        sq[k] = std::sqrt(h[k]); /* correctly rounded, as the device's */
    }
    if (launch.item_order == SPAC_ORDER_COST)
        std::stable_sort(items, items + n_items, [&](const int4& u, const int4& v) { return h[u.x] < h[v.x]; });
    if ((rc = sx.ensure(nb * 8))) return rc;
    if ((rc = sy.ensure(nb * 8))) return rc;
    if ((rc = sh.ensure(nb * 8))) return rc;
    if ((rc = sstream.ensure(nb * 4))) return rc;
    if ((rc = spaths.ensure(nb * 4))) return rc;
    if ((rc = soff.ensure((nb + 1) * 8))) return rc;
    if ((rc = scircle.ensure(nb * 24))) return rc;
    if ((rc = sitems.ensure(std::max<size_t>(n_items, 1) * sizeof(int4)))) return rc;
    if ((rc = ssq.ensure(nb * 8))) return rc;
    hipStream_t s = stream;
    SPAC_HIP_CHECK(hipMemcpyAsync(sx.ptr, sb + o_x, nb * 8, hipMemcpyHostToDevice, s));
    SPAC_HIP_CHECK(hipMemcpyAsync(sy.ptr, sb + o_y, nb * 8, hipMemcpyHostToDevice, s));
    SPAC_HIP_CHECK(hipMemcpyAsync(sh.ptr, sb + o_h, nb * 8, hipMemcpyHostToDevice, s));
    SPAC_HIP_CHECK(hipMemcpyAsync(sstream.ptr, sb + o_st, nb * 4, hipMemcpyHostToDevice, s));
    SPAC_HIP_CHECK(hipMemcpyAsync(spaths.ptr, sb + o_p, nb * 4, hipMemcpyHostToDevice, s));
    SPAC_HIP_CHECK(hipMemcpyAsync(soff.ptr, off, (nb + 1) * 8, hipMemcpyHostToDevice, s));
    if (circle) SPAC_HIP_CHECK(hipMemcpyAsync(scircle.ptr, sb + o_c, nb * 24, hipMemcpyHostToDevice, s));
    if (n_items > 0) SPAC_HIP_CHECK(hipMemcpyAsync(sitems.ptr, items, n_items * sizeof(int4), hipMemcpyHostToDevice, s));
    SPAC_HIP_CHECK(hipMemcpyAsync(ssq.ptr, sq, nb * 8, hipMemcpyHostToDevice, s));
    /* the staging block is rewritten by the next call: wait for the copies */
    SPAC_HIP_CHECK(hipStreamSynchronize(s));
    const int64_t total_paths = off[n];
    SolveSet ss;
    ss.n = n;
    ss.x0 = (const double*)sx.ptr;
    ss.y0 = (const double*)sy.ptr;
    ss.h = (const double*)sh.ptr;
    ss.stream = (const uint32_t*)sstream.ptr;
    ss.paths = (const uint32_t*)spaths.ptr;
    ss.rec_off = (const int64_t*)soff.ptr;
    ss.circle = circle ? (const double*)scircle.ptr : nullptr;
    ss.items = (const int4*)sitems.ptr;
    ss.n_items = (int32_t)n_items;
    ss.total_paths = total_paths;
    ss.chunk = chunk;
    ss.sqrth = (const double*)ssq.ptr;
    if ((rc = ensure_solve_scratch(n, &ss))) return rc;
    *out = ss;
    return SPAC_OK;
}

static spac_status check_problem(const spac_problem_desc* d) {
    if (d->coef_kind != SPAC_COEF_FRANKE_POISSON && d->coef_kind != SPAC_COEF_VARCOEF_FRANKE &&
        d->coef_kind != SPAC_COEF_FRANKE_ROBIN)
        return set_error(SPAC_E_UNREGISTERED_PLUGIN, "coefficient set %d has no device implementation", d->coef_kind);
    if (d->domain_kind != SPAC_BOUNDARY_RECTANGLE)
        return set_error(SPAC_E_UNREGISTERED_PLUGIN, "domain boundary kind %d has no device implementation", d->domain_kind);
    if (d->subdomain_kind != SPAC_BOUNDARY_NONE && d->subdomain_kind != SPAC_BOUNDARY_CIRCLE)
        return set_error(SPAC_E_UNREGISTERED_PLUGIN, "subdomain boundary kind %d has no device implementation",
                         d->subdomain_kind);
    if (d->precision != SPAC_FP64 && d->precision != SPAC_FP32)
        return set_error(SPAC_E_INVALID_ARG, "precision %d unknown", d->precision);
    if (!(d->domain_params[2] > d->domain_params[0]) || !(d->domain_params[3] > d->domain_params[1]))
        return set_error(SPAC_E_INVALID_ARG, "rectangle domain must have xmax > xmin and ymax > ymin");
    if (d->subdomain_kind == SPAC_BOUNDARY_CIRCLE && !(d->subdomain_params[0] > 0.0))
        return set_error(SPAC_E_INVALID_ARG, "circle subdomain radius must be > 0");
    if (d->reflecting & ~(SPAC_REFLECT_XMIN | SPAC_REFLECT_YMIN | SPAC_REFLECT_XMAX | SPAC_REFLECT_YMAX))
        return set_error(SPAC_E_INVALID_ARG, "reflecting mask 0x%x has bits outside SPAC_REFLECT_*", d->reflecting);
    if (d->reflecting && d->precision != SPAC_FP64)
        return set_error(SPAC_E_INVALID_ARG, "reflecting sides need SPAC_FP64 stepping");
    if (d->reflecting && !(d->reflect_rho >= 0.0 && d->reflect_rho < 1e300))
        return set_error(SPAC_E_INVALID_ARG, "reflect_rho must be finite and >= 0");
    return SPAC_OK;
}

/* Queue reset and the persistent path kernel (no synchronisation). */
template <int CV>
static void launch_exit_values(int coef, dim3 grid, hipStream_t s, const PathArgs& a, int64_t total) {
    if (coef == kCoefTabulated)
        hipLaunchKernelGGL((fk_exit_values_kernel<CoefTab<CoefVarFranke>, CV>), grid, dim3(256), 0, s, a, total);
    else if (coef == SPAC_COEF_VARCOEF_FRANKE)
        hipLaunchKernelGGL((fk_exit_values_kernel<CoefVarFranke, CV>), grid, dim3(256), 0, s, a, total);
    else
        hipLaunchKernelGGL((fk_exit_values_kernel<CoefFranke, CV>), grid, dim3(256), 0, s, a, total);
}

spac_status Problem::ensure_solve_scratch(int32_t n, SolveSet* ss) {
    spac_status rc;
    if ((rc = ssig.ensure((size_t)std::max(n, 1) * 8))) return rc;
    if ((rc = ssigf.ensure((size_t)std::max(n, 1) * 4))) return rc;
    ss->sig0 = (double*)ssig.ptr;
    ss->sig0f = (float*)ssigf.ptr;
    if (ss->items && ss->n_items > 1) { /* room for the tail-first order, whatever the launch order */
        if ((rc = sitems_tail.ensure((size_t)ss->n_items * sizeof(int4)))) return rc;
        if ((rc = stail_hist.ensure((size_t)kTailBlocks * kTailBuckets * 4))) return rc;
        ss->items_tail = (int4*)sitems_tail.ptr;
        ss->tail_hist = (int32_t*)stail_hist.ptr;
    }
    return SPAC_OK;
}

spac_status no_pending(const Problem* p) {
    if (p->pending.kind != Problem::PENDING_NONE)
        return set_error(SPAC_E_INVALID_ARG, "an asynchronous call is pending on this problem: call spac_fk_wait first");
    return SPAC_OK;
}

RecordSet Problem::own_records() const {
    RecordSet r;
    r.xt = (double*)rec_xt.ptr; r.yt = (double*)rec_yt.ptr;
    r.Y = (double*)rec_Y.ptr; r.Z = (double*)rec_Z.ptr; r.xi = (double*)rec_xi.ptr;
    r.tau = (double*)rec_tau.ptr; r.gval = (double*)rec_gval.ptr;
    r.steps = (int32_t*)rec_steps.ptr; r.outcome = (int8_t*)rec_outcome.ptr;
    r.queue = (int32_t*)queue.ptr;
    return r;
}

spac_status Problem::enqueue_paths(const SolveSet& ss, uint64_t seed, const uint64_t* sdev, int* grid_blocks,
                                   hipEvent_t after_path, const RecordSet* recs) {
    const RecordSet rs = recs ? *recs : own_records();
    PathArgs a;
    memset(&a, 0, sizeof a);
    for (int i = 0; i < 4; ++i) a.dom[i] = desc.domain_params[i];
    a.gobet = desc.gobet_constant;
    a.sigma_fresh = desc.gobet_constant >= 0.0 ? 1 : 0;
    a.seed = seed;
    a.seed_dev = sdev;
    a.max_steps = desc.max_steps > 0 ? desc.max_steps : 0x7fffffff;
    a.reflecting = desc.reflecting;
    a.rho = desc.reflect_rho;
    a.x0 = ss.x0; a.y0 = ss.y0;
    a.h = ss.h; a.h_uniform = ss.h_uniform;
    a.sqrth = ss.sqrth; a.sqrth_uniform = ss.sqrth_uniform;
    a.sig0 = ss.sig0; a.sig0f = ss.sig0f;
    a.stream = ss.stream;
    a.circle = ss.circle;
    for (int i = 0; i < 3; ++i) a.circle_uniform[i] = desc.subdomain_params[i];
    a.rec_off = ss.rec_off;
    a.paths_uniform = ss.paths_uniform;
    a.chunk = ss.chunk ? ss.chunk : (uint32_t)launch.chunk_paths;
    a.items = ss.items;
    a.n_items = ss.n_items;
    a.chunks_per_solve = ss.chunks_per_solve;
    a.queue = rs.queue;
    a.xt = rs.xt; a.yt = rs.yt;
    a.Y = rs.Y; a.Z = rs.Z; a.xi = rs.xi;
    a.tau = rs.tau; a.gval = rs.gval;
    a.steps = rs.steps; a.outcome = rs.outcome;
    if (cv_mode == SPAC_CV_TABLE || cv_mode == SPAC_CV_NETWORK) {
        a.cv.table = (const float*)cv_table.ptr;
        a.cv.t = cv_t;
        a.cv.lox = cv_dom[0];
        a.cv.loy = cv_dom[1];
        a.cv.ix = (double)(cv_t - 1) / (cv_dom[2] - cv_dom[0]);
        a.cv.iy = (double)(cv_t - 1) / (cv_dom[3] - cv_dom[1]);
        a.cv32.cells = (const float4*)cv_grad.ptr;
        a.cv32.t = cv_t;
        a.cv32.lox = (float)a.cv.lox;
        a.cv32.loy = (float)a.cv.loy;
        a.cv32.ix = (float)a.cv.ix;
        a.cv32.iy = (float)a.cv.iy;
        a.cv32.top = (float)(cv_t - 1);
    }
    if (cv_mode == SPAC_CV_NETWORK) {
        a.cvnet.P = (const float*)cv_net.ptr;
        a.cvnet.lox = cv_net_map[0];
        a.cvnet.loy = cv_net_map[1];
        a.cvnet.sx = cv_net_map[2];
        a.cvnet.sy = cv_net_map[3];
    }
    if (tab_mask) {
        a.tab.v = (const double*)tab.ptr;
        a.tab.t = tab_t;
        a.tab.mask = tab_mask;
        a.tab.lox = tab_dom[0];
        a.tab.loy = tab_dom[1];
        a.tab.ix = (double)(tab_t - 1) / (tab_dom[2] - tab_dom[0]);
        a.tab.iy = (double)(tab_t - 1) / (tab_dom[3] - tab_dom[1]);
        a.tab.need_F = cv_mode == SPAC_CV_ANALYTIC ? 1 : 0;
        a.tab32.v = (const float*)tab32.ptr;
        a.tab32.t = tab_t;
        a.tab32.mask = tab_mask;
        a.tab32.need_F = a.tab.need_F;
        a.tab32.lox = (float)a.tab.lox;
        a.tab32.loy = (float)a.tab.loy;
        a.tab32.ix = (float)a.tab.ix;
        a.tab32.iy = (float)a.tab.iy;
        a.tab32.top = (float)(tab_t - 1);
    }
    const int coef = coef_kernels();
    int sub = (ss.circle || desc.subdomain_kind == SPAC_BOUNDARY_CIRCLE) ? SUB_CIRCLE : SUB_NONE;
    int block = launch.block_size;
    int blocks = num_cus * launch.blocks_per_cu;
    int64_t max_blocks = std::max<int64_t>(1, ((int64_t)ss.n_items + (block / 64) - 1) / (block / 64));
    if (blocks > max_blocks) blocks = (int)max_blocks;
    if (!ss.sqrth && !(ss.sqrth_uniform > 0.0)) return set_error(SPAC_E_INVALID_ARG, "solve set without sqrt(h)");
    if (coef == SPAC_COEF_VARCOEF_FRANKE || coef == kCoefTabulated) {
        if (!ss.sig0 || !ss.sig0f) return set_error(SPAC_E_INVALID_ARG, "solve set without sigma(X0) slots");
        if (coef == kCoefTabulated)
            hipLaunchKernelGGL((fk_solve_sigma_kernel<CoefTab<CoefVarFranke>, CoefTab32<CoefVarFranke32>>),
                               dim3((ss.n + 255) / 256), dim3(256), 0, stream, ss.n, ss.x0, ss.y0, ss.sig0, ss.sig0f,
                               a.tab, a.tab32);
        else
            hipLaunchKernelGGL((fk_solve_sigma_kernel<CoefVarFranke, CoefVarFranke32>), dim3((ss.n + 255) / 256),
                               dim3(256), 0, stream, ss.n, ss.x0, ss.y0, ss.sig0, ss.sig0f, a.tab, a.tab32);
    }
    /* tail-first order: asked for, or by default where the longest paths set the end of the
       launch (few paths per lane; measured: learned CV at 64 x 64 circles, 33 paths per lane,
       9.3 -> 7.8 ms; an uneven batch of ~2,000 paths per lane 92 -> 99 ms, so node order there) */
    const bool tail_order = launch.item_order == SPAC_ORDER_TAIL ||
                            (launch.item_order == SPAC_ORDER_AUTO &&
                             ss.total_paths <= kTailAutoPathsPerLane * (int64_t)blocks * (int64_t)block);
    if (tail_order && ss.items && ss.items_tail && ss.tail_hist && ss.n_items > 1) {
        TailKey k;
        k.sig0 = (coef == SPAC_COEF_VARCOEF_FRANKE || coef == kCoefTabulated) ? ss.sig0 : nullptr;
        k.h = ss.h;
        k.h_uniform = ss.h_uniform;
        k.circle = ss.circle;
        k.R_uniform = (sub == SUB_CIRCLE && !ss.circle) ? desc.subdomain_params[0] : 1.0;
        hipLaunchKernelGGL(k_tail_hist, dim3(kTailBlocks), dim3(kTailBuckets), 0, stream, ss.n_items, ss.items, k,
                           ss.tail_hist);
        hipLaunchKernelGGL(k_tail_offsets, dim3(1), dim3(kTailBuckets), 0, stream, ss.tail_hist);
        hipLaunchKernelGGL(k_tail_scatter, dim3(kTailBlocks), dim3(kTailBuckets), 0, stream, ss.n_items, ss.items, k,
                           (const int32_t*)ss.tail_hist, ss.items_tail);
        a.items = ss.items_tail;
    }
    hipLaunchKernelGGL(zero_counter_kernel, dim3(1), dim3(64), 0, stream, a.queue);
    if (cv_mode == SPAC_CV_NETWORK && (desc.precision != SPAC_FP32 || coef != SPAC_COEF_VARCOEF_FRANKE))
        return set_error(SPAC_E_INVALID_ARG, "SPAC_CV_NETWORK needs FP32 stepping and the compiled variable-coefficient set");
    launch_paths(coef, desc.precision, sub, cv_mode, desc.reflecting != 0, dim3(blocks), dim3(block), stream, a);
    SPAC_HIP_CHECK(hipGetLastError());
    if (after_path) SPAC_HIP_CHECK(hipEventRecord(after_path, stream));
    if (ss.total_paths > 0) {
        const int64_t want = (ss.total_paths + 255) / 256;
        const int gblocks = (int)std::min<int64_t>(want, (int64_t)num_cus * 16);
        if (cv_mode == SPAC_CV_TABLE || cv_mode == SPAC_CV_NETWORK)
            launch_exit_values<CV_TABLE>(coef, dim3(gblocks), stream, a, ss.total_paths);
        else launch_exit_values<CV_ANALYTIC>(coef, dim3(gblocks), stream, a, ss.total_paths);
        SPAC_HIP_CHECK(hipGetLastError());
    }
    if (grid_blocks) *grid_blocks = blocks;
    return SPAC_OK;
}

/* Estimator reduction (no synchronisation). */
spac_status Problem::enqueue_reduce(const SolveSet& ss, spac_point_stats* d_stats, int node_mode,
                                    const RecordSet* recs) {
    const RecordSet rs = recs ? *recs : own_records();
    ReduceArgs r;
    memset(&r, 0, sizeof r);
    r.n_solves = ss.n;
    r.node_mode = node_mode;
    r.paths = ss.paths;
    r.paths_uniform = ss.paths_uniform;
    r.rec_off = ss.rec_off;
    r.Y = rs.Y; r.Z = rs.Z; r.xi = rs.xi;
    r.tau = rs.tau; r.gval = rs.gval;
    r.steps = rs.steps; r.outcome = rs.outcome;
    r.stats = d_stats;
    r.cv_fallback = (cv_mode == SPAC_CV_TABLE || cv_mode == SPAC_CV_NETWORK) ? cv_fallback : 0;
    r.cv_max_ratio = cv_max_ratio;
    hipLaunchKernelGGL(fk_reduce_kernel, dim3(ss.n), dim3(64), 0, stream, r);
    SPAC_HIP_CHECK(hipGetLastError());
    return SPAC_OK;
}

/* Launch the persistent path kernel and, optionally, the reduction; synchronous, timed. */
spac_status Problem::run(const SolveSet& ss, uint64_t seed, spac_point_stats* d_stats, int node_mode,
                         const RecordSet* recs) {
    int blocks = 0;
    spac_status rc = launch_run(ss, seed, d_stats, node_mode, recs, &blocks);
    if (rc) return rc;
    return finish_run(blocks);
}

spac_status Problem::launch_run(const SolveSet& ss, uint64_t seed, spac_point_stats* d_stats, int node_mode,
                            const RecordSet* recs, int* blocks) {
    spac_status rc;
    if (!recs) {
        if ((rc = ensure_records(ss.total_paths))) return rc;
        if ((rc = queue.ensure(64))) return rc;
    }
    SPAC_HIP_CHECK(hipEventRecord(ev[0], stream));
    if ((rc = enqueue_paths(ss, seed, nullptr, blocks, ev[1], recs))) return rc;
    if (d_stats && (rc = enqueue_reduce(ss, d_stats, node_mode, recs))) return rc;
    SPAC_HIP_CHECK(hipEventRecord(ev[2], stream));
    return SPAC_OK;
}

spac_status Problem::finish_run(int blocks) {
    SPAC_HIP_CHECK(hipStreamSynchronize(stream));
    float ms_path = 0.f, ms_red = 0.f;
    SPAC_HIP_CHECK(hipEventElapsedTime(&ms_path, ev[0], ev[1]));
    SPAC_HIP_CHECK(hipEventElapsedTime(&ms_red, ev[1], ev[2]));
    last.path_kernel_ms = ms_path;
    last.reduce_kernel_ms = ms_red;
    last.total_ms = (double)ms_path + (double)ms_red;
    last.path_kernel_launches = 1;
    last.grid_blocks = blocks;
    last.path_steps = 0;
    return SPAC_OK;
}

/* The (du/dx, du/dy) channels of a T x T float4 control-variate table, by cell (the layout of
 * CvTable32): cell (i, j) gets the corners (i, j), (i + 1, j), (i, j + 1), (i + 1, j + 1). */
__global__ void k_cv_gradient_cells(int t, const float4* __restrict__ tab, float4* __restrict__ cells) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = (t - 1) * (t - 1);
    if (k >= n) return;
    const int i = k % (t - 1), j = k / (t - 1);
    const float4 a = tab[j * t + i], b = tab[j * t + i + 1], c = tab[(j + 1) * t + i], d = tab[(j + 1) * t + i + 1];
    cells[2 * k] = make_float4(a.y, a.z, b.y, b.z);
    cells[2 * k + 1] = make_float4(c.y, c.z, d.y, d.z);
}

spac_status problem_set_cv(Problem* p, const spac_cv_config* cfg, const void* src, bool src_on_device) {
    if (spac_status rc = no_pending(p)) return rc;
    if (cfg->mode == SPAC_CV_NETWORK)
        return set_error(SPAC_E_INVALID_ARG, "SPAC_CV_NETWORK is installed by spac_mlp_attach_network");
    if (cfg->mode != SPAC_CV_ANALYTIC && cfg->mode != SPAC_CV_NONE && cfg->mode != SPAC_CV_TABLE)
        return set_error(SPAC_E_INVALID_ARG, "unknown control-variate mode %d", cfg->mode);
    if (cfg->fallback && !(cfg->max_ratio > 0.0))
        return set_error(SPAC_E_INVALID_ARG, "max_ratio must be positive");
    if (cfg->mode == SPAC_CV_TABLE) {
        if (!src) return set_error(SPAC_E_INVALID_ARG, "table mode needs a table");
        if (cfg->table_size < 2 || cfg->table_size > 8192)
            return set_error(SPAC_E_INVALID_ARG, "table_size must be in [2, 8192]");
        const double* d = cfg->table_domain;
        if (!(d[2] > d[0]) || !(d[3] > d[1]))
            return set_error(SPAC_E_INVALID_ARG, "table domain must have xmax > xmin and ymax > ymin");
        if (p->desc.coef_kind != SPAC_COEF_FRANKE_POISSON && p->desc.coef_kind != SPAC_COEF_VARCOEF_FRANKE &&
            p->desc.coef_kind != SPAC_COEF_FRANKE_ROBIN)
            return set_error(SPAC_E_INVALID_ARG, "table mode needs a scalar-diffusion coefficient set");
        size_t bytes = (size_t)cfg->table_size * (size_t)cfg->table_size * 16;
        spac_status rc;
        if ((rc = p->cv_table.ensure(bytes))) return rc;
        SPAC_HIP_CHECK(hipMemcpyAsync(p->cv_table.ptr, src, bytes,
                                      src_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, p->stream));
        /* the FP32 kernel reads only the two gradient channels, by cell: one line per lookup */
        const int nc = (cfg->table_size - 1) * (cfg->table_size - 1);
        if ((rc = p->cv_grad.ensure((size_t)nc * 32))) return rc;
        hipLaunchKernelGGL(k_cv_gradient_cells, dim3((nc + 255) / 256), dim3(256), 0, p->stream, cfg->table_size,
                           (const float4*)p->cv_table.ptr, (float4*)p->cv_grad.ptr);
        SPAC_HIP_CHECK(hipGetLastError());
        SPAC_HIP_CHECK(hipStreamSynchronize(p->stream));
        p->cv_t = cfg->table_size;
        for (int i = 0; i < 4; ++i) p->cv_dom[i] = d[i];
    }
    p->cv_mode = cfg->mode;
    p->cv_fallback = cfg->fallback ? 1 : 0;
    p->cv_max_ratio = cfg->max_ratio > 0.0 ? cfg->max_ratio : 1.0;
    return SPAC_OK;
}

/* SPAC_CV_NETWORK: the network's table as in SPAC_CV_TABLE (the exit scores and the fallback
 * read it), plus a device copy of the network's parameters and its input map (lox, loy, sx, sy)
 * for the per-step evaluation in the FP32 kernel. */
spac_status problem_set_cv_network(Problem* p, const spac_cv_config* cfg, const void* table_dev,
                                   const float* params_dev, const float* map) {
    if (spac_status rc = no_pending(p)) return rc;
    if (p->desc.precision != SPAC_FP32)
        return set_error(SPAC_E_INVALID_ARG, "SPAC_CV_NETWORK runs in the FP32 path kernel: the problem must be SPAC_FP32");
    if (p->desc.coef_kind != SPAC_COEF_VARCOEF_FRANKE || p->tab_mask || p->desc.reflecting)
        return set_error(SPAC_E_INVALID_ARG, "SPAC_CV_NETWORK needs the compiled variable-coefficient set, absorbing "
                                             "sides and no coefficient tables");
    spac_cv_config t = *cfg;
    t.mode = SPAC_CV_TABLE;
    spac_status rc = problem_set_cv(p, &t, table_dev, true);
    if (rc) return rc;
    if ((rc = p->cv_net.ensure((size_t)kNetParams * 4))) return rc;
    SPAC_HIP_CHECK(hipMemcpyAsync(p->cv_net.ptr, params_dev, (size_t)kNetParams * 4, hipMemcpyDeviceToDevice, p->stream));
    SPAC_HIP_CHECK(hipStreamSynchronize(p->stream));
    for (int i = 0; i < 4; ++i) p->cv_net_map[i] = map[i];
    p->cv_mode = SPAC_CV_NETWORK;
    return SPAC_OK;
}

/* ------------------------------------------------ static scratch block --
 * Device memory of spac_fk_solve_points_device_noalloc: the per-path records,
 * the per-solve sigma(X0) slots and the work-queue counter of one point solve
 * of at most SPAC_NOALLOC_MAX_PATHS paths over SPAC_NOALLOC_MAX_SOLVES points.
 * Every GPU the library runs on holds its own copy (at its own address) from
 * the code object's load until the process ends.  Host code never passes the
 * array's name to a kernel: it passes the address hipGetSymbolAddress returns
 * for the current device (noalloc_scratch_address). */
constexpr size_t kNaPaths = (size_t)SPAC_NOALLOC_MAX_PATHS, kNaSolves = (size_t)SPAC_NOALLOC_MAX_SOLVES;
constexpr size_t na_align(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t kNaXt = 0, kNaYt = kNaXt + 8 * kNaPaths, kNaY = kNaYt + 8 * kNaPaths, kNaZ = kNaY + 8 * kNaPaths,
                 kNaXi = kNaZ + 8 * kNaPaths, kNaTau = kNaXi + 8 * kNaPaths, kNaGval = kNaTau + 8 * kNaPaths,
                 kNaSteps = kNaGval + 8 * kNaPaths, kNaOutcome = kNaSteps + 4 * kNaPaths,
                 kNaSig = na_align(kNaOutcome + kNaPaths), kNaSigf = kNaSig + 8 * kNaSolves,
                 kNaQueue = na_align(kNaSigf + 4 * kNaSolves), kNaBytes = kNaQueue + 256;
static_assert(kNaBytes == (size_t)SPAC_NOALLOC_SCRATCH_BYTES, "scratch layout and the header's size disagree");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

/* This device's address of `scratch`, looked up once per device (the device must be
 * current); the status of the lookup is kept with it. */
static spac_status noalloc_scratch_address(int device, unsigned char** out) {
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static unsigned char* addr[kMaxDevices] = {};
    static hipError_t err[kMaxDevices] = {};
    static bool done[kMaxDevices] = {};
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices)
        return set_error(SPAC_E_INVALID_ARG, "device %d outside the scratch table (%d devices)", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(mu);
    if (!done[device]) {
        void* ptr = nullptr;
        err[device] = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        addr[device] = err[device] == hipSuccess ? (unsigned char*)ptr : nullptr;
        done[device] = true;
    }
    if (err[device] != hipSuccess || !addr[device])
        return set_error(SPAC_E_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                         hipGetErrorString(err[device]));
    *out = addr[device];
    return SPAC_OK;
}

/* completes a pending spac_fk_assemble_system_async (fk_rows.hip) */
spac_status finish_assemble_system(Problem* p, const Problem::Pending& pd);

} // namespace spac

using namespace spac;

/* ================================================================ C ABI == */
extern "C" {


int32_t spac_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

spac_status spac_fk_problem_create(const spac_problem_desc* desc, spac_problem** out) {
    if (!desc || !out) return set_error(SPAC_E_INVALID_ARG, "null argument");
    *out = nullptr;
    spac_status rc = check_problem(desc);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
        return set_error(SPAC_E_NO_DEVICE, "no HIP device visible");
    Problem* p = new Problem();
    p->desc = *desc;
    SPAC_HIP_CHECK(hipGetDevice(&p->device));
    hipDeviceProp_t prop;
    SPAC_HIP_CHECK(hipGetDeviceProperties(&prop, p->device));
    p->num_cus = prop.multiProcessorCount;
    SPAC_HIP_CHECK(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
    for (int i = 0; i < 4; ++i) SPAC_HIP_CHECK(hipEventCreate(&p->ev[i]));
    p->launch.block_size = 256;
    p->launch.blocks_per_cu = 8;
    p->launch.chunk_paths = 64; /* paths per work item: smaller items balance the tail of heavy nodes */
    p->launch.item_order = SPAC_ORDER_AUTO;
    /* set-up of the no-allocation entry point on this (current) device */
    p->scratch_status = noalloc_scratch_address(p->device, &p->scratch);
    *out = reinterpret_cast<spac_problem*>(p);
    return SPAC_OK;
}

spac_status spac_fk_problem_destroy(spac_problem* hp) {
    if (!hp) return SPAC_OK;
    Problem* p = reinterpret_cast<Problem*>(hp);
    (void)hipStreamSynchronize(p->stream);
    p->release_all();
    for (int i = 0; i < 4; ++i) (void)hipEventDestroy(p->ev[i]);
    (void)hipStreamDestroy(p->stream);
    delete p;
    return SPAC_OK;
}

spac_status spac_fk_set_launch(spac_problem* hp, const spac_launch_config* cfg) {
    if (!hp || !cfg) return set_error(SPAC_E_INVALID_ARG, "null argument");
    Problem* p = reinterpret_cast<Problem*>(hp);
    if (spac_status pr = no_pending(p)) return pr;
    if (cfg->block_size) {
        if (cfg->block_size % 64 || cfg->block_size > 256 || cfg->block_size < 64)
            return set_error(SPAC_E_INVALID_ARG, "block_size must be 64, 128, 192 or 256");
        p->launch.block_size = cfg->block_size;
    }
    if (cfg->blocks_per_cu) {
        if (cfg->blocks_per_cu < 1 || cfg->blocks_per_cu > 64)
            return set_error(SPAC_E_INVALID_ARG, "blocks_per_cu must be in [1, 64]");
        p->launch.blocks_per_cu = cfg->blocks_per_cu;
    }
    if (cfg->chunk_paths) {
        if (cfg->chunk_paths < 1) return set_error(SPAC_E_INVALID_ARG, "chunk_paths must be >= 1");
        p->launch.chunk_paths = cfg->chunk_paths;
    }
    if (cfg->item_order < SPAC_ORDER_NODE || cfg->item_order > SPAC_ORDER_AUTO)
        return set_error(SPAC_E_INVALID_ARG, "item_order must be one of SPAC_ORDER_NODE, _COST, _TAIL, _AUTO");
    p->launch.item_order = cfg->item_order;
    return SPAC_OK;
}

spac_status spac_fk_set_control_variate(spac_problem* hp, const spac_cv_config* cfg) {
    if (!hp || !cfg) return set_error(SPAC_E_INVALID_ARG, "null argument");
    return problem_set_cv(reinterpret_cast<Problem*>(hp), cfg, cfg->table, false);
}

spac_status spac_fk_set_coef_tables(spac_problem* hp, const spac_coef_tables* t) {
    if (!hp || !t) return set_error(SPAC_E_INVALID_ARG, "null argument");
    Problem* p = reinterpret_cast<Problem*>(hp);
    spac_status rc = no_pending(p);
    if (rc) return rc;
    const int32_t all = SPAC_TAB_SIGMA | SPAC_TAB_DRIFT | SPAC_TAB_KILLING | SPAC_TAB_SOURCE | SPAC_TAB_DIRICHLET |
                        SPAC_TAB_SOLUTION;
    if (t->mask & ~all) return set_error(SPAC_E_INVALID_ARG, "table mask 0x%x has bits outside SPAC_TAB_*", t->mask);
    if (t->mask == 0) {
        p->tab_mask = 0;
        return SPAC_OK;
    }
    if (p->desc.coef_kind != SPAC_COEF_VARCOEF_FRANKE)
        return set_error(SPAC_E_INVALID_ARG, "tabulated coefficients need the base set SPAC_COEF_VARCOEF_FRANKE");
    if (p->desc.reflecting) return set_error(SPAC_E_INVALID_ARG, "tabulated coefficients need absorbing sides");
    if (t->table_size < 2 || t->table_size > 4096) return set_error(SPAC_E_INVALID_ARG, "table_size must be in [2, 4096]");
    const double* d = t->table_domain;
    if (!(d[2] > d[0]) || !(d[3] > d[1]))
        return set_error(SPAC_E_INVALID_ARG, "table domain must have xmax > xmin and ymax > ymin");
    if (!t->values) return set_error(SPAC_E_INVALID_ARG, "no table values");
    const size_t n = (size_t)t->table_size * (size_t)t->table_size * SPAC_TAB_STRIDE;
    for (size_t k = 0; k < n; ++k)
        if (!std::isfinite(t->values[k])) return set_error(SPAC_E_NAN, "table value %zu is not finite", k);
    std::vector<float> vf(n);
    for (size_t k = 0; k < n; ++k) vf[k] = (float)t->values[k];
    if ((rc = p->tab.ensure(n * 8)) || (rc = p->tab32.ensure(n * 4))) return rc;
    SPAC_HIP_CHECK(hipMemcpyAsync(p->tab.ptr, t->values, n * 8, hipMemcpyHostToDevice, p->stream));
    SPAC_HIP_CHECK(hipMemcpyAsync(p->tab32.ptr, vf.data(), n * 4, hipMemcpyHostToDevice, p->stream));
    SPAC_HIP_CHECK(hipStreamSynchronize(p->stream));
    p->tab_mask = t->mask;
    p->tab_t = t->table_size;
    for (int i = 0; i < 4; ++i) p->tab_dom[i] = d[i];
    return SPAC_OK;
}

spac_status spac_fk_last_timing(spac_problem* hp, spac_timing* out) {
    if (!hp || !out) return set_error(SPAC_E_INVALID_ARG, "null argument");
    *out = reinterpret_cast<Problem*>(hp)->last;
    return SPAC_OK;
}

/* The uniform point solve's work items as a list (solve, p0, p1), so that the tail-first order
 * can sort them (the list is what the queue would hand out without one). */
__global__ void k_point_items(int32_t n, int32_t cps, uint32_t chunk, uint32_t paths, int4* __restrict__ items) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)n * cps) return;
    const int32_t k = (int32_t)(i / cps), c = (int32_t)(i % cps);
    const uint32_t p0 = (uint32_t)c * chunk;
    items[i] = make_int4(k, (int)p0, (int)min(p0 + chunk, paths), 0);
}

static spac_status solve_points_impl(Problem* p, const double* d_x0, const double* d_y0, int32_t n, uint32_t paths,
                                     double h, uint64_t seed, spac_point_stats* d_stats, bool async = false) {
    SolveSet ss;
    ss.n = n;
    ss.x0 = d_x0;
    ss.y0 = d_y0;
    ss.h_uniform = h;
    ss.sqrth_uniform = std::sqrt(h);
    ss.paths_uniform = paths;
    ss.total_paths = (int64_t)n * (int64_t)paths;
    ss.chunk = work_item_paths((uint32_t)p->launch.chunk_paths, ss.total_paths);
    ss.chunks_per_solve = (int32_t)((paths + ss.chunk - 1) / ss.chunk);
    int64_t items = (int64_t)ss.chunks_per_solve * n;
    if (items > kMaxWorkItems) return set_error(SPAC_E_INVALID_ARG, "too many work items (%lld)", (long long)items);
    ss.n_items = (int32_t)items;
    spac_status rc;
    /* with a variable sigma the solves' exit-time tails differ: list the items so that the
       tail-first order can put the slow ones first (a constant sigma and one h give every item
       the same key, and the implicit order stays) */
    const int coef = p->coef_kernels();
    if ((coef == SPAC_COEF_VARCOEF_FRANKE || coef == kCoefTabulated) && items > 1 &&
        (p->launch.item_order == SPAC_ORDER_TAIL || p->launch.item_order == SPAC_ORDER_AUTO)) {
        if ((rc = p->sitems.ensure((size_t)items * sizeof(int4)))) return rc;
        hipLaunchKernelGGL(k_point_items, dim3((unsigned)((items + 255) / 256)), dim3(256), 0, p->stream, n,
                           ss.chunks_per_solve, ss.chunk, paths, (int4*)p->sitems.ptr);
        SPAC_HIP_CHECK(hipGetLastError());
        ss.items = (const int4*)p->sitems.ptr;
    }
    if ((rc = p->ensure_solve_scratch(n, &ss))) return rc;
    if (async) {
        int blocks = 0;
        if ((rc = p->launch_run(ss, seed, d_stats, 0, nullptr, &blocks))) return rc;
        p->pending.kind = Problem::PENDING_POINTS;
        p->pending.n = n;
        p->pending.blocks = blocks;
        return SPAC_OK;
    }
    rc = p->run(ss, seed, d_stats, 0);
    if (rc) return rc;
    p->last.path_steps = -1;
    return SPAC_OK;
}

static spac_status check_points_args(const void* x, const void* y, int32_t n, uint32_t paths, double h, const void* st) {
    if (!x || !y || !st) return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (n < 1) return set_error(SPAC_E_INVALID_ARG, "n must be >= 1");
    if (paths < 1) return set_error(SPAC_E_INVALID_ARG, "paths must be >= 1");
    if (!(h > 0.0) || !(h < 1e30)) return set_error(SPAC_E_INVALID_ARG, "h must be positive and finite");
    return SPAC_OK;
}

spac_status spac_fk_solve_points_device(spac_problem* hp, const double* d_x0, const double* d_y0, int32_t n,
                                        uint32_t paths, double h, uint64_t seed, spac_point_stats* d_stats) {
    if (!hp) return set_error(SPAC_E_INVALID_ARG, "null problem");
    spac_status rc = check_points_args(d_x0, d_y0, n, paths, h, d_stats);
    if (rc) return rc;
    Problem* p = reinterpret_cast<Problem*>(hp);
    if ((rc = no_pending(p))) return rc;
    if (p->desc.precision != SPAC_FP64 && p->desc.precision != SPAC_FP32)
        return set_error(SPAC_E_INVALID_ARG, "bad precision");
    return solve_points_impl(p, d_x0, d_y0, n, paths, h, seed, d_stats);
}

spac_status spac_fk_solve_points_device_async(spac_problem* hp, const double* d_x0, const double* d_y0, int32_t n,
                                              uint32_t paths, double h, uint64_t seed, spac_point_stats* d_stats) {
    if (!hp) return set_error(SPAC_E_INVALID_ARG, "null problem");
    spac_status rc = check_points_args(d_x0, d_y0, n, paths, h, d_stats);
    if (rc) return rc;
    Problem* p = reinterpret_cast<Problem*>(hp);
    if ((rc = no_pending(p))) return rc;
    if (p->desc.precision != SPAC_FP64 && p->desc.precision != SPAC_FP32)
        return set_error(SPAC_E_INVALID_ARG, "bad precision");
    return solve_points_impl(p, d_x0, d_y0, n, paths, h, seed, d_stats, true);
}

spac_status spac_fk_query(spac_problem* hp, int32_t* done) {
    if (!hp || !done) return set_error(SPAC_E_INVALID_ARG, "null argument");
    Problem* p = reinterpret_cast<Problem*>(hp);
    if (p->pending.kind == Problem::PENDING_NONE) {
        *done = 1;
        return SPAC_OK;
    }
    hipError_t e = hipStreamQuery(p->stream);
    if (e == hipSuccess) *done = 1;
    else if (e == hipErrorNotReady) *done = 0;
    else return set_error(SPAC_E_HIP, "hipStreamQuery failed: %s", hipGetErrorString(e));
    return SPAC_OK;
}

spac_status spac_fk_wait(spac_problem* hp) {
    if (!hp) return set_error(SPAC_E_INVALID_ARG, "null problem");
    Problem* p = reinterpret_cast<Problem*>(hp);
    const Problem::Pending pd = p->pending;
    p->pending = Problem::Pending();
    if (pd.kind == Problem::PENDING_NONE) return SPAC_OK;
    if (pd.kind == Problem::PENDING_SYSTEM) return finish_assemble_system(p, pd);
    spac_status rc = p->finish_run(pd.blocks);
    if (rc) return rc;
    p->last.path_steps = -1;
    return SPAC_OK;
}

spac_status spac_fk_solve_points_device_noalloc(spac_problem* hp, const double* d_x0, const double* d_y0, int32_t n,
                                                uint32_t paths, double h, uint64_t seed, spac_point_stats* d_stats) {
    if (!hp) return set_error(SPAC_E_INVALID_ARG, "null problem");
    spac_status rc = check_points_args(d_x0, d_y0, n, paths, h, d_stats);
    if (rc) return rc;
    Problem* p = reinterpret_cast<Problem*>(hp);
    if ((rc = no_pending(p))) return rc;
    if (p->desc.precision != SPAC_FP64 && p->desc.precision != SPAC_FP32)
        return set_error(SPAC_E_INVALID_ARG, "bad precision");
    if ((size_t)n > kNaSolves)
        return set_error(SPAC_E_INVALID_ARG, "no-allocation solve: %d points exceed the scratch maximum of %d", n,
                         (int)SPAC_NOALLOC_MAX_SOLVES);
    const uint64_t total = (uint64_t)n * (uint64_t)paths;
    if (total > kNaPaths)
        return set_error(SPAC_E_INVALID_ARG, "no-allocation solve: %llu paths exceed the scratch maximum of %d",
                         (unsigned long long)total, (int)SPAC_NOALLOC_MAX_PATHS);
    if (!p->scratch) return set_error(p->scratch_status ? p->scratch_status : SPAC_E_HIP,
                                      "no scratch address for device %d", p->device);
    int cur = -1;
    SPAC_HIP_CHECK(hipGetDevice(&cur));
    if (cur != p->device)
        return set_error(SPAC_E_INVALID_ARG, "device %d is current, the problem lives on device %d", cur, p->device);
    unsigned char* b = p->scratch;
    RecordSet rs;
    rs.xt = (double*)(b + kNaXt); rs.yt = (double*)(b + kNaYt);
    rs.Y = (double*)(b + kNaY); rs.Z = (double*)(b + kNaZ); rs.xi = (double*)(b + kNaXi);
    rs.tau = (double*)(b + kNaTau); rs.gval = (double*)(b + kNaGval);
    rs.steps = (int32_t*)(b + kNaSteps); rs.outcome = (int8_t*)(b + kNaOutcome);
    rs.queue = (int32_t*)(b + kNaQueue); /* zeroed by a kernel before the path kernel */
    SolveSet ss;
    ss.n = n;
    ss.x0 = d_x0;
    ss.y0 = d_y0;
    ss.h_uniform = h;
    ss.sqrth_uniform = std::sqrt(h);
    ss.paths_uniform = paths;
    ss.total_paths = (int64_t)total;
    ss.chunk = work_item_paths((uint32_t)p->launch.chunk_paths, ss.total_paths);
    ss.chunks_per_solve = (int32_t)((paths + ss.chunk - 1) / ss.chunk);
    ss.n_items = ss.chunks_per_solve * n; /* <= total <= 2^19 */
    ss.sig0 = (double*)(b + kNaSig);
    ss.sig0f = (float*)(b + kNaSigf);
    if ((rc = p->run(ss, seed, d_stats, 0, &rs))) return rc;
    p->last.path_steps = -1;
    return SPAC_OK;
}

spac_status spac_fk_solve_points(spac_problem* hp, const double* x0, const double* y0, int32_t n, uint32_t paths,
                                 double h, uint64_t seed, spac_point_stats* stats) {
    if (!hp) return set_error(SPAC_E_INVALID_ARG, "null problem");
    if (spac_status pr = no_pending(reinterpret_cast<Problem*>(hp))) return pr;
    spac_status rc = check_points_args(x0, y0, n, paths, h, stats);
    if (rc) return rc;
    Problem* p = reinterpret_cast<Problem*>(hp);
    if ((rc = p->sx.ensure((size_t)n * 8))) return rc;
    if ((rc = p->sy.ensure((size_t)n * 8))) return rc;
    if ((rc = p->stats.ensure((size_t)n * sizeof(spac_point_stats)))) return rc;
    SPAC_HIP_CHECK(hipMemcpyAsync(p->sx.ptr, x0, (size_t)n * 8, hipMemcpyHostToDevice, p->stream));
    SPAC_HIP_CHECK(hipMemcpyAsync(p->sy.ptr, y0, (size_t)n * 8, hipMemcpyHostToDevice, p->stream));
    rc = solve_points_impl(p, (const double*)p->sx.ptr, (const double*)p->sy.ptr, n, paths, h, seed,
                           (spac_point_stats*)p->stats.ptr);
    if (rc) return rc;
    if ((rc = p->copy_out(stats, p->stats.ptr, (size_t)n * sizeof(spac_point_stats)))) return rc;
    spac_status worst = SPAC_OK;
    for (int i = 0; i < n; ++i)
        if (stats[i].status < worst) worst = stats[i].status;
    if (worst == SPAC_E_NAN) return set_error(SPAC_E_NAN, "NaN reached an estimator");
    if (worst == SPAC_E_MAX_STEPS) return set_error(SPAC_E_MAX_STEPS, "paths stopped by max_steps");
    return SPAC_OK;
}

spac_status spac_fk_simulate(spac_problem* hp, const spac_solve_batch* b, uint64_t seed, spac_path_records* rec,
                             spac_point_stats* stats) {
    if (!hp || !b) return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (spac_status pr = no_pending(reinterpret_cast<Problem*>(hp))) return pr;
    if (b->n < 1 || !b->x0 || !b->y0 || !b->paths || !b->h || !b->stream)
        return set_error(SPAC_E_INVALID_ARG, "batch needs n >= 1 and x0, y0, paths, h, stream");
    Problem* p = reinterpret_cast<Problem*>(hp);
    const int32_t n = b->n;
    for (int k = 0; k < n; ++k)
        if (b->paths[k] < 1) return set_error(SPAC_E_INVALID_ARG, "solve %d has no paths", k);
    SolveSet ss;
    spac_status rc = p->upload_solves(n, b->x0, b->y0, b->paths, b->h, b->stream, b->circle, &ss);
    if (rc) return rc;
    size_t nb = (size_t)n;
    if ((rc = p->stats.ensure(nb * sizeof(spac_point_stats)))) return rc;
    rc = p->run(ss, seed, stats ? (spac_point_stats*)p->stats.ptr : nullptr, b->circle ? 1 : 0);
    if (rc) return rc;
    size_t tp = (size_t)ss.total_paths;
    if (rec) {
        if (rec->x_tau && (rc = p->copy_out(rec->x_tau, p->rec_xt.ptr, tp * 8))) return rc;
        if (rec->y_tau && (rc = p->copy_out(rec->y_tau, p->rec_yt.ptr, tp * 8))) return rc;
        if (rec->Y && (rc = p->copy_out(rec->Y, p->rec_Y.ptr, tp * 8))) return rc;
        if (rec->Z && (rc = p->copy_out(rec->Z, p->rec_Z.ptr, tp * 8))) return rc;
        if (rec->xi && (rc = p->copy_out(rec->xi, p->rec_xi.ptr, tp * 8))) return rc;
        if (rec->tau && (rc = p->copy_out(rec->tau, p->rec_tau.ptr, tp * 8))) return rc;
        if (rec->gval && (rc = p->copy_out(rec->gval, p->rec_gval.ptr, tp * 8))) return rc;
        if (rec->steps && (rc = p->copy_out(rec->steps, p->rec_steps.ptr, tp * 4))) return rc;
        if (rec->outcome && (rc = p->copy_out(rec->outcome, p->rec_outcome.ptr, tp))) return rc;
    }
    if (stats && (rc = p->copy_out(stats, p->stats.ptr, nb * sizeof(spac_point_stats)))) return rc;
    return SPAC_OK;
}

} /* extern "C" */
