/*
 * fk_internal.h — library-private state behind the opaque spac_problem handle.
 */
#ifndef SPAC_FK_INTERNAL_H
#define SPAC_FK_INTERNAL_H

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <vector>

#include "../../include/spac_fk.h"

#include "spac_error.h"

namespace spac {

/* Grow-only device allocation owned by a handle. */
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t cap = 0;
    spac_status ensure(size_t bytes);
    void release();
};

/* Device-side description of a batch of solves.  Arrays are device pointers;
 * a null per-solve array means "uniform", taken from the scalar beside it. */
struct SolveSet {
    int32_t n = 0;
    const double* x0 = nullptr;
    const double* y0 = nullptr;
    const double* h = nullptr;
    double h_uniform = 0.0;
    const double* sqrth = nullptr; /* per solve sqrt(h), or nullptr: sqrth_uniform */
    double sqrth_uniform = 0.0;
    double* sig0 = nullptr;        /* per solve sigma(X0), FP64 and FP32 (filled on the device) */
    float* sig0f = nullptr;
    const uint32_t* stream = nullptr;
    const uint32_t* paths = nullptr;
    uint32_t paths_uniform = 0;
    const int64_t* rec_off = nullptr;
    const double* circle = nullptr;
    const int4* items = nullptr;
    int4* items_tail = nullptr;    /* the items in tail-first order (SPAC_ORDER_TAIL), or nullptr */
    int32_t* tail_hist = nullptr;  /* kTailBlocks x kTailBuckets counts, then offsets */
    int32_t n_items = 0;
    int32_t chunks_per_solve = 0;
    uint32_t chunk = 0;            /* paths per work item (see work_item_paths) */
    int64_t total_paths = 0;
};

/* Device arrays one path launch writes: the per-path records (SoA over the
 * batch's paths) and the work-queue counter.  A Problem's own buffers, or the
 * library's static scratch block (spac_fk_solve_points_device_noalloc). */
struct RecordSet {
    double *xt = nullptr, *yt = nullptr, *Y = nullptr, *Z = nullptr, *xi = nullptr, *tau = nullptr, *gval = nullptr;
    int32_t* steps = nullptr;
    int8_t* outcome = nullptr;
    int32_t* queue = nullptr;
};

struct DeviceSystem;

/* the kernel set of a problem with tabulated coefficients (CoefTab over the variable-coefficient
 * set), as opposed to the SPAC_COEF_* compiled sets */
constexpr int kCoefTabulated = 1000;

struct Problem {
    spac_problem_desc desc{};
    int device = 0;
    int num_cus = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    spac_launch_config launch{};
    spac_timing last{};
    /* per-path records (SoA) */
    DeviceBuffer rec_xt, rec_yt, rec_Y, rec_Z, rec_xi, rec_tau, rec_gval, rec_steps, rec_outcome;
    /* solve table, work queue, outputs */
    DeviceBuffer sx, sy, sh, sstream, spaths, soff, scircle, sitems, queue, stats;
    DeviceBuffer ssq, ssig, ssigf; /* per solve sqrt(h) and sigma(X0) */
    DeviceBuffer sitems_tail, stail_hist; /* the tail-first item order */
    /* interface tables and row outputs */
    DeviceBuffer ifc, ifnode, ifipsi, ifoff, rows_iface, rows_goff, rows_G, rows_part, rows_col;
    /* control variate (SPAC_CV_*) and its table (T x T x float4) */
    DeviceBuffer cv_table;
    DeviceBuffer cv_net;  /* SPAC_CV_NETWORK: the network's parameters (FP32) */
    float cv_net_map[4] = {0.0f, 0.0f, 1.0f, 1.0f}; /* its input map: lox, loy, sx, sy */
    DeviceBuffer cv_grad; /* float4 [2 (T - 1)^2]: the table's (du/dx, du/dy) by cell, read by the FP32 path kernel */
    int32_t cv_mode = SPAC_CV_ANALYTIC;
    int32_t cv_fallback = 0;
    double cv_max_ratio = 1.0;
    int32_t cv_t = 0;
    double cv_dom[4] = {0.0, 0.0, 1.0, 1.0};

    /* tabulated coefficients (spac_fk_set_coef_tables): T x T x 8 doubles and a float copy */
    DeviceBuffer tab, tab32;
    int32_t tab_mask = 0, tab_t = 0;
    double tab_dom[4] = {0.0, 0.0, 1.0, 1.0};

    DeviceBuffer seed_dev; /* seed read by captured kernels (plans) */
    /* this device's address of the static scratch block (fk_kernels.hip), looked up when the
     * handle is created; nullptr if the lookup failed (scratch_status holds its status) */
    unsigned char* scratch = nullptr;
    spac_status scratch_status = SPAC_OK;

    /* the coefficient kernels to launch: desc.coef_kind, or kCoefTabulated */
    int coef_kernels() const { return tab_mask ? kCoefTabulated : desc.coef_kind; }

    /* the handle's own record buffers (valid after ensure_records) */
    RecordSet own_records() const;

    /* The asynchronous call in flight on the handle's stream (spac_fk_*_async), which
     * spac_fk_wait completes; kind 0: none.  Every other call on the handle refuses to run
     * while one is pending. */
    struct Pending {
        int32_t kind = 0; /* PENDING_* */
        int32_t n = 0;    /* solves or jobs */
        int32_t blocks = 0;
        const DeviceSystem* sys = nullptr;
        const spac_point_stats* d_stats = nullptr;
        spac_point_stats* host_stats = nullptr;
    } pending;
    enum { PENDING_NONE = 0, PENDING_POINTS = 1, PENDING_SYSTEM = 2 };

    spac_status ensure_records(int64_t total_paths);
    /* per-solve sigma(X0) slots for n solves (allocated here, never during capture) */
    spac_status ensure_solve_scratch(int32_t n, SolveSet* ss);
    spac_status upload_solves(int32_t n, const double* x0, const double* y0, const uint32_t* paths, const double* h,
                              const uint32_t* strm, const double* circle, SolveSet* out);
    /* The solve table upload_solves would build for every job of a device system's mesh at one
     * (paths, h) (RNG stream = job index), built on the device from the system's resident job
     * arrays: nothing goes up from the host (fk_rows.hip). */
    spac_status device_solves_uniform(const DeviceSystem* sys, uint32_t paths, double h, SolveSet* out);
    /* Enqueue without synchronising (capturable): per-solve sigma(X0), queue
     * reset, path kernel (after_path recorded behind it when given), then the
     * exit values g(X_tau) / u(X_tau) of every record.  seed_dev != nullptr:
     * kernels read the seed from device memory instead of `seed`. */
    spac_status enqueue_paths(const SolveSet& ss, uint64_t seed, const uint64_t* seed_dev, int* grid_blocks,
                              hipEvent_t after_path = nullptr, const RecordSet* recs = nullptr);
    spac_status enqueue_reduce(const SolveSet& ss, spac_point_stats* d_stats, int node_mode,
                               const RecordSet* recs = nullptr);
    /* recs == nullptr: the handle's own records (grown here); otherwise the given arrays,
     * with nothing allocated */
    spac_status run(const SolveSet& ss, uint64_t seed, spac_point_stats* d_stats, int node_mode,
                    const RecordSet* recs = nullptr);
    /* run() in two halves: launch_run enqueues (events ev[0..2] around the path kernel and the
     * reduction) and returns; finish_run waits for the stream and fills `last` */
    spac_status launch_run(const SolveSet& ss, uint64_t seed, spac_point_stats* d_stats, int node_mode,
                       const RecordSet* recs, int* blocks);
    spac_status finish_run(int blocks);
    /* pinned host staging of upload_solves and copy_out (DMA through pinned memory instead of
     * pageable copies) */
    void* stage = nullptr;
    size_t stage_cap = 0;
    spac_status ensure_stage(size_t bytes);
    /* device -> host copy of `bytes` on the handle's stream, synchronous; large copies go
     * through the pinned staging block */
    spac_status copy_out(void* dst, const void* src, size_t bytes);
    void release_all() {
        if (stage) (void)hipHostFree(stage);
        stage = nullptr;
        stage_cap = 0;
        DeviceBuffer* all[] = {&rec_xt, &rec_yt, &rec_Y, &rec_Z, &rec_xi, &rec_tau, &rec_gval, &rec_steps,
                               &rec_outcome, &sx, &sy, &sh, &sstream, &spaths, &soff, &scircle, &sitems,
                               &ssq, &ssig, &ssigf, &sitems_tail, &stail_hist,
                               &queue, &stats, &ifc, &ifnode, &ifipsi, &ifoff, &rows_iface, &rows_goff,
                               &rows_G, &rows_part, &rows_col, &cv_table, &cv_grad, &cv_net, &seed_dev, &tab, &tab32};
        for (DeviceBuffer* b : all) b->release();
    }
};

/* The tail-first item order: a counting sort of the work items into kTailBuckets buckets of
 * log2(R^2 / (sigma(X0)^2 h)) in eighth octaves, over kTailBlocks slices of the item list. */
constexpr int kTailBuckets = 256;
constexpr int kTailBlocks = 64;
constexpr int64_t kTailAutoPathsPerLane = 256; /* SPAC_ORDER_AUTO sorts launches up to this */

/* Most work items one launch may queue: the queue counter runs past the last item by up to a
 * claim per wave (at most 2^18), and it is an int32. */
constexpr int64_t kMaxWorkItems = 0x7ff00000;

/* Paths per work item of a launch over total_paths paths: the launch's chunk, enlarged for very
 * large batches so that the queue holds at most ~2^20 items (fk_kernels.hip). */
uint32_t work_item_paths(uint32_t chunk, int64_t total_paths);

/* SPAC_OK when no asynchronous call is pending on p, else an error naming spac_fk_wait. */
spac_status no_pending(const Problem* p);

/* Install a control variate; the table (cfg->table ignored) is copied from
 * src, a host pointer or (src_on_device) a device pointer. */
spac_status problem_set_cv(Problem* p, const spac_cv_config* cfg, const void* src, bool src_on_device);
spac_status problem_set_cv_network(Problem* p, const spac_cv_config* cfg, const void* table_dev,
                                   const float* params_dev, const float* map);

/* A mesh's interface system held on the device: the CSR pattern, the values right after
 * SistemaLineal::Inicializa_CSR (diagonal = containing circles, 1e-23 elsewhere), each job's
 * first slot in its row, and its rows' jobs in job order.  Rows assembled on the device are
 * accumulated here and solved here (spac_fk_assemble_system, spac_gmres_solve_system). */
/* largest coarse space of the system GMRES: its Galerkin matrix is inverted on the device per
 * solve, by one workgroup */
constexpr int kMaxCoarse = 576;
/* rows per block of the coarse prolongation, and the most aggregates one block's rows may span
 * (coarse spaces beyond it prolong row by row) */
constexpr int kCoarseRowBlock = 256;
constexpr int kCoarseBlockAggs = 32;

struct DeviceSystem {
    int32_t n_rows = 0, n_jobs = 0;
    int64_t nnz = 0;
    int device = 0;
    DeviceBuffer Ip, J, init, vals, B; /* int32 [n_rows + 1], int32 [nnz], f64 [nnz] x 2, f64 [n_rows] */
    DeviceBuffer job_slot;             /* int64 [n_jobs]: first CSR slot of the job's G, -1: none */
    DeviceBuffer dslot;                /* int32 [n_rows]: each row's diagonal slot, -1: none */
    DeviceBuffer row_job_ptr, row_jobs; /* int32 [n_rows + 1], int32 [n_jobs]: jobs of a row, in order */
    std::vector<int32_t> h_job_row;    /* host copy: row of every job */
    std::vector<int32_t> h_job_iface, h_job_m, h_job_frontera; /* interface, G length, boundary flag */
    std::vector<int32_t> h_Ip, h_J;    /* host copy of the pattern (reads) */
    std::vector<char> h_table_sig; /* table_signature of the mesh's interface table (batch check) */
    /* the mesh's interface table (iface_table_upload layout), its jobs' interface and boundary
     * flag (int32 [n_jobs] each, one buffer) and G offsets (int64 [n_jobs + 1]) */
    DeviceBuffer t_ifc, t_ifoff, t_ifnode, t_ifipsi, t_jobs, t_goff;
    int32_t n_ifaces = 0, rbf_kind = 0, stencil_size = 0; /* of that table */
    double gauss_eps = 0.0;
    int32_t mmax = 1;                       /* its largest interface (nodes) */
    int64_t g_total = 0;                    /* G values of all jobs (t_goff[n_jobs]) */
    /* the solve table of the batch of every job at one (paths, h), which
     * spac_fk_assemble_system_uniform completes on the device: each job's node (x, then y;
     * f64 [n_jobs] each), its circle (f64 [3 n_jobs]: radius, cx, cy) and the number of live
     * (non-boundary) jobs before it (int32 [n_jobs + 1]) */
    DeviceBuffer u_xy, u_circle, u_live;
    int32_t n_live = 0;
    DeviceBuffer node_xy; /* the rows' nodes: x, then y (f64 [n_rows] each) */
    DeviceBuffer flag; /* int32: worst estimator status of the last assembly */
    /* optional coarse space of the GMRES preconditioner (spac_mesh_system_set_coarse):
     * coarse_k aggregates of rows, piecewise constant */
    int32_t coarse_k = 0;
    DeviceBuffer c_agg;          /* int32 [n_rows]: aggregate of each row */
    DeviceBuffer c_ptr, c_rows;  /* int32 [k + 1], [n_rows]: rows of each aggregate, ascending */
    DeviceBuffer c_efix;         /* int64 [k * k]: Z^T A Z in 2^-40 fixed point */
    DeviceBuffer c_einv;         /* f64 [k * k]: its inverse */
    DeviceBuffer c_work;         /* f64 [k * 2k]: [E | I] reduced in place by the device inversion, then int32
                                    flag (singular) */
    DeviceBuffer c_vec;          /* f64 [2 k]: aggregate sums, coarse correction */
    /* the aggregates of each kCoarseRowBlock-row block of rows, in order of first appearance
       (int32 [blocks + 1] offsets, int32 entries), and each row's index in its block's list
       (uint8 [n_rows]); c_blk_max: the longest list (<= kCoarseBlockAggs) */
    DeviceBuffer c_bptr, c_bagg, c_rslot;
    int32_t c_blk_max = 0;
    /* per-block partial sums of a vector over each block's aggregates (f64, one per entry of
       c_bagg), and for each aggregate its entries in block order (int32 [k + 1] offsets, int32
       entries): Z^T x is summed from these */
    DeviceBuffer c_part, c_aptr, c_aent;
    /* GMRES work state of the last solve (buffers, stream, events, pinned columns), kept for the
       next solve of the same size (gmres.hip); gm_free releases it */
    mutable void* gm = nullptr;
    mutable void (*gm_free)(void*) = nullptr;
    void release() {
        for (DeviceBuffer* b : {&Ip, &J, &init, &vals, &B, &job_slot, &dslot, &row_job_ptr, &row_jobs, &t_ifc, &t_ifoff,
                                &t_ifnode, &t_ifipsi, &t_jobs, &t_goff, &u_xy, &u_circle, &u_live, &node_xy, &flag, &c_agg,
                                &c_ptr, &c_rows, &c_efix, &c_einv, &c_work, &c_vec, &c_bptr, &c_bagg, &c_rslot,
                                &c_part, &c_aptr, &c_aent})
            b->release();
        if (gm && gm_free) gm_free(gm);
        gm = nullptr;
    }
};

/* The parts of an interface table a row assembly reads on the host, plus its node lists and
 * kinds, as bytes: two tables with equal signatures give the same rows (fk_rows.hip). */
std::vector<char> table_signature(const spac_interface_table* t);

/* Copy an interface table to the device (fk_rows.hip); returns once the copies are done. */
spac_status iface_table_upload(const spac_interface_table* t, hipStream_t s, DeviceBuffer& ifc, DeviceBuffer& ifoff,
                               DeviceBuffer& ifnode, DeviceBuffer& ifipsi);

} // namespace spac

#endif
