/*
 * spac_fk.h — C ABI of the MI355X Feynman–Kac path solver and interface-row
 * assembly.  Plain C types only (no STL, no Eigen, no torch): host pointers,
 * sizes and SoA arrays.  The library owns device buffers and one HIP stream per
 * handle; the caller owns every host buffer passed in or out.
 *
 * Reference interfaces replaced (SpAc, paths relative to the reference root):
 *   FeynmanKacSolver::Solve_OMP_Analytic(Vector2d X0, unsigned N, double h,
 *       double rho, BVP)                        Solvers/FeynmanKacSolver.hpp:95-96
 *   FeynmanKacSolver::Solve_OMP_Analytic(Nudo&, unsigned N, double h,
 *       double rho, BVP, Interfaz)              Solvers/FeynmanKacSolver.hpp:97-98
 *   public estimator fields phi, phiphi, phi_VR, phiphi_VR, xi, xixi, xiphi,
 *       APL, RNGC                               Solvers/FeynmanKacSolver.hpp:80-83
 *   per-path exit data X_tau, Y_tau, Z_tau, tau, xi_tau, outcome_tau
 *                                               Solvers/FeynmanKacSolver.hpp:46-49
 *   Interfaz::Update_G (G row weights)          Mesh/Interfaz.cpp:219-264
 *
 * Errors: every call returns spac_status (0 = OK, negative = error) and never
 * prints or exits; spac_last_error() gives a thread-local message.
 * Threading: calls on one handle are serialised by the caller; different
 * handles may be used from different threads.
 */
#ifndef SPAC_FK_H
#define SPAC_FK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SPAC_ABI_VERSION 3

typedef int32_t spac_status;
enum {
    SPAC_OK = 0,
    SPAC_E_INVALID_ARG = -1,
    SPAC_E_UNREGISTERED_PLUGIN = -2, /* coefficient/boundary with no device twin */
    SPAC_E_MAX_STEPS = -3,           /* at least one path hit max_steps */
    SPAC_E_NAN = -4,                 /* a NaN reached an estimator */
    SPAC_E_HIP = -5,                 /* HIP runtime error */
    SPAC_E_PATTERN = -6,             /* G column outside the CSR pattern */
    SPAC_E_BAD_INTERFACE = -7,       /* unknown interface code (Mesh/Interfaz.cpp:248-250) */
    SPAC_E_NO_DEVICE = -8,
    SPAC_E_OUT_OF_MEMORY = -9
};

/* Coefficient sets with a device implementation.  Each is the fused device
 * twin of a set of host plugin functions (u, g, c, f, psi, varphi, b, gradu,
 * F, mu, sigma) registered by the C++ plugin surface. */
enum {
    SPAC_COEF_FRANKE_POISSON = 1, /* BVPs/Franke_Poisson.hpp: sigma = 1.41421356237 I, b = c = 0 */
    SPAC_COEF_VARCOEF_FRANKE = 2, /* sigma = (1 + 0.5 sin(pi x) cos(pi y)) I, b = 0.3 (y, -x),
                                     c = -1 - x^2, u = Franke, f = -(L u) */
    SPAC_COEF_FRANKE_ROBIN = 3    /* Franke Poisson plus the Robin data of reflecting sides:
                                     varphi = -0.5, psi(X, N) = grad u(X).N + 0.5 u(X), so that
                                     du/dN - varphi u = psi holds for the Franke u */
};

/* Boundary distance plugins (Dominios/Rectangulo.hpp:7-229, Dominios/Circulo.hpp:7-22). */
enum {
    SPAC_BOUNDARY_NONE = 0,      /* no subdomain: distance -inf, never the nearest */
    SPAC_BOUNDARY_RECTANGLE = 1, /* params: xmin, ymin, xmax, ymax */
    SPAC_BOUNDARY_CIRCLE = 2     /* params: R, cx, cy */
};

/* Reflecting sides of the rectangle domain (spac_problem_desc.reflecting), the
 * device form of Frontera::Absorbente returning false on a side
 * (BVP/Frontera.hpp:23,34).  A corner reflects when both of its sides do. */
enum {
    SPAC_REFLECT_XMIN = 1,
    SPAC_REFLECT_YMIN = 2,
    SPAC_REFLECT_XMAX = 4,
    SPAC_REFLECT_YMAX = 8
};

enum { SPAC_FP64 = 0, SPAC_FP32 = 1 };

/* Path outcome (Solvers/FeynmanKacSolver.hpp:41) plus the step cap. */
enum {
    SPAC_OUT_INSIDE = 0,
    SPAC_OUT_DOMAIN = 1,    /* parada_dominio */
    SPAC_OUT_SUBDOMAIN = 2, /* parada_subdominio */
    SPAC_OUT_REFLECTED = 3,
    SPAC_OUT_MAX_STEPS = 4
};

/* Interpolation of the G row over a circle's nodes (Mesh/Interfaz.cpp:206-264). */
enum {
    SPAC_RBF_IMQ_GLOBAL = 0,    /* reference: IMQ over all arc nodes, iPsi = Psi^-1 */
    SPAC_RBF_GAUSS_STENCIL = 1  /* Gaussian, local stencils of stencil_size nodes */
};

typedef struct {
    int32_t coef_kind;
    int32_t domain_kind;
    double domain_params[4];
    int32_t subdomain_kind;     /* used by the point solve; node solves use their circle */
    int32_t precision;          /* SPAC_FP64 or SPAC_FP32 stepping */
    double subdomain_params[4];
    double gobet_constant;      /* -0.5826 in the reference (Solvers/FeynmanKacSolver.cpp:268,283) */
    int32_t max_steps;          /* cap on steps per path; the reference has none */
    int32_t reflecting;         /* SPAC_REFLECT_* mask; 0 (every side absorbing) is the reference's
                                   shipped configuration.  Nonzero needs SPAC_FP64. */
    double reflect_rho;         /* rho of Solve_OMP_Analytic (Solvers/FeynmanKacSolver.hpp:95-98):
                                   paths closer than rho to a reflecting side take the
                                   Lepingle step (Solvers/FeynmanKacSolver.cpp:19-67) */
} spac_problem_desc;

typedef struct spac_problem spac_problem;

/* A batch of independent solves: solve k starts paths[k] paths at (x0[k], y0[k])
 * with step h[k]; RNG stream id stream[k]; stopped by the problem's domain and
 * either the problem's subdomain (circle == NULL) or circle (R, cx, cy) at
 * circle[3k..3k+2]. */
typedef struct {
    int32_t n;
    int32_t reserved;
    const double* x0;
    const double* y0;
    const uint32_t* paths;
    const double* h;
    const uint32_t* stream;
    const double* circle;
} spac_solve_batch;

/* Per-path exit data, SoA over sum(paths) entries in solve order
 * (Solvers/FeynmanKacSolver.cpp:285-291).  Any pointer may be NULL. */
typedef struct {
    double* x_tau;   /* projected exit point (Npro) */
    double* y_tau;
    double* Y;
    double* Z;
    double* xi;
    double* tau;
    double* gval;    /* g(X_tau) for domain exits, u(X_tau) for subdomain exits */
    int32_t* steps;
    int8_t* outcome;
} spac_path_records;

/* Node estimator, the reference's public fields after Update()
 * (Solvers/FeynmanKacSolver.cpp:294-323).  xi/xixi are the true moments (the
 * reference's sums[8] aliasing is not reproduced); RNGC counts normals drawn. */
typedef struct {
    double phi, phiphi, phi_VR, phiphi_VR, xi, xixi, xiphi, APL, RNGC;
    double B;             /* node mode only */
    int64_t paths;
    int64_t path_steps;   /* executed Paso_Absorbente + Dentro pairs */
    int64_t n_max_steps;  /* paths stopped by the cap */
    int32_t status;
    int32_t n_domain_exits;
    int32_t cv_fallback;  /* 1: Var(score + xi) > max_ratio Var(score), xi left out of B */
    int32_t reserved;
} spac_point_stats;

/* Interfaces (circles) and their nodes, the device copy of Interfaz
 * (Mesh/Interfaz.hpp:8-30).  Node lists are CSR: interface k owns nodes
 * node_offset[k] .. node_offset[k+1]-1. */
typedef struct {
    int32_t n_interfaces;
    int32_t rbf_kind;           /* SPAC_RBF_* */
    const double* cx;
    const double* cy;
    const double* radius;
    const int32_t* code;        /* interior code: -11,-22,-33,-44,-1,-2,-3,-4 or 1 */
    const double* c2;           /* IMQ shape per interface */
    const int32_t* node_offset; /* [n_interfaces + 1] */
    const double* node_theta;   /* angular position of each circle node */
    const int32_t* node_global; /* global index = matrix column */
    const int64_t* ipsi_offset; /* [n_interfaces]: start of the M*M row-major iPsi block */
    const double* ipsi;         /* IMQ mode: iPsi blocks; stencil mode: stencil inverses */
    double gauss_eps;           /* stencil mode: Gaussian shape on stencil-normalised distance */
    int32_t stencil_size;       /* stencil mode: nodes per stencil (odd) */
    int32_t reserved;
} spac_interface_table;

/* Rows to assemble: node k lies inside interface iface[k] and needs its G/B
 * row (Solvers/FeynmanKacSolver.cpp:344-355).  frontera[k] != 0 gives B = g(x)
 * with no paths (:346-347). */
typedef struct {
    int32_t n;
    int32_t reserved;
    const double* x;
    const double* y;
    const int32_t* iface;
    const int32_t* frontera;
    const uint32_t* paths;
    const double* h;
    const uint32_t* stream;
} spac_node_batch;

/* Output rows: B[k] and G weights of node k over its interface's nodes at
 * G[G_offset[k] .. G_offset[k+1]), columns in G_col (caller sizes
 * G_offset[n] = sum over rows of the interface's node count). */
typedef struct {
    double* B;
    double* G;
    int32_t* G_col;
    int64_t* G_offset;
    spac_point_stats* stats;    /* optional */
} spac_row_block;

/* ---------------------------------------------------------------- API -- */
const char* spac_last_error(void);
int32_t spac_abi_version(void);
/* Number of visible devices (0 on a machine without a GPU; never fails). */
int32_t spac_device_count(void);

spac_status spac_fk_problem_create(const spac_problem_desc* desc, spac_problem** out);
spac_status spac_fk_problem_destroy(spac_problem* p);

/* Point solve for n starting points, paths each: Solve_OMP_Analytic(X0, N, h,
 * rho, BVP) (Solvers/FeynmanKacSolver.cpp:335-342) batched over points.
 * stats[k] receives point k's estimator. */
spac_status spac_fk_solve_points(spac_problem* p, const double* x0, const double* y0, int32_t n,
                                 uint32_t paths, double h, uint64_t seed, spac_point_stats* stats);

/* Same on device-resident inputs and outputs (d_x0, d_y0 [n], d_stats [n]);
 * the timed form for throughput measurements.  Synchronous. */
spac_status spac_fk_solve_points_device(spac_problem* p, const double* d_x0, const double* d_y0,
                                        int32_t n, uint32_t paths, double h, uint64_t seed,
                                        spac_point_stats* d_stats);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): spac_fk_solve_points_device on the same inputs and outputs
 * with the same results, but every buffer beyond them (per-path records,
 * per-point sigma(X0), the work-queue counter) lives in one static device block
 * of SPAC_NOALLOC_SCRATCH_BYTES, which each GPU the library runs on holds until
 * the process ends.  Nothing is allocated or freed by the call.  At most
 * SPAC_NOALLOC_MAX_SOLVES points and n * paths <= SPAC_NOALLOC_MAX_PATHS, else
 * SPAC_E_INVALID_ARG.  The problem's device must be current.  Every call on one
 * GPU shares that GPU's block (it keeps the last call's records): calls that
 * overlap on one GPU, from any handle or thread, race.  Synchronous. */
#define SPAC_NOALLOC_MAX_PATHS (1 << 19)
#define SPAC_NOALLOC_MAX_SOLVES (1 << 15)
#define SPAC_NOALLOC_SCRATCH_BYTES (61 * SPAC_NOALLOC_MAX_PATHS + 12 * SPAC_NOALLOC_MAX_SOLVES + 256)
spac_status spac_fk_solve_points_device_noalloc(spac_problem* p, const double* d_x0, const double* d_y0,
                                                int32_t n, uint32_t paths, double h, uint64_t seed,
                                                spac_point_stats* d_stats);

/* General batch of solves with per-path exit data (Simulacion_OMP,
 * Solvers/FeynmanKacSolver.cpp:137-197) and, optionally, per-solve estimators. */
spac_status spac_fk_simulate(spac_problem* p, const spac_solve_batch* batch, uint64_t seed,
                             spac_path_records* records, spac_point_stats* stats);

/* Interface rows: Solve_OMP_Analytic(Nudo&, ...) + Reduce_Analytic(..., Nudo&,
 * Interfaz) + Update(Nudo&) (Solvers/FeynmanKacSolver.cpp:233-292, 324-355)
 * for a batch of nodes in one launch. */
spac_status spac_fk_assemble_rows(spac_problem* p, const spac_interface_table* ifaces,
                                  const spac_node_batch* nodes, uint64_t seed, spac_row_block* out);

/* A mesh's interface system resident on the device (created by spac_mesh_system_create,
 * spac_mesh.h). */
typedef struct spac_system spac_system;

/* spac_fk_assemble_rows for every job of the system's mesh (nodes: the mesh's jobs in job
 * order), accumulated into the device-resident system instead of copied out: the rows path of
 * Reduce_Analytic + SistemaLineal::Actualiza_GB (Solvers/FeynmanKacSolver.cpp:233-292,
 * Mesh/SistemaLineal.cpp:68-88) without the host round trip.  stats (optional): the nodes->n
 * estimator records.  The timing of the call (spac_fk_timing) includes the accumulation. */
spac_status spac_fk_assemble_system(spac_problem* p, const spac_interface_table* ifaces,
                                    const spac_node_batch* nodes, uint64_t seed, spac_system* sys,
                                    spac_point_stats* stats);

/* spac_fk_assemble_system for the batch of every job of the system's mesh at one path count
 * and step: `paths` paths of step h for every non-boundary job, RNG stream = job index (the
 * node batch spac_mesh_jobs describes).  This is the reference's use of the solver, one N and
 * one h for every node of the mesh (FeynmanKacSolver::Solve_OMP_Analytic(Nudo&, ..., N, h),
 * Solvers/FeynmanKacSolver.cpp:233-260, as driven by main_test_mpi.cpp over every node).
 * The solve table and work items are built on the device from the mesh's job arrays, which
 * the system holds since spac_mesh_system_create, so nothing goes up from the host per call.
 * Results equal spac_fk_assemble_system with that node batch bit for bit.  The _async form
 * returns once the kernels are enqueued, as spac_fk_assemble_system_async (spac_fk_wait). */
spac_status spac_fk_assemble_system_uniform(spac_problem* p, spac_system* sys, uint32_t paths, double h,
                                            uint64_t seed, spac_point_stats* stats);
spac_status spac_fk_assemble_system_uniform_async(spac_problem* p, spac_system* sys, uint32_t paths, double h,
                                                  uint64_t seed, spac_point_stats* stats);

/* Rows of nodes inside full interior circles (interface code 1) from the pseudospectral
 * solver instead of paths: PseudoespectralCirculoSolver::Inicializa + Resuelve(BVP,
 * nudos_interior, nudos_periferia) (Solvers/PseudoespectralCirculoSolver.cpp:2-260,
 * BVP/Interpolador_PS.cpp:18-61), applied by the reference's drivers to every non-perimeter
 * circle (main_test_serial.cpp:67-79).  Deterministic; host code.  coef_kind: a registered
 * coefficient set whose sigma is sqrt(2) I and b is 0 on the circles (the operator is the
 * Laplacian plus c), e.g. SPAC_COEF_FRANKE_POISSON.  Each circle needs an even node count;
 * n_radio (odd) Chebyshev points span its diameter (the reference's N_radio, 43).  The rows
 * have the layout of spac_fk_assemble_rows (G over the interface's nodes, columns their global
 * indices); boundary nodes get B = g(x); stats, when given, are zero with status 0. */
typedef struct {
    int32_t circles;         /* interfaces solved */
    int32_t rows;            /* interior-node rows */
    int32_t factorizations;  /* operators factored (circles with equal grids and c share one) */
    int32_t solves;          /* right-hand sides solved */
    double seconds;          /* wall time of the call */
} spac_ps_report;
spac_status spac_ps_assemble_rows(int32_t coef_kind, const spac_interface_table* ifaces, const spac_node_batch* nodes,
                                  int32_t n_radio, spac_row_block* out, spac_ps_report* report);

/* Persistence and resume of row assemblies.  The reference persists only the finished
 * system and its solution (Escribe_GB_COO / Escribe_GB_CSR, Mesh/SistemaLineal.cpp:96-138;
 * solution_GMRES.txt, Solvers/SistemaLinealSolverGMRES_OMP.hpp:606); these keep the rows of
 * every node batch, so that a batched run restarts at the batch it stopped in. */
typedef struct {
    int32_t batches;   /* node batches in the call */
    int32_t computed;  /* assembled on the device (and written) */
    int32_t loaded;    /* read from their files */
    int32_t rejected;  /* files present but for other inputs or damaged: recomputed */
} spac_resume_report;

/* spac_fk_assemble_rows over `nodes` in batches of batch_nodes nodes.  Batch b's rows are
 * read from dir/rows_<b>.bin when that file holds the rows of exactly these inputs (the
 * interface table, the batch's nodes, seed; checked by hash) and is whole (payload hash);
 * otherwise the batch is assembled and its file written (to a .tmp name, then renamed).
 * Results equal spac_fk_assemble_rows on the whole batch bit for bit (rows depend only on
 * seed, node stream id and path id).  out->stats, when given, receives every node's
 * estimators. */
spac_status spac_fk_assemble_rows_resumable(spac_problem* p, const spac_interface_table* ifaces,
                                            const spac_node_batch* nodes, uint64_t seed, int32_t batch_nodes,
                                            const char* dir, spac_row_block* out, spac_resume_report* report);
/* One batch's row file (host only): rows->G_offset[0..n] index rows->G / rows->G_col, and
 * rows->stats (required) holds the n estimators; batch, seed and input_hash identify it. */
spac_status spac_rows_file_write(const char* path, int64_t batch, uint64_t seed, uint64_t input_hash, int32_t n,
                                 const spac_row_block* rows);
/* Read it back into rows (G from rows->G_offset[0] on, stats optional) when the file exists
 * and matches (batch, seed, input_hash, n, entries) and its payload hash: *found = 1; a
 * missing file gives 0, a mismatched or damaged one -1 (rows untouched in both cases). */
spac_status spac_rows_file_read(const char* path, int64_t batch, uint64_t seed, uint64_t input_hash, int32_t n,
                                int64_t entries, spac_row_block* rows, int32_t* found);

/* Asynchronous forms (the reference's calls block inside their OpenMP regions,
 * Solvers/FeynmanKacSolver.cpp:137-197): the same work as spac_fk_solve_points_device
 * and spac_fk_assemble_system, enqueued on the handle's stream; the call returns once the
 * inputs are checked and uploaded and the kernels are launched.  spac_fk_wait blocks until the
 * call in flight is done and returns its status (the estimator records of
 * spac_fk_assemble_system_async reach `stats` then; spac_fk_last_timing is valid after it);
 * spac_fk_query tells whether it is done without blocking.  One call may be in flight per
 * handle: every other call on the handle returns SPAC_E_INVALID_ARG until spac_fk_wait (or
 * spac_fk_problem_destroy, which waits).  Device inputs, outputs and the system must stay
 * valid and untouched until then. */
spac_status spac_fk_solve_points_device_async(spac_problem* p, const double* d_x0, const double* d_y0, int32_t n,
                                              uint32_t paths, double h, uint64_t seed, spac_point_stats* d_stats);
spac_status spac_fk_assemble_system_async(spac_problem* p, const spac_interface_table* ifaces,
                                          const spac_node_batch* nodes, uint64_t seed, spac_system* sys,
                                          spac_point_stats* stats);
spac_status spac_fk_wait(spac_problem* p);
/* *done = 1 when nothing is in flight on the handle or the call in flight has finished */
spac_status spac_fk_query(spac_problem* p, int32_t* done);

/* Timing of the last call's kernels, measured with HIP events on the handle's
 * stream: path kernel, reduction kernels (ms), and launches counted. */
typedef struct {
    double path_kernel_ms;
    double reduce_kernel_ms;
    double total_ms;
    int64_t path_steps;
    int32_t path_kernel_launches;
    int32_t grid_blocks;
} spac_timing;
spac_status spac_fk_last_timing(spac_problem* p, spac_timing* out);

/* Work-item order of the persistent kernel's queue. */
enum {
    SPAC_ORDER_NODE = 0, /* items in node order */
    SPAC_ORDER_COST = 1, /* items of the nodes with the smallest h (most steps per path) first */
    SPAC_ORDER_TAIL = 2, /* items of the solves whose exit times have the longest tail first: the
                            key R^2 / (sigma(X0)^2 h), R the solve's circle radius (1 without a
                            circle), bucketed by eighth octaves on the device */
    SPAC_ORDER_AUTO = 3  /* a new problem's default: SPAC_ORDER_TAIL for launches of at most 256
                            paths per lane (where the longest paths set the end of the launch),
                            SPAC_ORDER_NODE above */
};

/* Launch shape and refill schedule of the persistent path kernel (0 = library
 * default; results never depend on it). */
typedef struct {
    int32_t block_size;       /* threads per block (multiple of 64) */
    int32_t blocks_per_cu;    /* persistent blocks per compute unit */
    int32_t chunk_paths;      /* paths per work-queue item */
    int32_t item_order;       /* SPAC_ORDER_* (0 is SPAC_ORDER_NODE; a new problem starts with SPAC_ORDER_AUTO) */
} spac_launch_config;
spac_status spac_fk_set_launch(spac_problem* p, const spac_launch_config* cfg);

/* A row-assembly plan for repeated assemblies of one node batch (config-3
 * style batches: thousands of nodes with uneven paths and h).  Creation
 * uploads the batch, the interface table and the work list once into buffers
 * owned by the plan, and captures the whole device sequence of
 * spac_fk_assemble_rows (queue reset, persistent path kernel, estimator
 * reduction, B and G kernels) in one HIP graph; execution writes the seed to
 * device memory, replays the graph and copies the rows out.  The plan copies
 * the problem's launch configuration and control variate at creation.
 * Results equal spac_fk_assemble_rows with the same seed. */
typedef struct spac_rows_plan spac_rows_plan;
spac_status spac_fk_plan_create(spac_problem* p, const spac_interface_table* ifaces, const spac_node_batch* nodes,
                                spac_rows_plan** out);
/* device_ms (optional): time of the graph on the plan's stream (HIP events). */
spac_status spac_fk_plan_execute(spac_rows_plan* plan, uint64_t seed, spac_row_block* out, double* device_ms);
spac_status spac_fk_plan_destroy(spac_rows_plan* plan);

/* Autotuner of the persistent path kernel for a node batch: coordinate search
 * (two passes) over block_size {64, 128, 256}, blocks_per_cu {1, 2, 4, 8, 16},
 * chunk_paths {32 .. 1024} and item_order, each candidate timed (best of two
 * runs, HIP events) on the batch with its paths scaled by path_fraction.  The
 * fastest configuration is installed on p (spac_fk_set_launch).  tried /
 * tried_ms (optional, cap entries) receive every measured candidate. */
typedef struct {
    int32_t n_candidates;
    int32_t reserved;
    spac_launch_config chosen;
    double best_ms;     /* path kernel on the sample, chosen configuration */
    double default_ms;  /* path kernel on the sample, configuration at entry */
} spac_autotune_report;
spac_status spac_fk_autotune_rows(spac_problem* p, const spac_interface_table* ifaces, const spac_node_batch* nodes,
                                  double path_fraction, spac_launch_config* tried, double* tried_ms, int32_t cap,
                                  spac_autotune_report* report);

/* Tabulated coefficients: the device form of the reference's numeric coefficients
 * (Coeficiente_Escalar_Numerico / _Vectorial_ / _Matricial_, BVP/Coeficiente_Escalar_Numerico.hpp:6-19,
 * evaluated through an Interpolador over a TablaDeValores, BVP/Interpolador.hpp:7-19).  The
 * table replaces the coefficients its mask names; the others stay the compiled plugin's. */
enum {
    SPAC_TAB_SIGMA = 1,     /* sigma = s(x) I */
    SPAC_TAB_DRIFT = 2,     /* b */
    SPAC_TAB_KILLING = 4,   /* c */
    SPAC_TAB_SOURCE = 8,    /* f */
    SPAC_TAB_DIRICHLET = 16, /* g (domain exits, boundary nodes) */
    SPAC_TAB_SOLUTION = 32  /* u (the statistics' subdomain-exit score) */
};
typedef struct {
    int32_t mask;            /* SPAC_TAB_* bits; 0 removes the table */
    int32_t table_size;      /* T >= 2 */
    double table_domain[4];  /* xmin, ymin, xmax, ymax */
    const double* values;    /* host, T * T * 8 doubles: at node (i, j), (j T + i) * 8 + {0: s, 1: b_x,
                                2: b_y, 3: c, 4: f, 5: g, 6: u, 7: unused}; copied */
} spac_coef_tables;
/* Install (or remove) tabulated coefficients on p.  Needs the base set SPAC_COEF_VARCOEF_FRANKE
 * (any scalar-diffusion operator is a table of it) and absorbing sides; the FP64 kernel reads
 * the table in FP64 (bilinear, the bits of the CPU restatement), the FP32 kernel a float copy. */
spac_status spac_fk_set_coef_tables(spac_problem* p, const spac_coef_tables* tables);

/* Control variate F in xi += Y F(X).dW (Solvers/FeynmanKacSolver.cpp:13). */
enum {
    SPAC_CV_ANALYTIC = 0, /* the coefficient set's F plugin (reference behaviour; default) */
    SPAC_CV_NONE = 1,     /* F = 0: plain Monte Carlo, xi = 0 */
    SPAC_CV_TABLE = 2,    /* F = -sigma(X) grad u_hat(X), grad u_hat bilinear in a T x T table */
    SPAC_CV_NETWORK = 3   /* F = -sigma(X) grad u_hat(X), grad u_hat the learned network's gradient
                             heads evaluated at X on every step (FP32 stepping, variable-coefficient
                             set; installed by spac_mlp_attach_network only).  The per-step form of
                             the learned control variate, to measure what the table changes. */
};

/* A tabulated u_hat (the device form of a TablaDeValores + Interpolador,
 * BVP/Interpolador.hpp:15-19): table[(j T + i) 4 + c] holds, at
 * x = xmin + i (xmax - xmin) / (T - 1) and y = ymin + j (ymax - ymin) / (T - 1),
 * c = 0: u_hat, 1: d u_hat / dx, 2: d u_hat / dy, 3: unused.  Points outside
 * the table take the value at the nearest edge.  In table mode u_hat(X_tau)
 * replaces u(X_tau) in the subdomain-exit score of the statistics. */
typedef struct {
    int32_t mode;            /* SPAC_CV_* */
    int32_t fallback;        /* 1: per solve, drop xi where Var(score + xi) > max_ratio Var(score) */
    double max_ratio;        /* 1.0 = fall back wherever the control variate adds variance */
    int32_t table_size;      /* T >= 2 (table mode) */
    int32_t reserved;
    double table_domain[4];  /* xmin, ymin, xmax, ymax */
    const float* table;      /* host, T * T * 4 floats; copied */
} spac_cv_config;
spac_status spac_fk_set_control_variate(spac_problem* p, const spac_cv_config* cfg);

#ifdef __cplusplus
}
#endif
#endif /* SPAC_FK_H */
