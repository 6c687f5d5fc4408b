/*
 * spac_mesh.h — C ABI of the PDDSparse mesh, interface tables and the
 * assembled interface system G u = B.
 *
 * Reference interfaces replaced:
 *   Malla::Construye_Cuadrado(params, Nl, superposicion, theta_0, Nt, BVP, c2)
 *                                               Mesh/Malla.cpp:12-83
 *   Interfaz::Inicia_Cuadrado / Calcula_Psi     Mesh/Interfaz.cpp:2-217
 *   SistemaLineal::Actualiza_GB(Nudo&)          Mesh/SistemaLineal.cpp:68-88
 *   SistemaLineal::Escribe_GB_COO / Escribe_GB_CSR / Lee_GB_COO
 *                                               Mesh/SistemaLineal.cpp:96-164
 *   Malla::Escribe_Posiciones                   Mesh/Malla.cpp:99-112
 *   SistemaLinealSolverGMRES::Resuelve          Solvers/SistemaLinealSolverGMRES_OMP.hpp:159-616
 * The mesh lives in host memory; arrays returned by pointer stay valid until
 * spac_mesh_destroy.
 */
#ifndef SPAC_MESH_H
#define SPAC_MESH_H

#include <stdint.h>

#include "spac_fk.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double domain[4];           /* xmin, ymin, xmax, ymax */
    int32_t n_side;             /* Nl: circles per side */
    int32_t nodes_per_circle;   /* Nt */
    double overlap;             /* superposicion, in [1, 2] */
    double theta0;              /* angulo_inicial */
    double c2_corner;           /* IMQ shape of corner arcs ("esquina") */
    double c2_side;             /* IMQ shape of side arcs ("lado") */
    int32_t rbf_kind;           /* SPAC_RBF_IMQ_GLOBAL or SPAC_RBF_GAUSS_STENCIL */
    int32_t stencil_size;       /* Gaussian stencils: nodes per stencil */
    double gauss_eps;           /* Gaussian stencils: shape on stencil-normalised distance */
    int32_t fk_all_circles;     /* 1: rows of full interior circles by FK too; 0: perimeter only */
    int32_t reserved;
} spac_mesh_desc;

typedef struct spac_mesh spac_mesh;

typedef struct {
    int32_t n_nodes;            /* rows of G */
    int32_t n_interfaces;
    int32_t n_jobs;             /* FK row jobs (node inside a circle) */
    int32_t nnz;                /* CSR pattern size */
    int64_t total_G;            /* sum over jobs of the job interface's node count */
    double radius;
    double max_cond;            /* largest |iPsi|_F |Psi|_F over the interpolation blocks */
} spac_mesh_info;

spac_status spac_mesh_build_square(const spac_mesh_desc* desc, spac_mesh** out);
spac_status spac_mesh_destroy(spac_mesh* m);
spac_status spac_mesh_get_info(const spac_mesh* m, spac_mesh_info* out);
/* Interface table for spac_fk_assemble_rows (pointers owned by the mesh). */
spac_status spac_mesh_interfaces(const spac_mesh* m, spac_interface_table* out);
/* Per job: global row, position, interface, frontera flag; arrays of n_jobs. */
spac_status spac_mesh_jobs(const spac_mesh* m, int32_t* row, double* x, double* y, int32_t* iface, int32_t* frontera);
/* Per node (global index order): position and frontera flag; arrays of n_nodes. */
spac_status spac_mesh_nodes(const spac_mesh* m, double* x, double* y, int32_t* frontera);
/* Calcula_Psi's return value |iPsi|_F |Psi|_F for interface k (IMQ tables). */
spac_status spac_mesh_interface_cond(const spac_mesh* m, int32_t k, double* cond);
/* Reset B and G to the initial pattern values (diagonal = containing-circle count, off-diagonal 1e-23). */
spac_status spac_mesh_reset_system(spac_mesh* m);
/* Actualiza_GB for every job of rows (as produced by spac_fk_assemble_rows). */
spac_status spac_mesh_accumulate(spac_mesh* m, const spac_row_block* rows);
/* CSR export: G_i[n+1], G_j[nnz], G_ij[nnz], B[n] (any pointer may be NULL). */
spac_status spac_mesh_csr(const spac_mesh* m, int32_t* G_i, int32_t* G_j, double* G_ij, double* B);
/* Device-resident interface system: the mesh's CSR pattern, its initial values (as after
 * spac_mesh_reset_system) and each job's first slot in its row, on the device.  No counterpart
 * in the reference, whose rows travel through host memory into SistemaLineal::Actualiza_GB
 * (Mesh/SistemaLineal.cpp:68-88) before GMRES reads the host CSR.  spac_fk_assemble_system
 * accumulates rows into it with Actualiza_GB's arithmetic (each G slot: its initial value plus
 * the one job contribution it receives; B: the row's jobs summed in job order), and
 * spac_gmres_solve_system solves it in place.  Creation fails with SPAC_E_PATTERN for a mesh
 * whose jobs would add twice into one CSR slot (interface node sets are disjoint, so a square
 * mesh never does). */
spac_status spac_mesh_system_create(const spac_mesh* m, spac_system** out);
spac_status spac_mesh_system_destroy(spac_system* s);
/* Two-level preconditioner for spac_gmres_solve_system (no counterpart in the reference, whose
 * GMRES uses RAS through MUMPS, Solvers/SistemaLinealSolverGMRES_OMP.hpp:519-577): the rows are
 * grouped by their node's cell on an nx x ny grid over the nodes' bounding box (empty cells
 * dropped, at most 576 aggregates), and right preconditioning uses
 * M^-1 = D^-1 + Z (Z^T A Z)^-1 Z^T with Z the aggregates' indicator vectors (additive coarse
 * correction).  Z^T A Z is formed on the device per solve (fixed-point sums, so results do not
 * depend on scheduling) and inverted on the host.  nx = ny = 0 switches it off.  The converged
 * solution is that of the same system to the same tolerance; only the iterations differ. */
spac_status spac_mesh_system_set_coarse(const spac_mesh* m, spac_system* s, int32_t nx, int32_t ny);
/* Copy the system out: G_i[n+1], G_j[nnz], G_ij[nnz], B[n] (any pointer may be NULL). */
spac_status spac_mesh_system_read(const spac_system* s, int32_t* G_i, int32_t* G_j, double* G_ij, double* B);
spac_status spac_mesh_write_coo(const spac_mesh* m, const char* G_path, const char* B_path);
spac_status spac_mesh_write_csr(const spac_mesh* m, const char* dir);
/* Escribe_Posiciones with u = the registered coefficient set's exact solution. */
spac_status spac_mesh_write_positions(const spac_mesh* m, int32_t coef_kind, const char* path);

typedef struct {
    int32_t iterations;
    int32_t restarts;
    double residual;            /* ||B - G u|| / ||B|| of the returned u, always a true residual */
    int32_t converged;          /* 1 only when that true residual is below tol */
    int32_t reserved;
    double seconds;             /* wall time of the call */
    double device_ms;           /* GPU solver: kernel time on its stream (HIP events); host solver: 0 */
} spac_gmres_info;

/* Restarted FP64 GMRES (modified Gram–Schmidt, Givens rotations) with a
 * Jacobi preconditioner on the CSR system; u holds the initial guess on entry. */
spac_status spac_gmres_solve(int32_t n, const int32_t* G_i, const int32_t* G_j, const double* G_ij, const double* B,
                             double* u, double tol, int32_t max_iter, int32_t restart, spac_gmres_info* info);

/* The same GMRES(m) with right Jacobi preconditioning on the GPU
 * (SistemaLinealSolverGMRES::Resuelve, Solvers/SistemaLinealSolverGMRES_OMP.hpp:519-577,
 * without its RAS preconditioner): FP64 CSR SpMV (a few lanes per row),
 * classical Gram–Schmidt for the Arnoldi basis (one pass, two from the first
 * step with heavy cancellation), every reduction in a fixed order; the
 * (m+1) x m Hessenberg least-squares problem and its Givens rotations on the
 * host, a few steps behind the device.  Host arrays in and out; u holds the
 * initial guess on entry. */
spac_status spac_gmres_solve_gpu(int32_t n, const int32_t* G_i, const int32_t* G_j, const double* G_ij,
                                 const double* B, double* u, double tol, int32_t max_iter, int32_t restart,
                                 spac_gmres_info* info);
/* spac_gmres_solve_gpu on a device-resident system (no upload); u (host, n_rows) holds the
 * initial guess on entry and the solution on return.  The system's device must be current.
 * Calls that use one system (assembly, solve, read) must not overlap. */
spac_status spac_gmres_solve_system(spac_system* s, double* u, double tol, int32_t max_iter, int32_t restart,
                                    spac_gmres_info* info);
/* The solution file of the reference's GMRES driver: one value per line, fixed notation with
 * 12 decimals (Output/solution_GMRES.txt, Solvers/SistemaLinealSolverGMRES_OMP.hpp:606-611). */
spac_status spac_write_solution(const char* path, int32_t n, const double* u);
/* Per-phase wall times as "phase,rank,seconds" lines, the format of the reference's
 * Output/times.csv (main_test_mpi.cpp:62,73,114-187); append != 0 adds to an existing file. */
spac_status spac_write_times_csv(const char* path, int32_t n, const char* const* phases, const int32_t* ranks,
                                 const double* seconds, int32_t append);
/* Read a COO file ("nnz n" then "i j v") into CSR (Lee_GB_COO). Call with
 * G_i == NULL to get n and nnz first. */
spac_status spac_read_coo(const char* path, int32_t* n, int32_t* nnz, int32_t* G_i, int32_t* G_j, double* G_ij);

#ifdef __cplusplus
}
#endif
#endif
