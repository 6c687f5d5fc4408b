/*
 * mesh_capi.cpp — C ABI of include/spac_mesh.h over Malla / Interfaz /
 * SistemaLineal, the text formats and the host FP64 GMRES (host code only; the
 * device-resident system is in system_capi.cpp).
 */
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "../../include/spac_mesh.h"
#include "BVPs.hpp"
#include "Malla.hpp"
#include "mesh_handle.hpp"
#include "spac_error.h"


using namespace spac;

extern "C" {

spac_status spac_mesh_build_square(const spac_mesh_desc* d, spac_mesh** out) {
    if (!d || !out) return set_error(SPAC_E_INVALID_ARG, "null argument");
    *out = nullptr;
    if (d->rbf_kind != SPAC_RBF_IMQ_GLOBAL && d->rbf_kind != SPAC_RBF_GAUSS_STENCIL)
        return set_error(SPAC_E_INVALID_ARG, "unknown rbf_kind %d", d->rbf_kind);
    if (d->rbf_kind == SPAC_RBF_GAUSS_STENCIL && (d->stencil_size < 1 || !(d->gauss_eps > 0.0)))
        return set_error(SPAC_E_INVALID_ARG, "Gaussian stencils need stencil_size >= 1 and gauss_eps > 0");
    if (d->rbf_kind == SPAC_RBF_IMQ_GLOBAL && d->fk_all_circles)
        return set_error(SPAC_E_BAD_INTERFACE,
                         "global IMQ interpolation is defined on arcs only (Mesh/Interfaz.cpp:248-250); "
                         "use Gaussian stencils for full circles");
    MeshHandle* h = new MeshHandle();
    h->desc = *d;
    std::map<std::string, double> c2;
    c2["esquina"] = d->c2_corner;
    c2["lado"] = d->c2_side;
    std::vector<double> P(d->domain, d->domain + 4);
    int rc;
    try {
        rc = h->malla.Construye_Cuadrado(P, d->n_side, d->overlap, d->theta0, d->nodes_per_circle, c2);
    } catch (const std::exception& e) {
        delete h;
        return set_error(SPAC_E_INVALID_ARG, "mesh construction failed: %s", e.what());
    }
    if (rc) {
        delete h;
        return set_error(SPAC_E_INVALID_ARG, "bad mesh parameters (Nl >= 2, Nt >= 4, overlap in [1, 2], xmax > xmin)");
    }
    Malla& m = h->malla;
    h->jobs = m.Trabajos(d->fk_all_circles != 0);
    const int K = (int)m.Interfaces.size();
    h->node_offset.push_back(0);
    for (int k = 0; k < K; ++k) {
        Interfaz& it = m.Interfaces[k];
        h->cx.push_back(it.centro(0));
        h->cy.push_back(it.centro(1));
        h->radius.push_back(it.radio);
        h->code.push_back(it.interior);
        h->c2.push_back(it.c2);
        for (const Nudo& n : it.nudos_circunferencia) {
            h->theta.push_back(n.posicion_angular);
            h->node_global.push_back(n.indice_global);
        }
        h->node_offset.push_back((int32_t)h->theta.size());
        double cond = 0.0;
        try {
            if (d->rbf_kind == SPAC_RBF_GAUSS_STENCIL) {
                cond = it.Calcula_Stencils(d->gauss_eps, d->stencil_size);
                h->ipsi_offset.push_back((int64_t)h->ipsi.size());
                h->ipsi.insert(h->ipsi.end(), it.stencil_inverses.begin(), it.stencil_inverses.end());
            } else if (it.interior < 0) {
                cond = it.iPsi.norm() * it.Psi.norm();
                h->ipsi_offset.push_back((int64_t)h->ipsi.size());
                h->ipsi.insert(h->ipsi.end(), it.iPsi.d.begin(), it.iPsi.d.end());
            } else {
                h->ipsi_offset.push_back(-1);
            }
        } catch (const std::exception& e) {
            delete h;
            return set_error(SPAC_E_INVALID_ARG, "interpolation matrix of interface %d: %s", k, e.what());
        }
        h->cond.push_back(cond);
        if (cond > h->max_cond) h->max_cond = cond;
    }
    *out = reinterpret_cast<spac_mesh*>(h);
    return SPAC_OK;
}

spac_status spac_mesh_destroy(spac_mesh* m) {
    delete reinterpret_cast<MeshHandle*>(m);
    return SPAC_OK;
}

spac_status spac_mesh_get_info(const spac_mesh* hm, spac_mesh_info* out) {
    if (!hm || !out) return set_error(SPAC_E_INVALID_ARG, "null argument");
    const MeshHandle* h = reinterpret_cast<const MeshHandle*>(hm);
    out->n_nodes = h->malla.numero_total_nudos;
    out->n_interfaces = (int32_t)h->malla.Interfaces.size();
    out->n_jobs = (int32_t)h->jobs.size();
    out->nnz = h->malla.sistema.nnz;
    int64_t tg = 0;
    for (const TrabajoFila& j : h->jobs) tg += (int64_t)h->malla.Interfaces[j.interfaz].nudos_circunferencia.size();
    out->total_G = tg;
    out->radius = h->malla.radio;
    out->max_cond = h->max_cond;
    return SPAC_OK;
}

spac_status spac_mesh_interfaces(const spac_mesh* hm, spac_interface_table* t) {
    if (!hm || !t) return set_error(SPAC_E_INVALID_ARG, "null argument");
    const MeshHandle* h = reinterpret_cast<const MeshHandle*>(hm);
    std::memset(t, 0, sizeof *t);
    t->n_interfaces = (int32_t)h->cx.size();
    t->rbf_kind = h->desc.rbf_kind;
    t->cx = h->cx.data();
    t->cy = h->cy.data();
    t->radius = h->radius.data();
    t->code = h->code.data();
    t->c2 = h->c2.data();
    t->node_offset = h->node_offset.data();
    t->node_theta = h->theta.data();
    t->node_global = h->node_global.data();
    t->ipsi_offset = h->ipsi_offset.data();
    t->ipsi = h->ipsi.data();
    t->gauss_eps = h->desc.gauss_eps;
    t->stencil_size = h->desc.stencil_size;
    return SPAC_OK;
}

spac_status spac_mesh_jobs(const spac_mesh* hm, int32_t* row, double* x, double* y, int32_t* iface, int32_t* frontera) {
    if (!hm) return set_error(SPAC_E_INVALID_ARG, "null argument");
    const MeshHandle* h = reinterpret_cast<const MeshHandle*>(hm);
    for (size_t k = 0; k < h->jobs.size(); ++k) {
        const TrabajoFila& j = h->jobs[k];
        if (row) row[k] = j.nudo_global;
        if (x) x[k] = j.posicion(0);
        if (y) y[k] = j.posicion(1);
        if (iface) iface[k] = j.interfaz;
        if (frontera) frontera[k] = j.frontera ? 1 : 0;
    }
    return SPAC_OK;
}

spac_status spac_mesh_nodes(const spac_mesh* hm, double* x, double* y, int32_t* frontera) {
    if (!hm) return set_error(SPAC_E_INVALID_ARG, "null argument");
    const MeshHandle* h = reinterpret_cast<const MeshHandle*>(hm);
    for (const Interfaz& it : h->malla.Interfaces)
        for (const Nudo& n : it.nudos_circunferencia) {
            int g = n.indice_global;
            if (x) x[g] = n.posicion_cartesiana(0);
            if (y) y[g] = n.posicion_cartesiana(1);
            if (frontera) frontera[g] = n.frontera ? 1 : 0;
        }
    return SPAC_OK;
}

spac_status spac_mesh_interface_cond(const spac_mesh* hm, int32_t k, double* cond) {
    if (!hm || !cond) return set_error(SPAC_E_INVALID_ARG, "null argument");
    const MeshHandle* h = reinterpret_cast<const MeshHandle*>(hm);
    if (k < 0 || k >= (int32_t)h->cond.size()) return set_error(SPAC_E_INVALID_ARG, "interface %d out of range", k);
    *cond = h->cond[k];
    return SPAC_OK;
}

spac_status spac_mesh_reset_system(spac_mesh* hm) {
    if (!hm) return set_error(SPAC_E_INVALID_ARG, "null argument");
    reinterpret_cast<MeshHandle*>(hm)->malla.sistema.Inicializa_CSR();
    return SPAC_OK;
}

spac_status spac_mesh_accumulate(spac_mesh* hm, const spac_row_block* rows) {
    if (!hm || !rows || !rows->B || !rows->G || !rows->G_col || !rows->G_offset)
        return set_error(SPAC_E_INVALID_ARG, "null argument");
    MeshHandle* h = reinterpret_cast<MeshHandle*>(hm);
    SistemaLineal& s = h->malla.sistema;
    for (size_t k = 0; k < h->jobs.size(); ++k) {
        const TrabajoFila& j = h->jobs[k];
        int64_t o = rows->G_offset[k];
        int m = j.frontera ? 0 : (int)(rows->G_offset[k + 1] - o);
        int rc = s.Actualiza_GB(j.nudo_global, rows->B[k], rows->G + o, rows->G_col + o, m);
        if (rc) return set_error(SPAC_E_PATTERN, "row %d: G columns outside the CSR pattern", j.nudo_global);
    }
    return SPAC_OK;
}

spac_status spac_mesh_csr(const spac_mesh* hm, int32_t* G_i, int32_t* G_j, double* G_ij, double* B) {
    if (!hm) return set_error(SPAC_E_INVALID_ARG, "null argument");
    const SistemaLineal& s = reinterpret_cast<const MeshHandle*>(hm)->malla.sistema;
    if (G_i) std::memcpy(G_i, s.G_i.data(), sizeof(int32_t) * s.G_i.size());
    if (G_j) std::memcpy(G_j, s.G_j.data(), sizeof(int32_t) * s.G_j.size());
    if (G_ij) std::memcpy(G_ij, s.G_ij.data(), sizeof(double) * s.G_ij.size());
    if (B) std::memcpy(B, s.B.data(), sizeof(double) * s.B.size());
    return SPAC_OK;
}


spac_status spac_mesh_write_coo(const spac_mesh* hm, const char* G_path, const char* B_path) {
    if (!hm || !G_path || !B_path) return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (!reinterpret_cast<const MeshHandle*>(hm)->malla.sistema.Escribe_GB_COO(G_path, B_path))
        return set_error(SPAC_E_INVALID_ARG, "cannot write %s / %s", G_path, B_path);
    return SPAC_OK;
}

spac_status spac_mesh_write_csr(const spac_mesh* hm, const char* dir) {
    if (!hm || !dir) return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (!reinterpret_cast<const MeshHandle*>(hm)->malla.sistema.Escribe_GB_CSR(dir))
        return set_error(SPAC_E_INVALID_ARG, "cannot write CSR files under %s", dir);
    return SPAC_OK;
}

static double franke_xy(double x, double y) { return Franke_u(Vec2(x, y)); }

spac_status spac_mesh_write_positions(const spac_mesh* hm, int32_t coef_kind, const char* path) {
    if (!hm || !path) return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (coef_kind != SPAC_COEF_FRANKE_POISSON && coef_kind != SPAC_COEF_VARCOEF_FRANKE &&
        coef_kind != SPAC_COEF_FRANKE_ROBIN)
        return set_error(SPAC_E_UNREGISTERED_PLUGIN, "coefficient set %d unknown", coef_kind);
    if (!reinterpret_cast<const MeshHandle*>(hm)->malla.Escribe_Posiciones(path, franke_xy))
        return set_error(SPAC_E_INVALID_ARG, "cannot write %s", path);
    return SPAC_OK;
}

spac_status spac_write_solution(const char* path, int32_t n, const double* u) {
    if (!path || !u || n < 0) return set_error(SPAC_E_INVALID_ARG, "bad arguments");
    SistemaLineal s;
    s.u.assign(u, u + n);
    if (!s.Escribe_Solucion(path)) return set_error(SPAC_E_INVALID_ARG, "cannot write %s", path);
    return SPAC_OK;
}

spac_status spac_write_times_csv(const char* path, int32_t n, const char* const* phases, const int32_t* ranks,
                                 const double* seconds, int32_t append) {
    if (!path || n < 0 || (n > 0 && (!phases || !ranks || !seconds))) return set_error(SPAC_E_INVALID_ARG, "bad arguments");
    for (int i = 0; i < n; ++i) {
        if (!phases[i] || !phases[i][0]) return set_error(SPAC_E_INVALID_ARG, "phase %d has no name", i);
        for (const char* c = phases[i]; *c; ++c)
            if (*c == ',' || *c == '\n') return set_error(SPAC_E_INVALID_ARG, "phase %d: name holds ',' or a newline", i);
    }
    FILE* fp = std::fopen(path, append ? "a" : "w");
    if (!fp) return set_error(SPAC_E_INVALID_ARG, "cannot write %s", path);
    bool ok = true;
    /* std::ostream's default formatting of a double (6 significant digits), as the reference's
       time_file << MPI_Wtime() - t */
    for (int i = 0; i < n && ok; ++i) ok = std::fprintf(fp, "%s,%d,%g\n", phases[i], ranks[i], seconds[i]) > 0;
    ok = (std::fclose(fp) == 0) && ok;
    if (!ok) return set_error(SPAC_E_INVALID_ARG, "cannot write %s", path);
    return SPAC_OK;
}

spac_status spac_read_coo(const char* path, int32_t* n, int32_t* nnz, int32_t* G_i, int32_t* G_j, double* G_ij) {
    if (!path || !n || !nnz) return set_error(SPAC_E_INVALID_ARG, "null argument");
    SistemaLineal s;
    if (!s.Lee_GB_COO(path)) return set_error(SPAC_E_INVALID_ARG, "cannot read COO file %s", path);
    *n = s.n_filas;
    *nnz = s.nnz;
    if (G_i) std::memcpy(G_i, s.G_i.data(), sizeof(int32_t) * (size_t)(s.n_filas + 1));
    if (G_j) std::memcpy(G_j, s.G_j.data(), sizeof(int32_t) * (size_t)s.nnz);
    if (G_ij) std::memcpy(G_ij, s.G_ij.data(), sizeof(double) * (size_t)s.nnz);
    return SPAC_OK;
}

/* ------------------------------------------------------------ GMRES --
 * Restarted GMRES(m) with right Jacobi preconditioning: Arnoldi by modified
 * Gram–Schmidt, Givens rotations on the Hessenberg matrix, back substitution
 * (the loop of Solvers/SistemaLinealSolverGMRES_OMP.hpp:519-590 without its
 * METIS/MUMPS additive Schwarz preconditioner). */
static void spmv(int n, const int32_t* Ip, const int32_t* J, const double* V, const double* x, double* y) {
    for (int i = 0; i < n; ++i) {
        double s = 0.0;
        for (int k = Ip[i]; k < Ip[i + 1]; ++k) s += V[k] * x[J[k]];
        y[i] = s;
    }
}
static double dot(int n, const double* a, const double* b) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

spac_status spac_gmres_solve(int32_t n, const int32_t* Ip, const int32_t* J, const double* V, const double* B, double* u,
                             double tol, int32_t max_iter, int32_t restart, spac_gmres_info* info) {
    if (n < 1 || !Ip || !J || !V || !B || !u) return set_error(SPAC_E_INVALID_ARG, "bad GMRES arguments");
    const auto t_start = std::chrono::steady_clock::now();
    if (restart < 1) restart = 50;
    if (max_iter < 1) max_iter = 1000;
    std::vector<double> dinv(n, 1.0);
    for (int i = 0; i < n; ++i)
        for (int k = Ip[i]; k < Ip[i + 1]; ++k)
            if (J[k] == i && V[k] != 0.0) dinv[i] = 1.0 / V[k];
    const int m = restart;
    std::vector<double> Vb((size_t)(m + 1) * n), H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), w(n), z(n), r(n);
    double bnorm = std::sqrt(dot(n, B, B));
    if (bnorm == 0.0) bnorm = 1.0;
    int it = 0, restarts = 0;
    double res = 0.0;
    bool conv = false;
    while (it < max_iter) {
        spmv(n, Ip, J, V, u, r.data());
        for (int i = 0; i < n; ++i) r[i] = B[i] - r[i];
        double beta = std::sqrt(dot(n, r.data(), r.data()));
        res = beta / bnorm;
        if (res < tol) { conv = true; break; }
        for (int i = 0; i < n; ++i) Vb[i] = r[i] / beta;
        std::fill(g.begin(), g.end(), 0.0);
        g[0] = beta;
        int k = 0;
        for (; k < m && it < max_iter; ++k, ++it) {
            for (int i = 0; i < n; ++i) z[i] = dinv[i] * Vb[(size_t)k * n + i];
            spmv(n, Ip, J, V, z.data(), w.data());
            for (int j = 0; j <= k; ++j) {
                double hjk = dot(n, w.data(), &Vb[(size_t)j * n]);
                H[(size_t)j * m + k] = hjk;
                for (int i = 0; i < n; ++i) w[i] -= hjk * Vb[(size_t)j * n + i];
            }
            double hn = std::sqrt(dot(n, w.data(), w.data()));
            H[(size_t)(k + 1) * m + k] = hn;
            if (hn > 0.0)
                for (int i = 0; i < n; ++i) Vb[(size_t)(k + 1) * n + i] = w[i] / hn;
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
            if (res < tol || hn == 0.0) { ++k; ++it; break; }
        }
        /* back-substitute y and update u += D^-1 V y */
        std::vector<double> yv(k, 0.0);
        for (int i = k - 1; i >= 0; --i) {
            double s = g[i];
            for (int j = i + 1; j < k; ++j) s -= H[(size_t)i * m + j] * yv[j];
            yv[i] = s / H[(size_t)i * m + i];
        }
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = 0; j < k; ++j) s += Vb[(size_t)j * n + i] * yv[j];
            u[i] += dinv[i] * s;
        }
        restarts++;
        if (res < tol) {
            spmv(n, Ip, J, V, u, r.data());
            for (int i = 0; i < n; ++i) r[i] = B[i] - r[i];
            res = std::sqrt(dot(n, r.data(), r.data())) / bnorm;
            conv = res < tol;
            if (conv) break;
        }
    }
    /* max_iter ran out: report the true residual of the last iterate, not the Givens estimate */
    if (!conv) {
        spmv(n, Ip, J, V, u, r.data());
        for (int i = 0; i < n; ++i) r[i] = B[i] - r[i];
        res = std::sqrt(dot(n, r.data(), r.data())) / bnorm;
        conv = res < tol;
    }
    if (info) {
        info->iterations = it;
        info->restarts = restarts;
        info->residual = res;
        info->converged = conv ? 1 : 0;
        info->reserved = 0;
        info->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
        info->device_ms = 0.0;
    }
    return SPAC_OK;
}

} /* extern "C" */
