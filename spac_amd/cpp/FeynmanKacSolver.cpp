/*
 * FeynmanKacSolver.cpp — the reference's solver interface over the C ABI.
 */
#include "FeynmanKacSolver.hpp"

#include <cmath>
#include <cstring>

#include "spac_error.h"

namespace spac {

FeynmanKacSolver::FeynmanKacSolver() = default;
FeynmanKacSolver::FeynmanKacSolver(unsigned int MPIrank) : seed_(MPIrank) {}
FeynmanKacSolver::~FeynmanKacSolver() {
    if (prob_) spac_fk_problem_destroy(prob_);
}

/* The interpolators of the numeric coefficients named by mask, in a fixed order. */
static std::vector<InterpoladorPtr> table_key(const BVP& P, int mask) {
    std::vector<InterpoladorPtr> k;
    if (mask & SPAC_TAB_SIGMA)
        k.insert(k.end(), {P.sigma.interpolador_11, P.sigma.interpolador_12, P.sigma.interpolador_21,
                           P.sigma.interpolador_22});
    if (mask & SPAC_TAB_DRIFT) k.insert(k.end(), {P.b.interpolador_x, P.b.interpolador_y});
    if (mask & SPAC_TAB_KILLING) k.push_back(P.c.interpolador);
    if (mask & SPAC_TAB_SOURCE) k.push_back(P.f.interpolador);
    if (mask & SPAC_TAB_DIRICHLET) k.push_back(P.g.interpolador);
    if (mask & SPAC_TAB_SOLUTION) k.push_back(P.u.interpolador);
    return k;
}

int FeynmanKacSolver::sample_tables(const BVP& P, int mask, const double dom[4], bool* changed) {
    *changed = false;
    if (!mask) {
        *changed = coef_mask_ != 0;
        coef_mask_ = 0;
        coef_tab_.clear();
        coef_key_.clear();
        return SPAC_OK;
    }
    if (T_tabla < 2 || T_tabla > 8192) return set_error(SPAC_E_INVALID_ARG, "T_tabla must be in [2, 8192]");
    std::vector<InterpoladorPtr> key = table_key(P, mask);
    bool same = mask == coef_mask_ && T_tabla == coef_T_ && key.size() == coef_key_.size() && !coef_tab_.empty();
    for (int i = 0; same && i < 4; ++i) same = dom[i] == coef_dom_[i];
    for (size_t i = 0; same && i < key.size(); ++i) same = key[i].get() == coef_key_[i].get();
    if (same) return SPAC_OK;
    const int T = (int)T_tabla;
    std::vector<double> v((size_t)T * T * 8, 0.0);
    const double dx = (dom[2] - dom[0]) / (double)(T - 1), dy = (dom[3] - dom[1]) / (double)(T - 1);
    for (int j = 0; j < T; ++j)
        for (int i = 0; i < T; ++i) {
            /* the node positions of TablaDeValores::x(i), y(j) */
            const Vec2 X(dom[0] + (double)i * dx, dom[1] + (double)j * dy);
            double* o = &v[((size_t)j * T + i) * 8];
            if (mask & SPAC_TAB_SIGMA) {
                const Mat2 S = P.sigma.Evalua(X);
                const double tol = 1e-12 * (1.0 + std::fabs(S(0, 0)));
                if (!(std::fabs(S(0, 1)) <= tol && std::fabs(S(1, 0)) <= tol && std::fabs(S(1, 1) - S(0, 0)) <= tol))
                    return set_error(SPAC_E_INVALID_ARG,
                                     "numeric sigma must be s(x) I on the device (sigma at (%g, %g) is [%g %g; %g %g])",
                                     X(0), X(1), S(0, 0), S(0, 1), S(1, 0), S(1, 1));
                o[0] = S(0, 0);
            }
            if (mask & SPAC_TAB_DRIFT) {
                const Vec2 B = P.b.Evalua(X);
                o[1] = B(0);
                o[2] = B(1);
            }
            if (mask & SPAC_TAB_KILLING) o[3] = P.c.Evalua(X);
            if (mask & SPAC_TAB_SOURCE) o[4] = P.f.Evalua(X);
            if (mask & SPAC_TAB_DIRICHLET) o[5] = P.g.Evalua(X);
            if (mask & SPAC_TAB_SOLUTION) o[6] = P.u.Evalua(X);
            for (int c = 0; c < 7; ++c)
                if (!std::isfinite(o[c]))
                    return set_error(SPAC_E_NAN, "numeric coefficient %d is not finite at (%g, %g)", c, X(0), X(1));
        }
    coef_tab_ = std::move(v);
    coef_mask_ = mask;
    coef_key_ = std::move(key);
    coef_T_ = T_tabla;
    for (int i = 0; i < 4; ++i) coef_dom_[i] = dom[i];
    *changed = true;
    return SPAC_OK;
}

int FeynmanKacSolver::apply_tables() {
    if (!prob_) return SPAC_OK;
    spac_coef_tables t;
    std::memset(&t, 0, sizeof t);
    t.mask = coef_mask_;
    if (coef_mask_) {
        t.table_size = (int32_t)coef_T_;
        for (int i = 0; i < 4; ++i) t.table_domain[i] = coef_dom_[i];
        t.values = coef_tab_.data();
    }
    const int rc = spac_fk_set_coef_tables(prob_, &t);
    if (rc) { /* the device holds no tables: sample and install again on the next call */
        coef_mask_ = 0;
        coef_tab_.clear();
        coef_key_.clear();
        if (prob_) { spac_fk_problem_destroy(prob_); prob_ = nullptr; }
    }
    return rc;
}

int FeynmanKacSolver::ensure_problem(const BVP& P, bool with_subdomain, double rho) {
    spac_problem_desc d;
    std::memset(&d, 0, sizeof d);
    int kind = 0, tab_mask = 0;
    bool F_zero = false;
    if (resolve_numeric_coefficients(P, &kind, &tab_mask, &F_zero) != SPAC_OK)
        return set_error(SPAC_E_UNREGISTERED_PLUGIN,
                         "BVP coefficients are neither a registered device set nor numeric coefficients over the "
                         "variable-coefficient set");
    d.coef_kind = kind;
    int dk = 0;
    if (resolve_boundary(P.frontera_dominio, &dk) != SPAC_OK || dk != SPAC_BOUNDARY_RECTANGLE ||
        P.frontera_dominio.parametros.size() < 4)
        return set_error(SPAC_E_UNREGISTERED_PLUGIN, "domain frontier must be the registered Rectangulo");
    d.domain_kind = dk;
    for (int i = 0; i < 4; ++i) d.domain_params[i] = P.frontera_dominio.parametros[i];
    int mask = 0;
    if (resolve_reflecting(P.frontera_dominio, &mask) != SPAC_OK)
        return set_error(SPAC_E_UNREGISTERED_PLUGIN, "Absorbente must be constant along each side of the rectangle");
    d.reflecting = mask;
    d.reflect_rho = mask ? rho : 0.0; /* rho only matters with reflecting sides */
    d.subdomain_kind = SPAC_BOUNDARY_NONE;
    if (with_subdomain) {
        int sk = 0;
        if (resolve_boundary(P.frontera_subdominio, &sk) != SPAC_OK)
            return set_error(SPAC_E_UNREGISTERED_PLUGIN, "subdomain frontier is not registered");
        if (sk == SPAC_BOUNDARY_CIRCLE) {
            if (P.frontera_subdominio.parametros.size() < 3)
                return set_error(SPAC_E_INVALID_ARG, "Circulo needs (R, cx, cy)");
            d.subdomain_kind = sk;
            for (int i = 0; i < 3; ++i) d.subdomain_params[i] = P.frontera_subdominio.parametros[i];
        } else if (sk != SPAC_BOUNDARY_NONE) {
            return set_error(SPAC_E_UNREGISTERED_PLUGIN, "subdomain frontier kind %d not supported on the device", sk);
        }
    }
    d.gobet_constant = -0.5826; /* Solvers/FeynmanKacSolver.cpp:268,283 */
    d.max_steps = (int32_t)max_steps;
    d.precision = SPAC_FP64;
    if (tab_mask && mask)
        return set_error(SPAC_E_INVALID_ARG, "numeric coefficients need absorbing sides (no reflecting side)");
    if ((P.psi.Es_Numerico() || P.varphi.Es_Numerico()) && mask)
        return set_error(SPAC_E_UNREGISTERED_PLUGIN, "numeric Robin data (psi, varphi) has no device form");
    bool tables_changed = false;
    int rc = sample_tables(P, tab_mask, d.domain_params, &tables_changed);
    if (rc) return rc;
    const bool cv_changed = F_zero != F_zero_;
    F_zero_ = F_zero;
    if (prob_ && std::memcmp(&d, &desc_, sizeof d) == 0) {
        if (tables_changed && (rc = apply_tables())) return rc;
        return cv_changed ? apply_cv() : SPAC_OK;
    }
    if (prob_) { spac_fk_problem_destroy(prob_); prob_ = nullptr; }
    rc = spac_fk_problem_create(&d, &prob_);
    if (rc) return rc;
    desc_ = d;
    if ((rc = apply_tables())) return rc;
    return apply_cv();
}

int FeynmanKacSolver::apply_cv() {
    if (!prob_) return SPAC_OK; /* applied when the problem is created */
    spac_cv_config c = cv_;
    /* the plugin control variate with F = 0 (the default) is xi = 0: plain Monte Carlo */
    if (c.mode == SPAC_CV_ANALYTIC && F_zero_) c.mode = SPAC_CV_NONE;
    c.table = cv_table_.empty() ? nullptr : cv_table_.data();
    return spac_fk_set_control_variate(prob_, &c);
}

int FeynmanKacSolver::Variable_Control_Plugin() {
    std::memset(&cv_, 0, sizeof cv_);
    cv_.mode = SPAC_CV_ANALYTIC;
    cv_table_.clear();
    return apply_cv();
}

int FeynmanKacSolver::Sin_Variable_Control() {
    std::memset(&cv_, 0, sizeof cv_);
    cv_.mode = SPAC_CV_NONE;
    cv_table_.clear();
    return apply_cv();
}

int FeynmanKacSolver::Variable_Control(const Interpolador_Tabla& u_hat, bool fallback, double max_ratio) {
    if (u_hat.tabla().empty()) return set_error(SPAC_E_INVALID_ARG, "Interpolador_Tabla has no valid table");
    std::memset(&cv_, 0, sizeof cv_);
    cv_.mode = SPAC_CV_TABLE;
    cv_.fallback = fallback ? 1 : 0;
    cv_.max_ratio = max_ratio;
    cv_.table_size = u_hat.valores.T;
    cv_.table_domain[0] = u_hat.valores.xmin;
    cv_.table_domain[1] = u_hat.valores.ymin;
    cv_.table_domain[2] = u_hat.valores.xmax;
    cv_.table_domain[3] = u_hat.valores.ymax;
    cv_table_ = u_hat.tabla();
    return apply_cv();
}

void FeynmanKacSolver::copy_stats(const spac_point_stats& s) {
    phi = s.phi;
    phiphi = s.phiphi;
    phi_VR = s.phi_VR;
    phiphi_VR = s.phiphi_VR;
    xi = s.xi;
    xixi = s.xixi;
    xiphi = s.xiphi;
    APL = s.APL;
    RNGC = s.RNGC;
    path_steps = s.path_steps;
}

int FeynmanKacSolver::Solve_OMP_Analytic(Vec2 X0, unsigned int n, double dt, double rho, const BVP& P) {
    N = n;
    h = dt;
    sqrth = std::sqrt(dt);
    int rc = ensure_problem(P, true, rho);
    if (rc) return rc;
    double x = X0(0), y = X0(1);
    uint32_t stream = calls_++;
    spac_solve_batch b;
    std::memset(&b, 0, sizeof b);
    b.n = 1;
    b.x0 = &x;
    b.y0 = &y;
    b.paths = &n;
    b.h = &dt;
    b.stream = &stream;
    spac_point_stats st;
    rc = spac_fk_simulate(prob_, &b, seed_, nullptr, &st);
    if (rc) return rc;
    copy_stats(st);
    return st.status;
}

/* The device copy of one interface (Mesh/Interfaz.hpp:8-30). */
struct OneInterface {
    std::vector<double> cx, cy, R, c2, theta, ipsi;
    std::vector<int32_t> code, off, glob;
    std::vector<int64_t> ioff;
    spac_interface_table t{};
};

static int build_table(const std::vector<const Interfaz*>& ifs, OneInterface& o) {
    bool stencil = false, imq = false;
    for (const Interfaz* it : ifs) (it->stencil_size > 0 ? stencil : imq) = true;
    if (stencil && imq) return set_error(SPAC_E_INVALID_ARG, "interfaces mix IMQ and stencil interpolation");
    o.off.push_back(0);
    for (const Interfaz* it : ifs) {
        o.cx.push_back(it->centro(0));
        o.cy.push_back(it->centro(1));
        o.R.push_back(it->radio);
        o.c2.push_back(it->c2);
        o.code.push_back(it->interior);
        for (const Nudo& n : it->nudos_circunferencia) {
            o.theta.push_back(n.posicion_angular);
            o.glob.push_back(n.indice_global);
        }
        o.off.push_back((int32_t)o.theta.size());
        if (stencil) {
            o.ioff.push_back((int64_t)o.ipsi.size());
            o.ipsi.insert(o.ipsi.end(), it->stencil_inverses.begin(), it->stencil_inverses.end());
        } else if (it->iPsi.rows == (int)it->nudos_circunferencia.size() && it->iPsi.rows > 0) {
            o.ioff.push_back((int64_t)o.ipsi.size());
            o.ipsi.insert(o.ipsi.end(), it->iPsi.d.begin(), it->iPsi.d.end());
        } else {
            o.ioff.push_back(-1);
        }
    }
    if (o.ipsi.empty()) o.ipsi.push_back(0.0);
    spac_interface_table& t = o.t;
    t.n_interfaces = (int32_t)ifs.size();
    t.rbf_kind = stencil ? SPAC_RBF_GAUSS_STENCIL : SPAC_RBF_IMQ_GLOBAL;
    t.cx = o.cx.data(); t.cy = o.cy.data(); t.radius = o.R.data(); t.code = o.code.data(); t.c2 = o.c2.data();
    t.node_offset = o.off.data(); t.node_theta = o.theta.data(); t.node_global = o.glob.data();
    t.ipsi_offset = o.ioff.data(); t.ipsi = o.ipsi.data();
    t.gauss_eps = stencil ? ifs[0]->gauss_eps : 0.0;
    t.stencil_size = stencil ? ifs[0]->stencil_size : 0;
    return SPAC_OK;
}

int FeynmanKacSolver::Solve_Nodes(std::vector<Nudo*>& nudos, const std::vector<const Interfaz*>& ifs, unsigned int n,
                                  double dt, const BVP& P, double rho) {
    if (nudos.size() != ifs.size() || nudos.empty())
        return set_error(SPAC_E_INVALID_ARG, "one interface per node is needed");
    N = n;
    h = dt;
    sqrth = std::sqrt(dt);
    int rc = ensure_problem(P, false, rho);
    if (rc) return rc;
    /* distinct interfaces -> table rows */
    std::vector<const Interfaz*> uniq;
    std::vector<int32_t> iface(nudos.size());
    for (size_t k = 0; k < ifs.size(); ++k) {
        size_t u = 0;
        while (u < uniq.size() && uniq[u] != ifs[k]) ++u;
        if (u == uniq.size()) uniq.push_back(ifs[k]);
        iface[k] = (int32_t)u;
    }
    OneInterface o;
    if ((rc = build_table(uniq, o))) return rc;
    const size_t m = nudos.size();
    std::vector<double> x(m), y(m), hh(m, dt);
    std::vector<int32_t> fr(m);
    std::vector<uint32_t> paths(m, n), stream(m);
    for (size_t k = 0; k < m; ++k) {
        x[k] = nudos[k]->posicion_cartesiana(0);
        y[k] = nudos[k]->posicion_cartesiana(1);
        fr[k] = nudos[k]->frontera ? 1 : 0;
        stream[k] = calls_++;
    }
    spac_node_batch nb;
    std::memset(&nb, 0, sizeof nb);
    nb.n = (int32_t)m;
    nb.x = x.data(); nb.y = y.data(); nb.iface = iface.data(); nb.frontera = fr.data();
    nb.paths = paths.data(); nb.h = hh.data(); nb.stream = stream.data();
    int64_t total = 0;
    for (size_t k = 0; k < m; ++k) total += o.off[iface[k] + 1] - o.off[iface[k]];
    std::vector<double> B(m), G((size_t)total);
    std::vector<int32_t> col((size_t)total);
    std::vector<int64_t> goff(m + 1);
    std::vector<spac_point_stats> st(m);
    spac_row_block out;
    out.B = B.data(); out.G = G.data(); out.G_col = col.data(); out.G_offset = goff.data(); out.stats = st.data();
    rc = spac_fk_assemble_rows(prob_, &o.t, &nb, seed_, &out);
    if (rc && rc != SPAC_E_MAX_STEPS) return rc;
    for (size_t k = 0; k < m; ++k) {
        Nudo& nd = *nudos[k];
        nd.B = B[k];
        nd.G.clear();
        if (!nd.frontera)
            for (int64_t q = goff[k]; q < goff[k + 1]; ++q) nd.G[col[q]] = G[q];
    }
    copy_stats(st[m - 1]);
    return rc;
}

int FeynmanKacSolver::Solve_OMP_Analytic(Nudo& nudo, unsigned int n, double dt, double rho, const BVP& P,
                                         const Interfaz& interfaz) {
    if (nudo.frontera) { /* Solvers/FeynmanKacSolver.cpp:346-347 */
        nudo.B = P.g.Evalua(nudo.posicion_cartesiana);
        return SPAC_OK;
    }
    std::vector<Nudo*> ns{&nudo};
    std::vector<const Interfaz*> is{&interfaz};
    return Solve_Nodes(ns, is, n, dt, P, rho);
}

} // namespace spac
