/*
 * surface_driver.cpp — the reference's serial driver call pattern
 * (main_test_serial.cpp:11-83) on the drop-in C++ surface.
 *
 *   surface_driver cpu   registry, frontier resolution, mesh + CSR (no device)
 *   surface_driver gpu   FK rows of every perimeter node through
 *                        FeynmanKacSolver::Solve_OMP_Analytic(Nudo&, ...),
 *                        accumulated with SistemaLineal::Actualiza_GB
 *   surface_driver cv    one point solve with the plugin control variate, none,
 *                        and a tabulated u_hat (TablaDeValores + Interpolador_Tabla)
 *   surface_driver reflect  one point solve of the Robin problem with the left
 *                        side of the unit square reflecting (Absorbente false there)
 * Prints one JSON object on stdout.
 */
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "BVP.hpp"
#include "BVPs.hpp"
#include "Dominios.hpp"
#include "FeynmanKacSolver.hpp"
#include "Interpolador.hpp"
#include "Malla.hpp"

using namespace spac;

static double unregistered_f(Vec2 x) { return x(0); }
/* Absorbente predicates: the left side (x = 0) reflects; a side that changes type halfway */
static bool left_reflects(Vec2 x) { return !(x(0) <= 0.0); }
static bool mixed_left(Vec2 x) { return !(x(0) <= 0.0 && x(1) < 0.5); }

/* Franke u and its gradient on a T x T grid over [x0, x1]^2 */
static TablaDeValores franke_table(int T, double x0, double x1, bool with_derivatives) {
    TablaDeValores tab;
    std::vector<double> z((size_t)T * T), dx, dy;
    if (with_derivatives) { dx.resize(z.size()); dy.resize(z.size()); }
    tab.Inicializa_TablaDeValores(x0, x0, x1, x1, T, {});
    for (int j = 0; j < T; ++j)
        for (int i = 0; i < T; ++i) {
            Vec2 X(tab.x(i), tab.y(j));
            const size_t k = (size_t)j * T + i;
            z[k] = Franke_u(X);
            if (with_derivatives) { dx[k] = Franke_dudx(X); dy[k] = Franke_dudy(X); }
        }
    tab.Inicializa_TablaDeValores(x0, x0, x1, x1, T, z, dx, dy);
    return tab;
}

int main(int argc, char** argv) {
    std::string mode = argc > 1 ? argv[1] : "cpu";
    Coeficiente_Escalar u, g, c, f;
    Coeficiente_Vectorial b, F;
    Coeficiente_Matricial sigma;
    u.Inicializa(Franke_u);
    g.Inicializa(Franke_g);
    c.Inicializa(Franke_c);
    f.Inicializa(Franke_f);
    sigma.Inicializa(Franke_sigma);
    b.Inicializa(Franke_b);
    F.Inicializa(Franke_F);
    BVP bvp;
    bvp.u = u; bvp.g = g; bvp.c = c; bvp.f = f; bvp.sigma = sigma; bvp.b = b; bvp.F = F;
    std::vector<double> circ = {0.0, 0.0, 0.0};
    std::vector<double> rect = {0.0, 0.0, 1.25, 1.25};
    bvp.frontera_dominio.Inicializa(rect, Rectangulo, Stopping, Stopping);
    bvp.frontera_subdominio.Inicializa(circ, Circulo, Stopping, Stopping);
    Malla mesh;
    std::map<std::string, double> c2;
    c2["esquina"] = 1.5;
    c2["lado"] = 1.1;
    int mrc = mesh.Construye_Cuadrado(rect, 4, 1.0, 0.1428, 44, c2);

    int kind = -1, dk = -1, sk = -1;
    int rc_coef = resolve_coefficients(bvp, &kind);
    int rc_dom = resolve_boundary(bvp.frontera_dominio, &dk);
    int rc_sub = resolve_boundary(bvp.frontera_subdominio, &sk);
    BVP bad = bvp;
    bad.f.Inicializa(unregistered_f);
    int k2 = -1;
    int rc_bad = resolve_coefficients(bad, &k2);

    if (mode == "cpu") {
        /* host Update_G (direct form) on one perimeter node, the reference's own routine */
        const Interfaz& it = mesh.Interfaces[0];
        std::map<int, double> G;
        Vec2 X = it.centro + Vec2(it.radio * std::cos(0.3), it.radio * std::sin(0.3));
        int urc = it.Update_G(1.0, X, G);
        double sumG = 0.0;
        for (auto& kv : G) sumG += kv.second;
        /* TablaDeValores + Interpolador_Tabla on the host */
        Interpolador_Tabla itab, itfd;
        itab.Inicializa(franke_table(513, 0.0, 1.25, true));
        itfd.Inicializa(franke_table(513, 0.0, 1.25, false));
        double eu = 0.0, eg = 0.0, egfd = 0.0, gmax = 0.0;
        for (int k = 0; k < 400; ++k) {
            Vec2 X(0.0125 + 1.225 * std::fmod(0.618034 * k, 1.0), 0.0125 + 1.225 * std::fmod(0.414214 * k, 1.0));
            eu = std::fmax(eu, std::fabs(itab.Interpola(X) - Franke_u(X)));
            eg = std::fmax(eg, std::fmax(std::fabs(itab.Deriva_x(X) - Franke_dudx(X)), std::fabs(itab.Deriva_y(X) - Franke_dudy(X))));
            egfd = std::fmax(egfd, std::fmax(std::fabs(itfd.Deriva_x(X) - Franke_dudx(X)),
                                             std::fabs(itfd.Deriva_y(X) - Franke_dudy(X))));
            gmax = std::fmax(gmax, std::fmax(std::fabs(Franke_dudx(X)), std::fabs(Franke_dudy(X))));
        }
        /* reflecting sides from the Absorbente predicate */
        Frontera fl, fm;
        fl.Inicializa({0.0, 0.0, 1.0, 1.0}, Rectangulo, left_reflects, Stopping);
        fm.Inicializa({0.0, 0.0, 1.0, 1.0}, Rectangulo, mixed_left, Stopping);
        int refl_mask = -1, mixed_mask = -1, stop_mask = -1;
        int refl_rc = resolve_reflecting(fl, &refl_mask);
        int mixed_rc = resolve_reflecting(fm, &mixed_mask);
        int stop_rc = resolve_reflecting(bvp.frontera_dominio, &stop_mask);
        Interpolador base;
        const bool base_nan = std::isnan(base.Interpola(Vec2(0.1, 0.1))) && std::isnan(base.Deriva_x(Vec2(0.1, 0.1)));
        std::printf("{\"mesh_rc\": %d, \"nodes\": %d, \"nnz\": %d, \"coef_rc\": %d, \"coef_kind\": %d, "
                    "\"dom_rc\": %d, \"dom_kind\": %d, \"sub_rc\": %d, \"sub_kind\": %d, \"bad_rc\": %d, "
                    "\"update_g_rc\": %d, \"update_g_sum\": %.17g, \"interp_err_u\": %.3e, \"interp_err_grad\": %.3e, "
                    "\"interp_err_grad_fd\": %.3e, \"grad_max\": %.3e, \"base_nan\": %d, "
                    "\"refl\": [%d, %d], \"refl_mixed_rc\": %d, \"refl_stopping\": [%d, %d]}\n",
                    mrc, mesh.numero_total_nudos, mesh.sistema.nnz, rc_coef, kind, rc_dom, dk, rc_sub, sk, rc_bad, urc,
                    sumG, eu, eg, egfd, gmax, base_nan ? 1 : 0, refl_rc, refl_mask, mixed_rc, stop_rc, stop_mask);
        return 0;
    }

    if (mode == "reflect") {
        BVP pb = bvp;
        pb.psi.Inicializa(Robin_psi);
        pb.varphi.Inicializa(Robin_varphi);
        pb.frontera_dominio.Inicializa({0.0, 0.0, 1.0, 1.0}, Rectangulo, left_reflects, Stopping);
        pb.frontera_subdominio = Frontera();
        int kr = -1;
        int krc = resolve_coefficients(pb, &kr);
        FeynmanKacSolver s;
        int rc = s.Solve_OMP_Analytic(Vec2(0.05, 0.5), 4000, 1e-3, 0.08, pb);
        std::printf("{\"coef_rc\": %d, \"coef_kind\": %d, \"rc\": %d, \"phi\": %.17g, \"phi_VR\": %.17g, "
                    "\"steps\": %lld, \"u\": %.17g}\n",
                    krc, kr, rc, s.phi, s.phi_VR, (long long)s.path_steps, Franke_u(Vec2(0.05, 0.5)));
        return 0;
    }

    if (mode == "cv") {
        BVP pb = bvp;
        pb.frontera_subdominio = Frontera();
        Interpolador_Tabla it;
        it.Inicializa(franke_table(513, 0.0, 1.25, true));
        FeynmanKacSolver s_plugin, s_none, s_table;
        int r1 = s_plugin.Solve_OMP_Analytic(Vec2(0.6, 0.7), 4000, 1e-3, 0.1, pb);
        int r2 = s_none.Sin_Variable_Control();
        if (!r2) r2 = s_none.Solve_OMP_Analytic(Vec2(0.6, 0.7), 4000, 1e-3, 0.1, pb);
        int r3 = s_table.Variable_Control(it, false);
        if (!r3) r3 = s_table.Solve_OMP_Analytic(Vec2(0.6, 0.7), 4000, 1e-3, 0.1, pb);
        FeynmanKacSolver* S[3] = {&s_plugin, &s_none, &s_table};
        std::printf("{\"rc\": [%d, %d, %d], \"u\": %.17g, \"modes\": [", r1, r2, r3, Franke_u(Vec2(0.6, 0.7)));
        for (int k = 0; k < 3; ++k)
            std::printf("%s{\"phi\": %.17g, \"phiphi\": %.17g, \"phi_VR\": %.17g, \"phiphi_VR\": %.17g, "
                        "\"xixi\": %.17g}", k ? ", " : "", S[k]->phi, S[k]->phiphi, S[k]->phi_VR, S[k]->phiphi_VR,
                        S[k]->xixi);
        std::printf("]}\n");
        return 0;
    }

    FeynmanKacSolver solver_fk; /* seed 0 */
    std::vector<double> Bs, Gsum;
    std::vector<int> rows;
    int worst = 0, calls = 0;
    const unsigned paths = 400;
    const double h = 1e-3;
    for (Interfaz& it : mesh.Interfaces) {
        if (!it.es_perimeter) continue;
        circ[0] = it.radio; circ[1] = it.centro(0); circ[2] = it.centro(1);
        bvp.frontera_subdominio.Inicializa(circ, Circulo, Stopping, Stopping);
        for (Nudo& nd : it.nudos_interior) {
            int rc = solver_fk.Solve_OMP_Analytic(nd, paths, h, 1.0, bvp, it);
            if (rc < worst) worst = rc;
            int arc = mesh.sistema.Actualiza_GB(nd);
            if (arc < worst) worst = arc;
            double s = 0.0;
            for (auto& kv : nd.G) s += kv.second;
            rows.push_back(nd.indice_global);
            Bs.push_back(nd.B);
            Gsum.push_back(s);
            if (!nd.frontera) calls++;
        }
    }
    /* point solve, the commented test of main_test_serial.cpp:90-100 */
    FeynmanKacSolver point;
    BVP pb = bvp;
    pb.frontera_subdominio = Frontera();
    int prc = point.Solve_OMP_Analytic(Vec2(0.6, 0.7), 2000, 1e-3, 0.1, pb);
    std::printf("{\"worst_rc\": %d, \"calls\": %d, \"rows\": [", worst, calls);
    for (size_t k = 0; k < rows.size(); ++k) std::printf("%s%d", k ? ", " : "", rows[k]);
    std::printf("], \"B\": [");
    for (size_t k = 0; k < Bs.size(); ++k) std::printf("%s%.17g", k ? ", " : "", Bs[k]);
    std::printf("], \"Gsum\": [");
    for (size_t k = 0; k < Gsum.size(); ++k) std::printf("%s%.17g", k ? ", " : "", Gsum[k]);
    std::printf("], \"point_rc\": %d, \"point_phi\": %.17g, \"point_phi_VR\": %.17g, \"point_steps\": %lld, "
                "\"csr_B_sum\": %.17g}\n",
                prc, point.phi, point.phi_VR, (long long)point.path_steps,
                [&] { double s = 0; for (double v : mesh.sistema.B) s += v; return s; }());
    return 0;
}
