/*
 * surface_driver.cpp — the reference's serial driver call pattern
 * (main_test_serial.cpp:11-83) on the drop-in C++ surface.
 *
 *   surface_driver cpu   registry, frontier resolution, mesh + CSR (no device)
 *   surface_driver gpu   FK rows of every perimeter node through
 *                        FeynmanKacSolver::Solve_OMP_Analytic(Nudo&, ...),
 *                        accumulated with SistemaLineal::Actualiza_GB
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
#include "Malla.hpp"

using namespace spac;

static double unregistered_f(Vec2 x) { return x(0); }

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
        std::printf("{\"mesh_rc\": %d, \"nodes\": %d, \"nnz\": %d, \"coef_rc\": %d, \"coef_kind\": %d, "
                    "\"dom_rc\": %d, \"dom_kind\": %d, \"sub_rc\": %d, \"sub_kind\": %d, \"bad_rc\": %d, "
                    "\"update_g_rc\": %d, \"update_g_sum\": %.17g}\n",
                    mrc, mesh.numero_total_nudos, mesh.sistema.nnz, rc_coef, kind, rc_dom, dk, rc_sub, sk, rc_bad, urc,
                    sumG);
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
