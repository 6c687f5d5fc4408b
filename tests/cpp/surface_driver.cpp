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
 *   surface_driver rbf   Interpolador_RBF and the numeric coefficient classes (no device)
 *   surface_driver numeric [table.bin]  point solves with Coeficiente_*_Numerico
 *                        (the variable-coefficient set as Interpolador_Tabla / RBF data),
 *                        the sampled device table written to table.bin
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
#include "Interpolador_RBF.hpp"
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

    if (mode == "rbf") {
        /* Interpolador_RBF on scattered Halton centres of the unit square */
        auto halton = [](int k, int b) { double f = 1.0, r = 0.0; for (int i = k; i > 0; i /= b) { f /= b; r += f * (i % b); } return r; };
        auto centres = [&](int n) {
            std::vector<Vec2> c;
            for (int k = 1; k <= n; ++k) c.emplace_back(halton(k, 2), halton(k, 3));
            return c;
        };
        auto lin = [](Vec2 X) { return 1.0 + 2.0 * X(0) - 3.0 * X(1); };
        std::vector<Vec2> c400 = centres(400), c1600 = centres(1600);
        std::vector<double> zl, z400, z1600;
        for (const Vec2& X : c400) { zl.push_back(lin(X)); z400.push_back(Franke_u(X)); }
        for (const Vec2& X : c1600) z1600.push_back(Franke_u(X));
        Interpolador_RBF rl, r400, r1600, bad;
        int rc = rl.Inicializa(c400, zl) | r400.Inicializa(c400, z400) | r1600.Inicializa(c1600, z1600);
        std::vector<Vec2> dup = {Vec2(0.1, 0.1), Vec2(0.1, 0.1), Vec2(0.5, 0.2), Vec2(0.3, 0.9)};
        int bad_rc = bad.Inicializa(dup, {1.0, 2.0, 3.0, 4.0});
        const bool bad_nan = std::isnan(bad.Interpola(Vec2(0.2, 0.2)));
        double at_centres = 0.0, lin_err = 0.0, lin_d1 = 0.0, lin_d2 = 0.0, e400 = 0.0, e1600 = 0.0, fd_err = 0.0,
               polar_err = 0.0;
        for (int k = 0; k < 400; ++k) at_centres = std::fmax(at_centres, std::fabs(r400.Interpola(c400[k]) - z400[k]));
        const double hh = 1e-4;
        for (int k = 0; k < 300; ++k) {
            Vec2 X(0.05 + 0.9 * std::fmod(0.618034 * k + 0.1, 1.0), 0.05 + 0.9 * std::fmod(0.414214 * k + 0.3, 1.0));
            lin_err = std::fmax(lin_err, std::fabs(rl.Interpola(X) - lin(X)));
            lin_d1 = std::fmax(lin_d1, std::fmax(std::fabs(rl.Deriva_x(X) - 2.0), std::fabs(rl.Deriva_y(X) + 3.0)));
            lin_d2 = std::fmax(lin_d2, std::fmax(std::fabs(rl.Deriva_x2(X)),
                                                 std::fmax(std::fabs(rl.Deriva_y2(X)), std::fabs(rl.Deriva_xy(X)))));
            e400 = std::fmax(e400, std::fabs(r400.Interpola(X) - Franke_u(X)));
            e1600 = std::fmax(e1600, std::fabs(r1600.Interpola(X) - Franke_u(X)));
            /* analytic derivatives of the interpolant against central differences of it */
            const Vec2 ex(hh, 0.0), ey(0.0, hh);
            auto S = [&](Vec2 P) { return r400.Interpola(P); };
            const double fx = (S(X + ex) - S(X - ex)) / (2 * hh), fy = (S(X + ey) - S(X - ey)) / (2 * hh);
            const double fxx = (S(X + ex) - 2 * S(X) + S(X - ex)) / (hh * hh);
            const double fyy = (S(X + ey) - 2 * S(X) + S(X - ey)) / (hh * hh);
            const double fxy = (S(X + ex + ey) - S(X + ex - ey) - S(X - ex + ey) + S(X - ex - ey)) / (4 * hh * hh);
            fd_err = std::fmax(fd_err, std::fmax(std::fabs(r400.Deriva_x(X) - fx), std::fabs(r400.Deriva_y(X) - fy)));
            fd_err = std::fmax(fd_err, 1e-3 * std::fmax(std::fabs(r400.Deriva_x2(X) - fxx),
                                                        std::fmax(std::fabs(r400.Deriva_y2(X) - fyy),
                                                                  std::fabs(r400.Deriva_xy(X) - fxy))));
            /* polar derivatives about the centre: differences along the ray and the circle */
            const Vec2 C = r400.centro, d = X - C;
            const double r = d.norm(), t = std::atan2(d(1), d(0));
            auto Pp = [&](double rr, double tt) { return S(C + Vec2(rr * std::cos(tt), rr * std::sin(tt))); };
            const double dr = (Pp(r + hh, t) - Pp(r - hh, t)) / (2 * hh);
            const double dt = (Pp(r, t + hh) - Pp(r, t - hh)) / (2 * hh);
            const double drr = (Pp(r + hh, t) - 2 * Pp(r, t) + Pp(r - hh, t)) / (hh * hh);
            const double dtt = (Pp(r, t + hh) - 2 * Pp(r, t) + Pp(r, t - hh)) / (hh * hh);
            const double dtr = (Pp(r + hh, t + hh) - Pp(r + hh, t - hh) - Pp(r - hh, t + hh) + Pp(r - hh, t - hh)) /
                               (4 * hh * hh);
            polar_err = std::fmax(polar_err, std::fmax(std::fabs(r400.Deriva_r(X) - dr), std::fabs(r400.Deriva_t(X) - dt)));
            polar_err = std::fmax(polar_err, 1e-3 * std::fmax(std::fabs(r400.Deriva_r2(X) - drr),
                                                              std::fmax(std::fabs(r400.Deriva_t2(X) - dtt),
                                                                        std::fabs(r400.Deriva_tr(X) - dtr))));
        }
        /* the numeric surface: Es_Numerico, a numeric coefficient copied into a BVP field */
        auto ip = std::make_shared<Interpolador_RBF>(r400);
        Coeficiente_Escalar_Numerico cn(ip);
        Coeficiente_Escalar sliced = cn;
        Coeficiente_Matricial_Numerico sn(ip);
        Coeficiente_Vectorial_Numerico vn(ip, nullptr);
        const Vec2 P0(0.3, 0.6);
        const Mat2 S0 = sn.Evalua(P0);
        const Vec2 V0 = vn.Evalua(P0);
        BVP nb = bvp;
        nb.c = cn;
        int nk = -1, nmask = -1, kk = -1;
        bool fz = false;
        int num_rc_poisson = resolve_numeric_coefficients(nb, &nk, &nmask, &fz); /* Franke Poisson base: refused */
        int plain_rc_numeric = resolve_coefficients(nb, &kk);
        BVP vb;
        vb.u.Inicializa(Franke_u); vb.g.Inicializa(Franke_g); vb.c = cn; vb.f.Inicializa(VarCoef_f);
        vb.b.Inicializa(VarCoef_b); vb.sigma = sn;
        int vk = -1, vmask = -1;
        bool vfz = false;
        int num_rc = resolve_numeric_coefficients(vb, &vk, &vmask, &vfz);
        BVP vf = vb;
        vf.F = vn;
        int fk = -1, fmask = -1;
        bool ffz = false;
        int numF_rc = resolve_numeric_coefficients(vf, &fk, &fmask, &ffz);
        std::printf("{\"rc\": %d, \"bad_rc\": %d, \"bad_nan\": %d, \"eps\": %.6g, \"at_centres\": %.3e, "
                    "\"lin_err\": %.3e, \"lin_d1\": %.3e, \"lin_d2\": %.3e, \"e400\": %.3e, \"e1600\": %.3e, "
                    "\"fd_err\": %.3e, \"polar_err\": %.3e, \"numerico\": [%d, %d, %d, %d, %d, %d], "
                    "\"sliced_value\": %.17g, \"direct_value\": %.17g, \"matrix\": [%.17g, %.17g, %.17g, %.17g], "
                    "\"vector\": [%.17g, %.17g], \"num_rc_poisson\": %d, \"plain_rc_numeric\": %d, "
                    "\"num_rc\": %d, \"num_kind\": %d, \"num_mask\": %d, \"num_F_zero\": %d, \"numF_rc\": %d}\n",
                    rc, bad_rc, bad_nan ? 1 : 0, r400.epsilon(), at_centres, lin_err, lin_d1, lin_d2, e400, e1600,
                    fd_err, polar_err, cn.Es_Numerico() ? 1 : 0, sliced.Es_Numerico() ? 1 : 0,
                    sn.Es_Numerico() ? 1 : 0, vn.Es_Numerico() ? 1 : 0, bvp.c.Es_Numerico() ? 1 : 0,
                    Coeficiente_Escalar(Franke_c).Es_Numerico() ? 1 : 0, sliced.Evalua(P0), r400.Interpola(P0),
                    S0(0, 0), S0(0, 1), S0(1, 0), S0(1, 1), V0(0), V0(1), num_rc_poisson, plain_rc_numeric, num_rc, vk,
                    vmask, vfz ? 1 : 0, numF_rc);
        return 0;
    }

    if (mode == "numeric") {
        /* the variable-coefficient set, compiled and as numeric coefficients (Interpolador_Tabla
           of each coefficient at the solver's own grid), one point solve each (same paths) */
        const int T = 513;
        auto tab = [&](auto fn) {
            TablaDeValores t;
            t.Inicializa_TablaDeValores(0.0, 0.0, 1.0, 1.0, T, {});
            std::vector<double> z((size_t)T * T);
            for (int j = 0; j < T; ++j)
                for (int i = 0; i < T; ++i) z[(size_t)j * T + i] = fn(Vec2(t.x(i), t.y(j)));
            t.Inicializa_TablaDeValores(0.0, 0.0, 1.0, 1.0, T, z);
            auto it = std::make_shared<Interpolador_Tabla>();
            it->Inicializa(t);
            return std::shared_ptr<const Interpolador>(it);
        };
        BVP vc;
        vc.u.Inicializa(Franke_u); vc.g.Inicializa(Franke_g); vc.c.Inicializa(VarCoef_c); vc.f.Inicializa(VarCoef_f);
        vc.b.Inicializa(VarCoef_b); vc.F.Inicializa(VarCoef_F); vc.sigma.Inicializa(VarCoef_sigma);
        vc.frontera_dominio.Inicializa({0.0, 0.0, 1.0, 1.0}, Rectangulo, Stopping, Stopping);
        BVP nv = vc;
        nv.sigma = Coeficiente_Matricial_Numerico(tab([](Vec2 X) { return VarCoef_s(X); }));
        nv.b = Coeficiente_Vectorial_Numerico(tab([](Vec2 X) { return VarCoef_b(X)(0); }),
                                              tab([](Vec2 X) { return VarCoef_b(X)(1); }));
        nv.c = Coeficiente_Escalar_Numerico(tab([](Vec2 X) { return VarCoef_c(X); }));
        nv.f = Coeficiente_Escalar_Numerico(tab([](Vec2 X) { return VarCoef_f(X); }));
        nv.g = Coeficiente_Escalar_Numerico(tab([](Vec2 X) { return Franke_g(X); }));
        nv.u = Coeficiente_Escalar_Numerico(tab([](Vec2 X) { return Franke_u(X); }));
        const Vec2 X0(0.6, 0.7);
        const unsigned n = 4000;
        FeynmanKacSolver s_c, s_n;
        s_n.T_tabla = T;
        int r1 = s_c.Sin_Variable_Control();
        if (!r1) r1 = s_c.Solve_OMP_Analytic(X0, n, 1e-3, 0.0, vc);
        int r2 = s_n.Sin_Variable_Control();
        if (!r2) r2 = s_n.Solve_OMP_Analytic(X0, n, 1e-3, 0.0, nv);
        /* a second call reuses the sampled table; then a numeric c from scattered RBF data */
        const double n_phi = s_n.phi, n_phiphi = s_n.phiphi, n_phi_VR = s_n.phi_VR;
        const long long n_steps = (long long)s_n.path_steps;
        int r3 = s_n.Solve_OMP_Analytic(X0, n, 1e-3, 0.0, nv);
        const bool same_table = s_n.Tabla_Coeficientes().size() == (size_t)T * T * 8;
        if (argc > 2) { /* the sampled table, for the restatement */
            FILE* fp = std::fopen(argv[2], "wb");
            if (fp) {
                std::fwrite(s_n.Tabla_Coeficientes().data(), sizeof(double), s_n.Tabla_Coeficientes().size(), fp);
                std::fclose(fp);
            }
        }
        std::vector<Vec2> cs;
        std::vector<double> zc;
        for (int j = 0; j < 21; ++j)
            for (int i = 0; i < 21; ++i) {
                cs.emplace_back(i / 20.0, j / 20.0);
                zc.push_back(VarCoef_c(cs.back()));
            }
        auto rbf = std::make_shared<Interpolador_RBF>();
        int rbf_rc = rbf->Inicializa(cs, zc);
        BVP rb = vc;
        rb.c = Coeficiente_Escalar_Numerico(rbf);
        FeynmanKacSolver s_r;
        s_r.T_tabla = 129;
        int r4 = s_r.Sin_Variable_Control();
        if (!r4) r4 = s_r.Solve_OMP_Analytic(X0, n, 1e-3, 0.0, rb);
        double rbf_tab_err = 0.0;
        for (int j = 0; j < 129; j += 8)
            for (int i = 0; i < 129; i += 8)
                rbf_tab_err = std::fmax(rbf_tab_err, std::fabs(s_r.Tabla_Coeficientes()[((size_t)j * 129 + i) * 8 + 3] -
                                                               rbf->Interpola(Vec2(i / 128.0, j / 128.0))));
        /* refusals: sigma not s I; a numeric F; a reflecting side; F = 0 makes the plugin CV plain MC */
        BVP an = nv;
        an.sigma = Coeficiente_Matricial_Numerico(tab([](Vec2 X) { return VarCoef_s(X); }), nullptr,
                                                  tab([](Vec2 X) { return 0.1 * X(0); }),
                                                  tab([](Vec2 X) { return VarCoef_s(X); }));
        FeynmanKacSolver s_bad;
        int bad_sigma = s_bad.Solve_OMP_Analytic(X0, 100, 1e-3, 0.0, an);
        BVP nf = nv;
        nf.F = Coeficiente_Vectorial_Numerico(tab([](Vec2 X) { return X(0); }), nullptr);
        int bad_F = s_bad.Solve_OMP_Analytic(X0, 100, 1e-3, 0.0, nf);
        BVP rf = nv;
        rf.frontera_dominio.Inicializa({0.0, 0.0, 1.0, 1.0}, Rectangulo, left_reflects, Stopping);
        int bad_refl = s_bad.Solve_OMP_Analytic(X0, 100, 1e-3, 0.08, rf);
        BVP z = nv;
        z.F = Coeficiente_Vectorial();
        FeynmanKacSolver s_z; /* default: the plugin control variate */
        int r5 = s_z.Solve_OMP_Analytic(X0, 1000, 1e-3, 0.0, z);
        std::printf("{\"rc\": [%d, %d, %d, %d, %d], \"rbf_rc\": %d, \"mask\": %d, \"same_table\": %d, "
                    "\"compiled\": {\"phi\": %.17g, \"phiphi\": %.17g, \"steps\": %lld}, "
                    "\"numeric\": {\"phi\": %.17g, \"phiphi\": %.17g, \"phi_VR\": %.17g, \"steps\": %lld, "
                    "\"second_phi\": %.17g}, \"rbf\": {\"phi\": %.17g, \"tab_err\": %.3e}, \"u\": %.17g, "
                    "\"bad_sigma\": %d, \"bad_F\": %d, \"bad_refl\": %d, \"F_zero\": {\"phi\": %.17g, "
                    "\"phi_VR\": %.17g, \"xixi\": %.17g}}\n",
                    r1, r2, r3, r4, r5, rbf_rc, s_n.Mascara_Coeficientes(), same_table ? 1 : 0, s_c.phi, s_c.phiphi,
                    (long long)s_c.path_steps, n_phi, n_phiphi, n_phi_VR, n_steps, s_n.phi, s_r.phi, rbf_tab_err, Franke_u(X0), bad_sigma, bad_F, bad_refl, s_z.phi, s_z.phi_VR,
                    s_z.xixi);
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
