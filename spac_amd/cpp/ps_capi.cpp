/*
 * ps_capi.cpp — spac_ps_assemble_rows: G/B rows of nodes inside full interior circles from the
 * pseudospectral solver, in the row layout of spac_fk_assemble_rows (host code).
 *
 * The reference hands every non-perimeter circle to PseudoespectralCirculoSolver
 * (main_test_serial.cpp:67-79, main_test_mpi.cpp:216-224): Inicializa with the circle's node
 * count and first node angle, then Resuelve for its interior nodes.  Circles of one mesh share
 * radius, node count and first angle, so where c is the same on their grids the operator is
 * the same: it is factored once and its cardinal solutions are reused.
 */
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/spac_fk.h"
#include "BVP.hpp"
#include "Nudo.hpp"
#include "PseudoespectralCirculoSolver.hpp"
#include "spac_error.h"

using namespace spac;

extern "C" spac_status spac_ps_assemble_rows(int32_t coef_kind, const spac_interface_table* t,
                                             const spac_node_batch* nb, int32_t n_radio, spac_row_block* out,
                                             spac_ps_report* report) {
    const auto t0 = std::chrono::steady_clock::now();
    if (!t || !nb || !out || !out->B || !out->G || !out->G_col || !out->G_offset)
        return set_error(SPAC_E_INVALID_ARG, "null argument");
    if (n_radio < 3 || n_radio % 2 != 1) return set_error(SPAC_E_INVALID_ARG, "n_radio must be odd and >= 3");
    if (t->n_interfaces < 1 || !t->cx || !t->cy || !t->radius || !t->code || !t->node_offset || !t->node_theta ||
        !t->node_global)
        return set_error(SPAC_E_INVALID_ARG, "interface table is incomplete");
    if (nb->n < 1 || !nb->x || !nb->y || !nb->iface || !nb->frontera)
        return set_error(SPAC_E_INVALID_ARG, "node batch is incomplete");
    CoefSetEntry cs;
    if (!coefficient_set_of_kind(coef_kind, &cs))
        return set_error(SPAC_E_UNREGISTERED_PLUGIN, "coefficient set %d is not registered", coef_kind);
    BVP bvp;
    bvp.u.Inicializa(cs.u);
    bvp.g.Inicializa(cs.g);
    bvp.c.Inicializa(cs.c);
    bvp.f.Inicializa(cs.f);
    bvp.b.Inicializa(cs.b);
    bvp.sigma.Inicializa(cs.sigma);
    const int n = nb->n;
    for (int k = 0; k < n; ++k) {
        const int f = nb->iface[k];
        if (f < 0 || f >= t->n_interfaces) return set_error(SPAC_E_INVALID_ARG, "node %d: interface %d out of range", k, f);
        if (nb->frontera[k]) continue;
        if (t->code[f] != 1)
            return set_error(SPAC_E_BAD_INTERFACE,
                             "node %d: interface %d is not a full interior circle (the pseudospectral solver serves "
                             "those only)", k, f);
        const int M = t->node_offset[f + 1] - t->node_offset[f];
        if (M < 2 || M % 2)
            return set_error(SPAC_E_INVALID_ARG, "interface %d has %d nodes: the pseudospectral grid needs an even count",
                             f, M);
    }
    /* the row layout of spac_fk_assemble_rows: each node's G over its interface's nodes */
    out->G_offset[0] = 0;
    for (int k = 0; k < n; ++k) {
        const int f = nb->iface[k];
        out->G_offset[k + 1] = out->G_offset[k] + (t->node_offset[f + 1] - t->node_offset[f]);
    }
    std::map<int, std::vector<int>> by_iface; /* live nodes of each interface, in batch order */
    for (int k = 0; k < n; ++k) {
        const int f = nb->iface[k];
        const int64_t o = out->G_offset[k];
        const int M = t->node_offset[f + 1] - t->node_offset[f];
        for (int q = 0; q < M; ++q) {
            out->G_col[o + q] = t->node_global[t->node_offset[f] + q];
            out->G[o + q] = 0.0;
        }
        if (out->stats) {
            std::memset(&out->stats[k], 0, sizeof(spac_point_stats));
            out->stats[k].status = SPAC_OK;
        }
        if (nb->frontera[k]) {
            out->B[k] = bvp.g.Evalua(Vec2(nb->x[k], nb->y[k])); /* Solvers/FeynmanKacSolver.cpp:346-347 */
            continue;
        }
        by_iface[f].push_back(k);
    }
    spac_ps_report rep;
    std::memset(&rep, 0, sizeof rep);
    PseudoespectralCirculoSolver ps;
    bool have = false;
    double key_R = 0.0, key_t0 = 0.0;
    int key_M = 0;
    std::vector<double> key_angles;
    std::vector<std::vector<double>> cardinal; /* cached cardinal solutions, one per circle node */
    for (const auto& kv : by_iface) {
        const int f = kv.first;
        const int o = t->node_offset[f], M = t->node_offset[f + 1] - o;
        const Vec2 centro(t->cx[f], t->cy[f]);
        const double R = t->radius[f], th0 = t->node_theta[o];
        std::vector<double> angles(t->node_theta + o, t->node_theta + o + M);
        /* the operator L = Laplacian + c assumes sigma = sqrt(2) I and b = 0 on the circle */
        for (int q = 0; q < 64; ++q) {
            const double a = 2.0 * M_PI * q / 64, s = (q % 8 + 1) / 8.0;
            const Vec2 x(centro(0) + s * R * std::cos(a), centro(1) + s * R * std::sin(a));
            const Mat2 sg = bvp.sigma.Evalua(x);
            const Vec2 bx = bvp.b.Evalua(x);
            if (sg(0, 1) != 0.0 || sg(1, 0) != 0.0 || std::fabs(0.5 * sg(0, 0) * sg(0, 0) - 1.0) > 1e-9 ||
                std::fabs(0.5 * sg(1, 1) * sg(1, 1) - 1.0) > 1e-9 || bx(0) != 0.0 || bx(1) != 0.0)
                return set_error(SPAC_E_INVALID_ARG,
                                 "coefficient set %d: the pseudospectral solver needs sigma = sqrt(2) I and b = 0 "
                                 "(its operator is the Laplacian plus c)", coef_kind);
        }
        bool reuse = have && R == key_R && M == key_M && th0 == key_t0 && ps.Reubica(centro, bvp);
        if (!reuse) {
            if (ps.Inicializa(centro, R, (unsigned)M, (unsigned)n_radio, th0, bvp))
                return set_error(SPAC_E_INVALID_ARG, "interface %d: the pseudospectral operator is singular", f);
            have = true;
            key_R = R;
            key_M = M;
            key_t0 = th0;
            cardinal.clear();
            rep.factorizations++;
        }
        if (!reuse || angles != key_angles) {
            cardinal.assign(M, std::vector<double>());
            const int nh = M / 2;
            for (int j = 0; j < M; ++j) {
                std::vector<double> z(ps.size(), 0.0);
                for (int l = 0; l < M; ++l) z[l] = bvp.fcardinal_Fourier(ps.interpolador.t[l], angles[j], nh);
                ps.solve(z);
                cardinal[j] = std::move(z);
            }
            key_angles = angles;
            rep.solves += M;
        }
        /* B: zero boundary data, source f (Resuelve, :213-226) */
        std::vector<double> zb(ps.size(), 0.0);
        for (int q = M; q < ps.size(); ++q) zb[q] = -bvp.f.Evalua(ps.punto(q));
        ps.solve(zb);
        rep.solves += 1;
        ps.interpolador.z = zb;
        for (int k : kv.second) out->B[k] = ps.interpolador.Interpola(Vec2(nb->x[k], nb->y[k]));
        /* G: minus the cardinal solutions at the nodes (:240-258) */
        for (int j = 0; j < M; ++j) {
            ps.interpolador.z = cardinal[j];
            for (int k : kv.second)
                out->G[out->G_offset[k] + j] = -ps.interpolador.Interpola(Vec2(nb->x[k], nb->y[k]));
        }
        rep.circles++;
        rep.rows += (int32_t)kv.second.size();
    }
    for (int k = 0; k < n; ++k)
        if (!std::isfinite(out->B[k])) return set_error(SPAC_E_NAN, "node %d: the interpolated B is not finite", k);
    rep.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (report) *report = rep;
    return SPAC_OK;
}
