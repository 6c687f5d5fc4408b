/*
 * sanitize_driver.cpp — runs the host code (mesh and interface tables, CSR pattern,
 * Actualiza_GB, host GMRES, text and row-file formats, the pseudospectral rows, the RBF
 * interpolator) and the CPU restatement of the path solver under AddressSanitizer and UndefinedBehaviorSanitizer.  Built with g++/gcc
 * only (no HIP): tests/test_sanitize.py compiles and runs it; any report aborts it.
 *
 * It follows the pipeline of the reference's serial driver (main_test_serial.cpp:11-83):
 * build the circle mesh, assemble every row (here by the restatement, few paths), update
 * the system, solve it and write the output files.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "Interpolador_RBF.hpp"
#include "spac_fk.h"
#include "spac_mesh.h"

extern "C" {
int oracle_assemble_rows(const spac_problem_desc* desc, const spac_interface_table* t, const spac_node_batch* nb,
                         uint64_t seed, int nthreads, int direct, double* B, double* G, spac_point_stats* stats);
int oracle_simulate(const spac_problem_desc* desc, const spac_solve_batch* batch, uint64_t seed,
                    spac_path_records* rec, int nthreads);
}

#define CHECK(cond, ...)                                                   \
    do {                                                                   \
        if (!(cond)) {                                                     \
            std::fprintf(stderr, "check failed at %d: ", __LINE__);        \
            std::fprintf(stderr, __VA_ARGS__);                             \
            std::fprintf(stderr, " (%s)\n", spac_last_error());            \
            std::exit(2);                                                  \
        }                                                                  \
    } while (0)

static spac_problem_desc varcoef_desc() {
    spac_problem_desc d;
    std::memset(&d, 0, sizeof d);
    d.coef_kind = SPAC_COEF_VARCOEF_FRANKE;
    d.domain_kind = SPAC_BOUNDARY_RECTANGLE;
    d.domain_params[0] = 0.0; d.domain_params[1] = 0.0; d.domain_params[2] = 1.0; d.domain_params[3] = 1.0;
    d.subdomain_kind = SPAC_BOUNDARY_NONE;
    d.precision = SPAC_FP64;
    d.gobet_constant = -0.5826;
    d.max_steps = 1000000;
    return d;
}

static void run_mesh(int n_side, int nt, int rbf_kind, const std::string& dir, int paths, int threads) {
    spac_mesh_desc md;
    std::memset(&md, 0, sizeof md);
    md.domain[2] = 1.0; md.domain[3] = 1.0;
    md.n_side = n_side;
    md.nodes_per_circle = nt;
    md.overlap = 1.2;
    md.theta0 = M_PI / nt;
    md.c2_corner = 1.5;
    md.c2_side = 1.1;
    md.rbf_kind = rbf_kind;
    md.stencil_size = 15;
    md.gauss_eps = 3.0;
    md.fk_all_circles = rbf_kind == SPAC_RBF_GAUSS_STENCIL ? 1 : 0;
    spac_mesh* m = nullptr;
    CHECK(spac_mesh_build_square(&md, &m) == SPAC_OK, "mesh %d x %d", n_side, nt);
    spac_mesh_info info;
    CHECK(spac_mesh_get_info(m, &info) == SPAC_OK, "info");
    spac_interface_table t;
    CHECK(spac_mesh_interfaces(m, &t) == SPAC_OK, "interfaces");
    const int nj = info.n_jobs, n = info.n_nodes;
    std::vector<int32_t> row(nj), iface(nj), front(nj);
    std::vector<double> x(nj), y(nj);
    CHECK(spac_mesh_jobs(m, row.data(), x.data(), y.data(), iface.data(), front.data()) == SPAC_OK, "jobs");
    std::vector<double> nx(n), ny(n);
    std::vector<int32_t> nf(n);
    CHECK(spac_mesh_nodes(m, nx.data(), ny.data(), nf.data()) == SPAC_OK, "nodes");
    if (rbf_kind == SPAC_RBF_IMQ_GLOBAL) {
        for (int k = 0; k < t.n_interfaces; ++k) {
            double c = 0.0;
            if (t.code[k] != 1) CHECK(spac_mesh_interface_cond(m, k, &c) == SPAC_OK && c > 1.0, "cond %d", k);
        }
    }
    /* every row by the restatement (rows of IMQ meshes: the arcs only) */
    std::vector<uint32_t> pth(nj, (uint32_t)paths), strm(nj);
    std::vector<double> hh(nj, 5e-4);
    for (int k = 0; k < nj; ++k) strm[k] = (uint32_t)k;
    spac_node_batch nb;
    std::memset(&nb, 0, sizeof nb);
    nb.n = nj;
    nb.x = x.data(); nb.y = y.data(); nb.iface = iface.data(); nb.frontera = front.data();
    nb.paths = pth.data(); nb.h = hh.data(); nb.stream = strm.data();
    std::vector<int64_t> goff(nj + 1, 0);
    for (int k = 0; k < nj; ++k) goff[k + 1] = goff[k] + (t.node_offset[iface[k] + 1] - t.node_offset[iface[k]]);
    CHECK(goff[nj] == info.total_G, "G entries %lld vs %lld", (long long)goff[nj], (long long)info.total_G);
    std::vector<double> B(nj), G(goff[nj]);
    std::vector<int32_t> col(goff[nj]);
    for (int k = 0; k < nj; ++k)
        for (int64_t q = goff[k]; q < goff[k + 1]; ++q) col[q] = t.node_global[t.node_offset[iface[k]] + (q - goff[k])];
    std::vector<spac_point_stats> st(nj);
    spac_problem_desc d = varcoef_desc();
    CHECK(oracle_assemble_rows(&d, &t, &nb, 5, threads, 0, B.data(), G.data(), st.data()) == 0, "oracle rows");
    for (int k = 0; k < nj; ++k) CHECK(st[k].status == 0 && std::isfinite(B[k]), "row %d", k);
    spac_row_block rb;
    rb.B = B.data(); rb.G = G.data(); rb.G_col = col.data(); rb.G_offset = goff.data(); rb.stats = st.data();
    CHECK(spac_mesh_reset_system(m) == SPAC_OK, "reset");
    CHECK(spac_mesh_accumulate(m, &rb) == SPAC_OK, "accumulate");
    std::vector<int32_t> Gi(n + 1), Gj(info.nnz);
    std::vector<double> Gv(info.nnz), Bs(n), u(n, 0.0);
    CHECK(spac_mesh_csr(m, Gi.data(), Gj.data(), Gv.data(), Bs.data()) == SPAC_OK, "csr");
    spac_gmres_info gi;
    CHECK(spac_gmres_solve(n, Gi.data(), Gj.data(), Gv.data(), Bs.data(), u.data(), 1e-10, 5000, 100, &gi) == SPAC_OK &&
              gi.converged,
          "gmres (residual %.3e)", gi.residual);
    double err = 0.0;
    for (int i = 0; i < n; ++i) err = std::fmax(err, std::fabs(u[i]));
    CHECK(err < 10.0, "solution bounded");
    /* the reference's output files, and a COO round trip */
    CHECK(spac_mesh_write_coo(m, (dir + "/G.txt").c_str(), (dir + "/B.txt").c_str()) == SPAC_OK, "coo");
    CHECK(spac_mesh_write_csr(m, dir.c_str()) == SPAC_OK, "csr files");
    CHECK(spac_mesh_write_positions(m, SPAC_COEF_VARCOEF_FRANKE, (dir + "/knot_position.txt").c_str()) == SPAC_OK,
          "positions");
    CHECK(spac_write_solution((dir + "/solution_GMRES.txt").c_str(), n, u.data()) == SPAC_OK, "solution");
    const char* phases[2] = {"G_and_B_computation", "Solving_LS"};
    int32_t ranks[2] = {0, 0};
    double secs[2] = {1.5, gi.seconds};
    CHECK(spac_write_times_csv((dir + "/times.csv").c_str(), 2, phases, ranks, secs, 0) == SPAC_OK, "times");
    {
        FILE* fg = std::fopen((dir + "/G.txt").c_str(), "r");
        FILE* fo = std::fopen((dir + "/G_in.txt").c_str(), "w");
        CHECK(fg && fo, "G_in");
        std::fprintf(fo, "%d %d\n", info.nnz, n);
        char buf[256];
        while (std::fgets(buf, sizeof buf, fg)) std::fputs(buf, fo);
        std::fclose(fg);
        std::fclose(fo);
        int32_t rn = 0, rnnz = 0;
        CHECK(spac_read_coo((dir + "/G_in.txt").c_str(), &rn, &rnnz, nullptr, nullptr, nullptr) == SPAC_OK, "coo size");
        std::vector<int32_t> Ri(rn + 1), Rj(rnnz);
        std::vector<double> Rv(rnnz);
        CHECK(spac_read_coo((dir + "/G_in.txt").c_str(), &rn, &rnnz, Ri.data(), Rj.data(), Rv.data()) == SPAC_OK, "coo");
        CHECK(rn == n && rnnz == info.nnz && Ri == Gi && Rj == Gj, "coo round trip");
    }
    /* the rows as one batch file, read back, and a damaged copy refused */
    const std::string rf = dir + "/rows_000000.bin";
    CHECK(spac_rows_file_write(rf.c_str(), 0, 5, 77, nj, &rb) == SPAC_OK, "rows file");
    std::vector<double> B2(nj), G2(goff[nj]);
    std::vector<int32_t> col2(goff[nj]);
    std::vector<int64_t> go2(nj + 1, 0);
    std::vector<spac_point_stats> st2(nj);
    spac_row_block r2;
    r2.B = B2.data(); r2.G = G2.data(); r2.G_col = col2.data(); r2.G_offset = go2.data(); r2.stats = st2.data();
    int32_t found = 0;
    CHECK(spac_rows_file_read(rf.c_str(), 0, 5, 77, nj, goff[nj], &r2, &found) == SPAC_OK && found == 1, "read");
    CHECK(B2 == B && G2 == G && col2 == col && go2 == goff, "rows round trip");
    CHECK(std::memcmp(st2.data(), st.data(), st.size() * sizeof(spac_point_stats)) == 0, "stats round trip");
    FILE* fp = std::fopen(rf.c_str(), "r+b");
    CHECK(fp && std::fseek(fp, 100, SEEK_SET) == 0 && std::fputc(0x5a, fp) != EOF, "damage");
    std::fclose(fp);
    CHECK(spac_rows_file_read(rf.c_str(), 0, 5, 77, nj, goff[nj], &r2, &found) == SPAC_OK && found == -1, "damaged");
    CHECK(spac_mesh_destroy(m) == SPAC_OK, "destroy");
    std::printf("mesh %dx%d nt %d rbf %d: %d rows, %d nodes, nnz %d, GMRES %d iterations, residual %.2e\n", n_side,
                n_side, nt, rbf_kind, nj, n, info.nnz, gi.iterations, gi.residual);
}

static void run_points(int threads) {
    /* the Franke point workload: 16 grid points, 200 paths, every record and estimator */
    spac_problem_desc d = varcoef_desc();
    d.coef_kind = SPAC_COEF_FRANKE_POISSON;
    d.domain_params[0] = -1.0; d.domain_params[1] = -1.0;
    const int n = 16, paths = 200;
    std::vector<double> x0(n), y0(n), h(n, 1e-3);
    std::vector<uint32_t> p(n, paths), s(n);
    for (int i = 0; i < n; ++i) {
        x0[i] = -0.8 + 0.1 * i;
        y0[i] = 0.7 - 0.09 * i;
        s[i] = (uint32_t)i;
    }
    spac_solve_batch b;
    std::memset(&b, 0, sizeof b);
    b.n = n;
    b.x0 = x0.data(); b.y0 = y0.data(); b.paths = p.data(); b.h = h.data(); b.stream = s.data();
    const size_t tp = (size_t)n * paths;
    std::vector<double> xt(tp), yt(tp), Y(tp), Z(tp), xi(tp), tau(tp), gv(tp);
    std::vector<int32_t> steps(tp);
    std::vector<int8_t> out(tp);
    spac_path_records r;
    r.x_tau = xt.data(); r.y_tau = yt.data(); r.Y = Y.data(); r.Z = Z.data(); r.xi = xi.data(); r.tau = tau.data();
    r.gval = gv.data(); r.steps = steps.data(); r.outcome = out.data();
    CHECK(oracle_simulate(&d, &b, 3, &r, threads) == 0, "simulate");
    long long total = 0;
    for (size_t i = 0; i < tp; ++i) {
        CHECK(out[i] == SPAC_OUT_DOMAIN && steps[i] >= 1, "path %zu", i);
        total += steps[i];
    }
    std::printf("points: %zu paths, %lld path-steps\n", tp, total);
}

static void run_pseudospectral_and_rbf() {
    /* pseudospectral rows of every interior-circle node of a 4 x 4 mesh (44 nodes per circle),
       twice (the second call reuses the factorisation), and an RBF fit of scattered data */
    spac_mesh_desc md;
    std::memset(&md, 0, sizeof md);
    md.domain[2] = 1.25; md.domain[3] = 1.25;
    md.n_side = 4; md.overlap = 1.0; md.theta0 = 0.1428; md.nodes_per_circle = 44;
    md.c2_corner = 1.5; md.c2_side = 1.1; md.rbf_kind = SPAC_RBF_GAUSS_STENCIL;
    md.gauss_eps = 3.0; md.stencil_size = 15; md.fk_all_circles = 1;
    spac_mesh* m = nullptr;
    CHECK(spac_mesh_build_square(&md, &m) == SPAC_OK, "ps mesh");
    spac_mesh_info info;
    CHECK(spac_mesh_get_info(m, &info) == SPAC_OK, "ps info");
    spac_interface_table t;
    CHECK(spac_mesh_interfaces(m, &t) == SPAC_OK, "ps interfaces");
    const int nj = info.n_jobs;
    std::vector<int32_t> row(nj), iface(nj), front(nj);
    std::vector<double> x(nj), y(nj);
    CHECK(spac_mesh_jobs(m, row.data(), x.data(), y.data(), iface.data(), front.data()) == SPAC_OK, "ps jobs");
    std::vector<double> jx, jy;
    std::vector<int32_t> ji, jf;
    int64_t total = 0;
    for (int k = 0; k < nj; ++k)
        if (t.code[iface[k]] == 1 && !front[k]) {
            jx.push_back(x[k]); jy.push_back(y[k]); ji.push_back(iface[k]); jf.push_back(0);
            total += t.node_offset[iface[k] + 1] - t.node_offset[iface[k]];
        }
    const int n = (int)jx.size();
    CHECK(n > 0, "interior-circle jobs");
    std::vector<uint32_t> paths(n, 1), stream(n, 0);
    std::vector<double> hh(n, 1e-3), B(n), G(total);
    std::vector<int32_t> col(total);
    std::vector<int64_t> goff(n + 1);
    spac_node_batch nb;
    std::memset(&nb, 0, sizeof nb);
    nb.n = n; nb.x = jx.data(); nb.y = jy.data(); nb.iface = ji.data(); nb.frontera = jf.data();
    nb.paths = paths.data(); nb.h = hh.data(); nb.stream = stream.data();
    spac_row_block rb;
    std::memset(&rb, 0, sizeof rb);
    rb.B = B.data(); rb.G = G.data(); rb.G_col = col.data(); rb.G_offset = goff.data();
    spac_ps_report rep;
    for (int pass = 0; pass < 2; ++pass)
        CHECK(spac_ps_assemble_rows(SPAC_COEF_FRANKE_POISSON, &t, &nb, 21, &rb, &rep) == SPAC_OK, "ps rows");
    for (int k = 0; k < n; ++k) CHECK(std::isfinite(B[k]), "ps B %d", k);
    CHECK(spac_mesh_destroy(m) == SPAC_OK, "ps destroy");
    std::vector<spac::Vec2> cs;
    std::vector<double> z;
    for (int k = 0; k < 200; ++k) {
        cs.emplace_back(std::fmod(0.618034 * k, 1.0), std::fmod(0.414214 * k + 0.2, 1.0));
        z.push_back(std::sin(3.0 * cs.back()(0)) * cs.back()(1));
    }
    spac::Interpolador_RBF r;
    CHECK(r.Inicializa(cs, z) == 0, "rbf fit");
    double e = 0.0;
    for (int k = 0; k < 200; ++k) e = std::fmax(e, std::fabs(r.Interpola(cs[k]) - z[k]));
    CHECK(e < 1e-8 && std::isfinite(r.Deriva_t2(spac::Vec2(0.3, 0.4))), "rbf at centres %.2e", e);
    std::printf("pseudospectral: %d rows, %d factorisations; rbf: 200 centres\n", n, rep.factorizations);
}

int main(int argc, char** argv) {
    const std::string dir = argc > 1 ? argv[1] : ".";
    const int threads = argc > 2 ? std::atoi(argv[2]) : 2;
    run_points(threads);
    run_mesh(4, 24, SPAC_RBF_GAUSS_STENCIL, dir, 40, threads);
    run_mesh(4, 24, SPAC_RBF_IMQ_GLOBAL, dir, 40, threads);
    run_mesh(6, 20, SPAC_RBF_GAUSS_STENCIL, dir, 20, threads);
    run_pseudospectral_and_rbf();
    std::printf("sanitized run ok\n");
    return 0;
}
