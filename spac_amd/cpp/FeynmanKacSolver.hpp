/*
 * FeynmanKacSolver.hpp — drop-in for the reference's solver class
 * (Solvers/FeynmanKacSolver.hpp:43-99): same public estimator fields and the
 * same two Solve_OMP_Analytic overloads, implemented over the C ABI of
 * include/spac_fk.h, so the paths run on the GPU.
 *
 * Differences a caller can see:
 *   - the BVP and Interfaz are taken by const reference (the reference copies
 *     both per call, :335-355);
 *   - every solve returns a spac_status (0 = OK) instead of printing ERROR;
 *   - RNG: one Philox stream per call (seed = MPI rank, stream = call count),
 *     so a run is reproducible for any thread count;
 *   - xi/xixi are the true moments and RNGC counts normals drawn;
 *   - a default (unset) subdomain frontier means "no subdomain", not a
 *     distance-0 boundary that stops every path after one step;
 *   - a domain Rectangulo whose Absorbente predicate is false on a side makes
 *     that side reflecting (the Lepingle step, Solvers/FeynmanKacSolver.cpp:19-67,
 *     restated without its defects); rho is the band of that step and is
 *     ignored when every side absorbs, the setting of the shipped drivers.
 */
#ifndef SPAC_FEYNMANKACSOLVER_HPP
#define SPAC_FEYNMANKACSOLVER_HPP

#include <cstdint>
#include <vector>

#include "../../include/spac_fk.h"
#include "BVP.hpp"
#include "Interfaz.hpp"
#include "Interpolador.hpp"
#include "Nudo.hpp"

namespace spac {

class FeynmanKacSolver {
public:
    double h = 0.0, sqrth = 0.0;
    unsigned int N = 0;
    double eps = 0.0;
    double phi = 0, phiphi = 0, phi_sublinear = 0, phiphi_sublinear = 0, phi_VR = 0, phiphi_VR = 0,
           phi_sublinearVR = 0, phiphi_sublinearVR = 0, xi = 0, xixi = 0, xi_sublinear = 0, xixi_sublinear = 0,
           xiphi = 0, xiphi_sublinear = 0, phi_num = 0, phiphi_num = 0, phi_sublinear_num = 0,
           phiphi_sublinear_num = 0, phi_VR_num = 0, phiphi_VR_num = 0, phi_sublinearVR_num = 0,
           phiphi_sublinearVR_num = 0, xi_num = 0, xixi_num = 0, xi_sublinear_num = 0, xixi_sublinear_num = 0,
           xiphi_num = 0, xiphi_sublinear_num = 0, APL = 0, RNGC = 0;
    int64_t path_steps = 0;     /* executed Paso_Absorbente + Dentro pairs of the last solve */
    uint32_t max_steps = 0;     /* cap per path (0: none) */

    FeynmanKacSolver();
    explicit FeynmanKacSolver(unsigned int MPIrank);
    ~FeynmanKacSolver();
    FeynmanKacSolver(const FeynmanKacSolver&) = delete;
    FeynmanKacSolver& operator=(const FeynmanKacSolver&) = delete;

    /* Solvers/FeynmanKacSolver.cpp:335-342 */
    int Solve_OMP_Analytic(Vec2 X0, unsigned int numero_trayectorias, double discretizacion_temporal, double rho,
                           const BVP& Problema);
    /* Solvers/FeynmanKacSolver.cpp:344-355: fills nudo.G (over the interface's
     * nodes) and nudo.B */
    int Solve_OMP_Analytic(Nudo& nudo, unsigned int numero_trayectorias, double discretizacion_temporal, double rho,
                           const BVP& Problema, const Interfaz& interfaz);
    /* Control variate F of xi += Y F(X).dW (Solvers/FeynmanKacSolver.cpp:13):
     * the BVP's F plugin (the reference's behaviour, default), none (plain Monte
     * Carlo), or F = -sigma grad u_hat from a tabulated u_hat read on the device;
     * fallback drops xi per node where it adds variance. */
    int Variable_Control_Plugin();
    int Sin_Variable_Control();
    int Variable_Control(const Interpolador_Tabla& u_hat, bool fallback = true, double max_ratio = 1.0);

    /* Batched form: every (node, interface) pair in one GPU launch. */
    int Solve_Nodes(std::vector<Nudo*>& nudos, const std::vector<const Interfaz*>& interfaces,
                    unsigned int numero_trayectorias, double discretizacion_temporal, const BVP& Problema,
                    double rho = 0.0);

    /* Numeric coefficients (Coeficiente_*_Numerico: sigma = s I, b, c, f, g, u) run on the
     * device as tables: each is sampled once, through its Evalua, at the nodes of a
     * T_tabla x T_tabla grid over the domain rectangle, and read bilinearly by the path
     * kernels.  The samples are kept while the BVP holds the same interpolator objects, the
     * same T_tabla and the same rectangle; Retabula_Coeficientes() drops them (after an
     * interpolator was changed in place).  The other coefficients must be those of the
     * registered variable-coefficient set; F may also be left at its default (zero), which
     * makes the plugin control variate xi = 0, i.e. plain Monte Carlo. */
    unsigned int T_tabla = 257;
    void Retabula_Coeficientes() { coef_key_.clear(); coef_tab_.clear(); }
    /* The sampled table (T_tabla^2 x 8 doubles, the layout of spac_coef_tables) and its mask. */
    const std::vector<double>& Tabla_Coeficientes() const { return coef_tab_; }
    int Mascara_Coeficientes() const { return coef_mask_; }

private:
    spac_problem* prob_ = nullptr;
    spac_problem_desc desc_{};
    uint64_t seed_ = 0;
    uint32_t calls_ = 0;
    spac_cv_config cv_{};           /* applied to every problem this solver creates */
    std::vector<float> cv_table_;
    bool F_zero_ = false;           /* the BVP's F is the default zero: the plugin CV is plain MC */
    std::vector<double> coef_tab_;  /* sampled numeric coefficients */
    int coef_mask_ = 0;
    std::vector<InterpoladorPtr> coef_key_; /* held, so no other object can take their addresses */
    unsigned int coef_T_ = 0;
    double coef_dom_[4] = {0.0, 0.0, 0.0, 0.0};
    int apply_cv();
    int apply_tables();
    int sample_tables(const BVP& P, int mask, const double dom[4], bool* changed);
    int ensure_problem(const BVP& P, bool with_subdomain, double rho);
    void copy_stats(const spac_point_stats& s);
};

} // namespace spac

#endif
