/*
 * spac_mlp.h — C ABI of the learned control variate.
 *
 * The reference's control variate is the plugin F = -sigma^T grad u in the
 * step xi += Y F(X).dW (Solvers/FeynmanKacSolver.cpp:13, BVP/BVP.hpp:17-28,
 * BVPs/Franke_Poisson.hpp:67-73); it needs grad u, which only a manufactured
 * problem knows.  Here an MLP 2 -> 64 -> 64 -> 64 -> 3 (tanh; outputs u_hat,
 * d u_hat/dx, d u_hat/dy) is trained on the GPU from plain Monte-Carlo point
 * estimates, tabulated on a T x T grid with BF16 MFMA inference, and attached
 * to a problem as its F.  Any adapted F keeps the estimator unbiased; a
 * per-solve variance-ratio test falls back to plain MC where it does not pay.
 *
 * spac_mlp_* entry points are the "spac_mlp_* (train, infer, load, save)"
 * functions of the proposed boundary; parameters load and save through
 * spac_mlp_get_params / spac_mlp_set_params as 8707 floats.
 */
#ifndef SPAC_MLP_H
#define SPAC_MLP_H

#include "spac_fk.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SPAC_MLP_N_PARAMS 8707 /* [W1 64x2, b1, W2 64x64, b2, W3 64x64, b3, W4 3x64, b4], W row-major [out][in] */

typedef struct spac_mlp spac_mlp;

typedef struct {
    double domain[4];         /* xmin, ymin, xmax, ymax: inputs are scaled to [-1, 1]^2 over it */
    uint64_t seed;            /* of the initialisation: Glorot-uniform weights, zero biases */
    double first_layer_scale; /* > 0: the first layer's weights and biases uniform in (-s, s)
                                 instead (tanh features of varied slope and offset across the
                                 domain from the first step); 0: Glorot, zero biases */
} spac_mlp_desc;

/* Adam on the full batch (beta1 0.9, beta2 as given, 0: 0.999); learning rate
 * decays geometrically from lr to lr_final (lr_final <= 0: constant).  Loss
 *   L = mean_k (o0_k - u_k)^2 + grad_weight ((o1_k - gx_k)^2 + (o2_k - gy_k)^2)
 * with (gx, gy) the forward-mode gradient of o0, held fixed per step. */
typedef struct {
    int32_t steps;
    int32_t eager;   /* 0: steps replayed from captured HIP graphs; 1: launched one by one (same results) */
    double lr;
    double lr_final;
    double grad_weight;
    double beta2;    /* Adam second-moment decay in (0, 1); 0 selects 0.999 */
} spac_mlp_train_cfg;

typedef struct {
    double final_loss;
    double train_ms; /* device time of the training steps (HIP events) */
    int32_t steps;
    int32_t reserved;
} spac_mlp_train_report;

spac_status spac_mlp_create(const spac_mlp_desc* desc, spac_mlp** out);
spac_status spac_mlp_destroy(spac_mlp* m);
spac_status spac_mlp_get_params(spac_mlp* m, float* params /* SPAC_MLP_N_PARAMS */);
/* Also resets the Adam moments. */
spac_status spac_mlp_set_params(spac_mlp* m, const float* params);

/* Loss and its gradient w.r.t. the parameters over n samples (FP32 kernels).
 * grad may be NULL (loss only). */
spac_status spac_mlp_loss_grad(spac_mlp* m, const double* x, const double* y, const double* u, int32_t n,
                               double grad_weight, float* grad, double* loss);
spac_status spac_mlp_train(spac_mlp* m, const double* x, const double* y, const double* u, int32_t n,
                           const spac_mlp_train_cfg* cfg, spac_mlp_train_report* report);

/* FP32 forward of the training kernels: out3 = [3][n] (u, gx, gy heads). */
spac_status spac_mlp_forward_fp32(spac_mlp* m, const double* x, const double* y, int32_t n, float* out3);
/* BF16-MFMA inference (hidden layers on v_mfma_f32_32x32x8bf16_1k, FP32
 * accumulate): out4 = [n][4] = (u, gx, gy, 0). */
spac_status spac_mlp_infer(spac_mlp* m, const double* x, const double* y, int32_t n, float* out4);

/* Tabulate the heads on a t x t grid over the domain with the MFMA kernel;
 * entry (i, j) at x = xmin + i (xmax - xmin) / (t - 1), y likewise, index j t + i. */
spac_status spac_mlp_build_table(spac_mlp* m, int32_t t, double* infer_ms);
spac_status spac_mlp_read_table(spac_mlp* m, float* out4 /* t*t*4 */);

/* Attach the table as problem p's control variate (mode SPAC_CV_TABLE; the
 * table is copied, m may be destroyed afterwards). */
spac_status spac_mlp_attach(spac_mlp* m, spac_problem* p, int32_t fallback, double max_ratio);

/* The network's u head at every node of a device-resident system (spac_mesh_system_create),
 * evaluated on the device by the BF16-MFMA inference: u[i] = u_hat(x_i, y_i) for the system's
 * n_rows nodes.  The learned method's initial guess for the interface solve
 * (spac_gmres_solve_system): the reference's GMRES starts from zero
 * (Solvers/SistemaLinealSolverGMRES_OMP.hpp:519-577); a guess near the solution only changes
 * how many iterations reach the same tolerance. */
spac_status spac_mlp_predict_system(spac_mlp* m, const spac_system* sys, double* u);

/* Attach the network itself (mode SPAC_CV_NETWORK): F = -sigma grad u_hat with the gradient heads
 * evaluated in FP32 at every step of every path, instead of read from the table.  The per-step
 * form of the learned control variate (Solvers/FeynmanKacSolver.cpp:13 with a learned F), kept
 * to measure what the table changes: variance and throughput against spac_mlp_attach.  Needs a
 * table (spac_mlp_build_table: the exit scores and the fallback read it), an FP32 problem of
 * the compiled variable-coefficient set, absorbing sides and no coefficient tables.  The
 * parameters and table are copied; m may be destroyed afterwards. */
spac_status spac_mlp_attach_network(spac_mlp* m, spac_problem* p, int32_t fallback, double max_ratio);

#ifdef __cplusplus
}
#endif
#endif /* SPAC_MLP_H */
