"""Algorithmic FP64 work per path-step, counted from the formulas the kernels run.

Convention: add/sub/mul/neg = 1, FMA = 2, and each exp/log/sin/cos/sqrt/div
(and comparison) = 1.  Integer Philox work is not counted.  One path-step is
one Paso_Absorbente + Dentro pair (Solvers/FeynmanKacSolver.cpp:171-185).

Franke Poisson, no subdomain (the plain FP64 point workload):
  Box–Muller normals: 1-u, log, *-2, sqrt, 2u, sin, cos, 2 products, 2 scalings  11
  dW = sqrt(h) n                                                                  2
  Franke Gaussians: 7 affine args (2 each) = 14; F1,F3,F4: 2 squares, add,
    scale, exp = 5 each; F2: square, scale, affine, scale, sub, exp = 7           36
  f = -(u_xx + u_yy): u_xx 4 terms x 5 + 3 adds = 23; u_yy 3 x 5 + 1 + 3 = 19;
    negate + subtract = 2                                                        44
  F = -sigma grad u: u_x 4 x 2 + 3 = 11; u_y 3 x 2 + 1 + 3 = 10; 2 x (mul, neg) 25
  updates: Z (2 mul, add) 3, xi (2 mul, add, mul, add) 5, X (2 x mul, add) 4,
    t 1                                                                         13
  Dentro on the rectangle: centre 4, half sides 2, |x - c| 4, branch 2,
    distance 1, N^T sigma 2, norm 3 + sqrt, threshold 2, compare 1                22
                                                                        total  153
Variable-coefficient operator, circle subdomain (the PDDSparse row workload):
  the Franke total above, plus
  sigma at X: sinpi, cospi, 2 mul, add = 5; sigma again in Dentro = 5            10
  drift b = 0.3 (y, -x): 2; X update gains b h: 2 mul + 2 add                      6
  killing c = -1 - x^2: 2; Y update: 2 mul + add                                   5
  f replaces -(u_xx+u_yy) (2) by -((s^2/2 lap + b.grad u) + c u): lap 1,
    s^2/2 * lap 3, b.grad u 3, u 7, c u 1, two adds + negate 3 = 18              +16
  circle distance: 2 sub, 2 mul + add + sqrt, max, 2 div, 2 (mul, add),
    sub = 14; nearest-boundary compare 1                                          15
                                                                        total  205
FP32 stepping (fk_path_kernel_fp32), same operator and circle subdomain:
  the variable-coefficient total above, less t += h per step (tau = steps x h
  at the exit)                                                                    -1
  plain Monte Carlo (no control variate): no F = -sigma grad u (25), no xi
  update (5)                                                                     -30
                                                                 total (plain)   174
  learned control variate (a tabulated grad u_hat, bilinear): no F plugin (25);
    cell (x - lo) * inv for x and y 4, clamps 4, fractions and complements 4,
    two channels x (4 mul + 2 add + 2 mul + add) 18 = 30; xi update
    -s Y (gx dWx + gy dWy) 6 instead of 5                                         +6
                                                                 total (table)   210
"""

FLOPS_PER_STEP = {
    "franke_point_fp64": 153,
    "varcoef_circle_fp64": 205,
    "varcoef_circle_fp32_plain": 174,
    "varcoef_circle_fp32_table": 210,
}

# MI355X published peaks (AMD): FP64 vector 78.6 TFLOP/s, FP32 vector 157.3 TFLOP/s,
# BF16 dense MFMA 2.5 PFLOP/s, HBM3E 8 TB/s.
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_BF16_MFMA_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
