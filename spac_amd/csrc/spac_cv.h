/*
 * spac_cv.h — lookup in a tabulated control variate (u_hat, d u_hat/dx,
 * d u_hat/dy on a T x T grid of floats, 4 per node), compiled as C for the CPU
 * restatement and as HIP for the FP64 path kernel so both read the same bits.
 *
 * The table is the device form of the reference's TablaDeValores +
 * Interpolador pair (BVP/Interpolador.hpp:15-19): bilinear interpolation, the
 * nearest edge value outside the grid.  Index (i, j) -> table[(j T + i) 4 + c].
 */
#ifndef SPAC_CV_H
#define SPAC_CV_H

#include "spac_math.h"

typedef struct {
    const float* table;
    int32_t t;
    double lox, loy; /* grid origin */
    double ix, iy;   /* (T - 1) / (xmax - xmin), (T - 1) / (ymax - ymin) */
} spac_cv_table;

/* Cell and weights of (x, y); NaN maps to the origin cell. */
SPAC_HD void spac_cv_cell(const spac_cv_table* c, double x, double y, int32_t* k00, double* ax, double* ay) {
    double fx = (x - c->lox) * c->ix, fy = (y - c->loy) * c->iy;
    double top = (double)(c->t - 1);
    fx = (fx > 0.0) ? fx : 0.0;
    fy = (fy > 0.0) ? fy : 0.0;
    fx = (fx < top) ? fx : top;
    fy = (fy < top) ? fy : top;
    int32_t i = (int32_t)fx, j = (int32_t)fy;
    i = (i < c->t - 2) ? i : c->t - 2;
    j = (j < c->t - 2) ? j : c->t - 2;
    *ax = fx - (double)i;
    *ay = fy - (double)j;
    *k00 = j * c->t + i;
}

/* Component comp (0: u_hat, 1: du/dx, 2: du/dy) at (x, y), in FP64. */
SPAC_HD double spac_cv_eval(const spac_cv_table* c, int32_t k00, double ax, double ay, int comp) {
    const float* p = c->table + (int64_t)k00 * 4 + comp;
    const int64_t row = (int64_t)c->t * 4;
    double v00 = (double)p[0], v10 = (double)p[4], v01 = (double)p[row], v11 = (double)p[row + 4];
    double bx = 1.0 - ax, by = 1.0 - ay;
    return by * (bx * v00 + ax * v10) + ay * (bx * v01 + ax * v11);
}

#endif
