/*
 * spac_tab.h — tabulated coefficients: the device form of the reference's numeric
 * coefficients (Coeficiente_{Escalar,Vectorial,Matricial}_Numerico, whose Evalua is an
 * Interpolador over a TablaDeValores: BVP/Coeficiente_Escalar_Numerico.hpp:6-19,
 * BVP/Interpolador.hpp:7-19).  Compiled as C for the CPU restatement and as HIP for the path
 * kernels, so that both read the same bits.
 *
 * A table holds, at every node of a uniform T x T grid over [xmin, xmax] x [ymin, ymax],
 * eight doubles: s (sigma = s I), b_x, b_y, c, f, g, u, 0 — node (i, j) at v[(j T + i) 8].
 * Values between nodes are bilinear, outside the grid those of the nearest edge.  A mask
 * says which coefficients the table replaces (SPAC_TAB_* of spac_fk.h); the others stay the
 * compiled plugin's.
 */
#ifndef SPAC_TAB_H
#define SPAC_TAB_H

#include "spac_math.h"

#define SPAC_TAB_STRIDE 8
enum { SPAC_TABC_S = 0, SPAC_TABC_BX = 1, SPAC_TABC_BY = 2, SPAC_TABC_C = 3, SPAC_TABC_F = 4, SPAC_TABC_G = 5,
       SPAC_TABC_U = 6 };

typedef struct {
    const double* v;  /* T * T * 8 doubles */
    int32_t t;
    int32_t mask;     /* SPAC_TAB_* bits */
    double lox, loy;  /* grid origin */
    double ix, iy;    /* (T - 1) / (xmax - xmin), (T - 1) / (ymax - ymin) */
    int32_t need_F;   /* the control variate reads the compiled F (analytic mode) */
    int32_t reserved;
} spac_coef_table;

/* Cell and weights of (x, y); NaN maps to the origin cell. */
SPAC_HD void spac_tab_cell(const spac_coef_table* c, double x, double y, int32_t* k00, double* ax, double* ay) {
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

/* Coefficient comp (SPAC_TABC_*) at the cell's point, in FP64. */
SPAC_HD double spac_tab_eval(const spac_coef_table* c, int32_t k00, double ax, double ay, int comp) {
    const double* p = c->v + (int64_t)k00 * SPAC_TAB_STRIDE + comp;
    const int64_t row = (int64_t)c->t * SPAC_TAB_STRIDE;
    double v00 = p[0], v10 = p[SPAC_TAB_STRIDE], v01 = p[row], v11 = p[row + SPAC_TAB_STRIDE];
    double bx = 1.0 - ax, by = 1.0 - ay;
    return by * (bx * v00 + ax * v10) + ay * (bx * v01 + ax * v11);
}

SPAC_HD double spac_tab_at(const spac_coef_table* c, double x, double y, int comp) {
    int32_t k00;
    double ax, ay;
    spac_tab_cell(c, x, y, &k00, &ax, &ay);
    return spac_tab_eval(c, k00, ax, ay, comp);
}

#endif
