/*
 * dense_solve.h -- direct solve of a small assembled system, the 6-argument
 * solver(P, M, FS, FF, B, X) that the time-integration drivers call and the
 * reference tree leaves unimplemented.  TEST INFRASTRUCTURE ONLY: used by
 * the oracle and by the reference harnesses, never by the product.
 *
 * Gaussian elimination with partial pivoting (the largest |a_kj| of column
 * j at or below row j; ties keep the first row), row operations left to
 * right, back substitution from the last row.  The same code in the oracle
 * and in the harness gives the same bits for the same assembled matrix.
 */
#ifndef ORACLE_DENSE_SOLVE_H
#define ORACLE_DENSE_SOLVE_H

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* A: n*n row-major, overwritten; b: right-hand side in, solution out.
 * Returns 0, or 1 on a zero pivot. */
static inline int orc_dense_solve(int n, double *A, double *b)
{
    for (int j = 0; j < n; j++) {
        int p = j;
        double best = fabs(A[(size_t)j * n + j]);
        for (int k = j + 1; k < n; k++) {
            const double v = fabs(A[(size_t)k * n + j]);
            if (v > best) {
                best = v;
                p = k;
            }
        }
        if (best == 0.0) return 1;
        if (p != j) {
            for (int c = 0; c < n; c++) {
                const double t = A[(size_t)j * n + c];
                A[(size_t)j * n + c] = A[(size_t)p * n + c];
                A[(size_t)p * n + c] = t;
            }
            const double t = b[j];
            b[j] = b[p];
            b[p] = t;
        }
        const double piv = A[(size_t)j * n + j];
        for (int k = j + 1; k < n; k++) {
            const double f = A[(size_t)k * n + j] / piv;
            if (f == 0.0) continue;
            for (int c = j; c < n; c++) A[(size_t)k * n + c] -= f * A[(size_t)j * n + c];
            b[k] -= f * b[j];
        }
    }
    for (int j = n - 1; j >= 0; j--) {
        double s = b[j];
        for (int c = j + 1; c < n; c++) s -= A[(size_t)j * n + c] * b[c];
        b[j] = s / A[(size_t)j * n + j];
    }
    return 0;
}

/* X = (D A)^{-1} B for an element-by-element operator: A[e*k*k + a*k + b]
 * of element e with equation numbers lm[e*k + .] (neq = prescribed, not
 * assembled), D = diag(dscale) or the identity (dscale == NULL).
 * X[neq] = 0.  Returns 0, or 1 on a zero pivot. */
static inline int orc_dense_solve_ebe(int neq, int nel, int k, const double *A, const int *lm,
                                      const double *dscale, const double *B, double *X)
{
    double *M = (double *)calloc((size_t)neq * neq, sizeof(double));
    if (!M) return 1;
    for (int e = 0; e < nel; e++)
        for (int a = 0; a < k; a++) {
            const int r = lm[(size_t)e * k + a];
            if (r == neq) continue;
            for (int b = 0; b < k; b++) {
                const int c = lm[(size_t)e * k + b];
                if (c == neq) continue;
                M[(size_t)r * neq + c] += A[(size_t)e * k * k + a * k + b];
            }
        }
    if (dscale)
        for (int r = 0; r < neq; r++)
            for (int c = 0; c < neq; c++) M[(size_t)r * neq + c] *= dscale[r];
    memcpy(X, B, sizeof(double) * neq);
    const int rc = orc_dense_solve(neq, M, X);
    X[neq] = 0.0;
    free(M);
    return rc;
}

#endif
