/*
 * fga_core.h -- shared state of the drop-in adapters (struct-agnostic).
 *
 * Every adapter (ss2d, ss3d, tr2d, eu2d) is compiled with its own driver's
 * header, so the reference structs differ between them; what they share is
 * kept here as static helpers over plain pointers: the device context, the
 * error report, and the rule that decides when the element / edge / CSR
 * values are uploaded again.
 *
 * Upload rule.  The drivers rebuild the matrix on the host before every
 * solve (Process.c of the steady drivers; Predictor_Old.c:71-75: Build,
 * scaling, precond_setup, solver), and the products are called by the
 * solver, which runs on the device.  So the values are uploaded
 *   - by the first setup, scaling or solver entry of the adapter after the
 *     previous solve ended (the driver's Build lies in between), and
 *   - by a product symbol called before any values were uploaded, or after
 *     femgpu_values_changed();
 * a product symbol called again before the next solve reuses the device
 * values.  A driver that rebuilds the matrix and then calls a product symbol
 * with no setup, scaling or solver call in between calls
 * femgpu_values_changed() after its Build (no reference driver does).
 * Scalings and block scalings rewrite the device values in place: the host
 * matrix keeps the unscaled values, and every routine that reads the scaled
 * matrix is exported by the adapter.
 */
#ifndef FGA_CORE_H
#define FGA_CORE_H

#include <stdio.h>
#include <string.h>

#include "femgpu.h"

static fg_ctx *fga_ctx;
static int fga_built;      /* device values hold the host's current build   */
static int fga_have;       /* any values uploaded since the operator exists  */

static int fga_fail(const char *where)
{
    printf("femgpu %s: %s\n", where, fg_last_error());
    return 1;
}

static int fga_ensure_ctx(void)
{
    if (fga_ctx) return 0;
    if (fg_ctx_create(0, &fga_ctx) != FG_OK) return fga_fail("context");
    return 0;
}

/* a setup, scaling or solver entry: does it have to upload? */
static int fga_needs_build_upload(void) { return !fga_built; }
/* a product symbol: does it have to upload? */
static int fga_needs_product_upload(void) { return !fga_have; }
static void fga_uploaded(void) { fga_built = 1; fga_have = 1; }
/* a solve ended: the next setup / scaling / solver entry follows a Build */
static void fga_solve_done(void) { fga_built = 0; }
/* a new operator: nothing uploaded yet */
static void fga_new_operator(void) { fga_built = 0; fga_have = 0; }

/* no reference counterpart: for a driver that rebuilds the matrix and calls
 * a product symbol with no setup, scaling or solver call in between */
void femgpu_values_changed(void)
{
    fga_built = 0;
    fga_have = 0;
}

/* host <-> device copies of neq+1 doubles into a cached pair of vectors */
static double *fga_dv[2];
static int fga_dv_n;
static int fga_vectors(int n)
{
    if (fga_dv_n >= n && fga_dv[0]) return 0;
    for (int i = 0; i < 2; i++) {
        if (fga_dv[i]) fg_free(fga_ctx, fga_dv[i]);
        fga_dv[i] = NULL;
    }
    fga_dv_n = 0;
    if (fg_malloc(fga_ctx, sizeof(double) * (size_t)n, (void **)&fga_dv[0]) != FG_OK ||
        fg_malloc(fga_ctx, sizeof(double) * (size_t)n, (void **)&fga_dv[1]) != FG_OK)
        return fga_fail("vectors");
    fga_dv_n = n;
    return 0;
}

/* GMRES options from the driver's parameters */
static void fga_opts(fg_gmres_opts *o, int kmax, int lmax, double tol)
{
    memset(o, 0, sizeof *o);
    o->kmax = kmax;
    o->lmax = lmax;
    o->tol = tol;
    o->precision = FG_GMRES_FP64;
}

/* the device solve of A X = B (B already left-preconditioned by the setup,
 * pgmres.c:49-53 semantics), host B in, host X out (neq+1 entries); the
 * reference zeroes X first (pgmres.c:46), the device solver overwrites it */
static int fga_solve(fg_op *op, fg_prec *pc, int neq, const fg_gmres_opts *o, const double *B,
                     double *X, fg_gmres_info *info)
{
    if (fga_vectors(neq + 1)) return 1;
    size_t bytes = sizeof(double) * (size_t)(neq + 1);
    if (fg_memcpy_h2d(fga_ctx, fga_dv[0], B, bytes) != FG_OK) return fga_fail("B");
    int rc = fg_gmres(fga_ctx, op, pc, fga_dv[0], fga_dv[1], o, info);
    fga_solve_done();
    if (rc != FG_OK && rc != FG_ERR_NOCONV) return fga_fail("gmres");
    if (fg_memcpy_d2h(fga_ctx, X, fga_dv[1], bytes) != FG_OK) return fga_fail("X");
    X[neq] = 0.0;   /* the sink slot reads as zero */
    return 0;
}

/* Q = A P through the device operator, host P and Q (neq+1) */
static int fga_product(fg_op *op, int neq, double *P, double *Q)
{
    if (fga_vectors(neq + 1)) return 1;
    size_t bytes = sizeof(double) * (size_t)(neq + 1);
    if (fg_memcpy_h2d(fga_ctx, fga_dv[0], P, bytes) || fg_op_apply_f64(op, fga_dv[0], fga_dv[1]) ||
        fg_memcpy_d2h(fga_ctx, Q, fga_dv[1], bytes))
        return fga_fail("matvec");
    return 0;
}

/* z = M^{-1} p through a device preconditioner, host p and z (may alias) */
static int fga_precond(fg_prec *pc, int neq, const double *p, double *z)
{
    if (fga_vectors(neq + 1)) return 1;
    size_t bytes = sizeof(double) * (size_t)(neq + 1);
    if (fg_memcpy_h2d(fga_ctx, fga_dv[0], p, bytes) ||
        fg_prec_apply_f64(pc, fga_dv[0], fga_dv[1]) ||
        fg_memcpy_d2h(fga_ctx, z, fga_dv[1], bytes))
        return fga_fail("preconditioner");
    z[neq] = 0.0;
    return 0;
}

/* F (host, neq+1) through a device routine that works in place on F */
static int fga_f_roundtrip(int neq, double *F)
{
    if (fga_vectors(neq + 1)) return 1;
    return fg_memcpy_h2d(fga_ctx, fga_dv[0], F, sizeof(double) * (size_t)(neq + 1)) != FG_OK
               ? fga_fail("F")
               : 0;
}
static int fga_f_back(int neq, double *F)
{
    return fg_memcpy_d2h(fga_ctx, F, fga_dv[0], sizeof(double) * (size_t)(neq + 1)) != FG_OK
               ? fga_fail("F")
               : 0;
}

#endif
