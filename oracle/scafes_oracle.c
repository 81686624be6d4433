/*
 * scafes_oracle.c -- CPU restatement of ScaFES's explicit time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker; see scafes_oracle.h).  Built with
 * -O2 -fopenmp -ffp-contract=off so every FP64 operation rounds exactly as
 * written, in the order the reference's hooks write it.
 */
#include "scafes_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static int g_threads = 0;

void oracle_set_threads(int n) { g_threads = n; }

int oracle_get_threads(void)
{
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

#ifdef _OPENMP
#define NTHR (g_threads > 0 ? g_threads : omp_get_max_threads())
#endif

/* memAll extent along axis p (DataField ctor: memAll = partition +- stencilWidth,
 * Problem.hpp:1495-1514). */
static long ext(const oracle_grid* g, int p)
{
    return p < g->dim ? (long)g->n[p] + 2L * g->hw : 1L;
}

long oracle_nelems(const oracle_grid* g) { return ext(g, 0) * ext(g, 1) * ext(g, 2); }

long oracle_nnodes(const oracle_grid* g) { return (long)g->n[0] * g->n[1] * g->n[2]; }

/* h_p = (x_last - x_first) / (n_p - 1)   (Grid.hpp:547-555). */
double oracle_gridsize(const oracle_grid* g, int p)
{
    return (g->x1[p] - g->x0[p]) / (g->n[p] - 1);
}

/* Fortran scalar index of node (i,j,k) in memAll (Grid.hpp:955-968). */
static long pos(const oracle_grid* g, int i, int j, int k)
{
    long o0 = g->dim > 0 ? g->hw : 0, o1 = g->dim > 1 ? g->hw : 0, o2 = g->dim > 2 ? g->hw : 0;
    return (i + o0) + ext(g, 0) * ((j + o1) + ext(g, 1) * (k + o2));
}

/* Node index in a stencil-0 field (known dfs F, G: no ghost layer). */
static long npos(const oracle_grid* g, int i, int j, int k)
{
    return i + (long)g->n[0] * (j + (long)g->n[1] * k);
}

/* GLOBAL_BORDER if within nLayersAtBorder of the global edge along any
 * active axis, else GLOBAL_INNER (Problem.hpp:3902-3980). */
static int is_border(const oracle_grid* g, int i, int j, int k)
{
    int idx[3] = {i, j, k};
    for (int p = 0; p < g->dim; ++p) {
        if (idx[p] < g->nlayers || idx[p] > g->n[p] - 1 - g->nlayers) return 1;
    }
    return 0;
}

/* Neighbour memory offset along axis p (Grid::neigh, Grid.hpp:885-921; the
 * inner-node stencil never reaches the sentinel branch). */
static long stride(const oracle_grid* g, int p)
{
    return p == 0 ? 1L : (p == 1 ? ext(g, 0) : ext(g, 0) * ext(g, 1));
}

/* coordinates(idxNode) = x_first + h*(i - first)  (Grid.hpp:719-733). */
static void coords(const oracle_grid* g, int i, int j, int k, double* x)
{
    int idx[3] = {i, j, k};
    for (int p = 0; p < g->dim; ++p) x[p] = g->x0[p] + oracle_gridsize(g, p) * idx[p];
}

/*---------------------------------------------------------------------------*/
/* Constant-coefficient heat / Pennes system-test update.
 * Inner: HeatEqnFDM::updateInner (tests/system/HeatEqnFDM/HeatEqnFDM.hpp:204-218)
 *        and PennesBioheatEqnFDM::updateInner (PennesBioheatEqnFDM.hpp:153-168).
 * Border: Dirichlet::updateBorder, vNew = knownDf(G) (Dirichlet/Dirichlet.hpp:124-130).
 * Copy-old-to-new before the hooks: setVectDfDepToDfIndep (Problem.hpp:6226-6256)
 * -> DataField::assignValues (DataField.hpp:1147-1158). */
/* Flux border node (no index 0, some index n_p - 1): the border classes'
 * updateBorder in the reference's order: per axis, the one-sided term where
 * the index is n_p - 1 (outer neighbour replaced by the inner one plus the
 * flux built from G_old), central differences elsewhere; then perfusion and
 * source as on inner nodes. */
static double flux_border_update(const oracle_grid* g, const double* uo, long m, const int* idx,
                                 const long* st, const double* hh, double gold, double fold,
                                 int border_mode, const oracle_heat7_coef* c)
{
    const double u = uo[m];
    double v = u;
    for (int p = 0; p < g->dim; ++p) {
        const double w = uo[m - st[p]];
        if (idx[p] == g->n[p] - 1) {
            const double a2 = 2.0 * oracle_gridsize(g, p) / c->lambda;
            double t = w + w;
            if (border_mode == ORACLE_BORDER_NEUMANN) {
                t = t - a2 * gold;
            } else if (border_mode == ORACLE_BORDER_CAUCHY) {
                t = t - a2 * c->alpha * u;
                t = t + a2 * gold;
            } else {
                t = t - a2 * c->epsilon * c->sigma * u * u * u * u;
                t = t + a2 * gold;
            }
            t = t - 2.0 * u;
            v += c->tau * c->coeff_a * t / (hh[p]);
        } else {
            v += c->tau * c->coeff_a * (w + uo[m + st[p]] - 2.0 * u) / (hh[p]);
        }
    }
    if (c->has_perf) v -= c->tau * c->perf_c * c->perf_w * u;
    if (c->has_source) v += c->tau * c->rhoc_inv * fold;
    return v;
}

void oracle_heat7_step(const oracle_grid* g, const double* uo, double* un,
                       const double* fo, const double* go, const double* gn, int border_mode,
                       double border_value, const oracle_heat7_coef* c)
{
    const long ne = oracle_nelems(g);
    memcpy(un, uo, sizeof(double) * (size_t)ne);
    double hh[3];
    long st[3];
    for (int p = 0; p < 3; ++p) {
        hh[p] = p < g->dim ? oracle_gridsize(g, p) * oracle_gridsize(g, p) : 1.0;
        st[p] = stride(g, p);
    }
    const int n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
#pragma omp parallel for collapse(2) schedule(static) num_threads(NTHR)
    for (int k = 0; k < n2; ++k) {
        for (int j = 0; j < n1; ++j) {
            for (int i = 0; i < n0; ++i) {
                const long m = pos(g, i, j, k);
                if (!is_border(g, i, j, k)) {
                    double v = uo[m];
                    for (int p = 0; p < g->dim; ++p) {
                        v += c->tau * c->coeff_a * (uo[m - st[p]] + uo[m + st[p]] - 2.0 * uo[m]) /
                             (hh[p]);
                    }
                    if (c->has_perf) v -= c->tau * c->perf_c * c->perf_w * uo[m];
                    if (c->has_source && fo) v += c->tau * c->rhoc_inv * fo[npos(g, i, j, k)];
                    un[m] = v;
                } else {
                    if (border_mode == ORACLE_BORDER_FIELD) {
                        un[m] = gn[npos(g, i, j, k)];
                    } else if (border_mode == ORACLE_BORDER_VALUE) {
                        un[m] = border_value;
                    } else if (border_mode >= ORACLE_BORDER_NEUMANN) {
                        const int idx[3] = {i, j, k};
                        int dir = 0;
                        for (int p = 0; p < g->dim; ++p)
                            if (idx[p] == 0) dir = 1;
                        const long q = npos(g, i, j, k);
                        un[m] = dir ? gn[q]
                                    : flux_border_update(g, uo, m, idx, st, hh, go[q],
                                                         (c->has_source && fo) ? fo[q] : 0.0,
                                                         border_mode, c);
                    }
                }
            }
        }
    }
}

/* Known fields of the TimeLinSpaceLin system tests at time t:
 *   heat   evalInner/evalBorder: tests/system/HeatEqnFDM/Dirichlet/TimeLinSpaceLin/TimeLinSpaceLin.hpp:75-123
 *   Pennes evalInner/evalBorder: tests/system/PennesBioheatEqnFDM/Dirichlet/TimeLinSpaceLin/TimeLinSpaceLin.hpp:75-124
 * F, G, Y are stencil-0 fields. */
static void eval_known_timelin(const oracle_grid* g, int known, int border, double t,
                               const oracle_heat7_coef* c, double* F, double* G, double* Y)
{
    const int n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
    (void)c;
#pragma omp parallel for collapse(2) schedule(static) num_threads(NTHR)
    for (int k = 0; k < n2; ++k) {
        for (int j = 0; j < n1; ++j) {
            for (int i = 0; i < n0; ++i) {
                double x[3] = {0, 0, 0};
                coords(g, i, j, k, x);
                const long m = npos(g, i, j, k);
                if (!is_border(g, i, j, k)) {
                    double f = 1.0;
                    for (int p = 0; p < g->dim; ++p) f += x[p];
                    f *= 1.0 * 1.0; /* RHO * C */
                    if (known == ORACLE_KNOWN_PENNES_TIMELIN) {
                        double tmp = 1.0;
                        for (int p = 0; p < g->dim; ++p) tmp += x[p];
                        tmp *= (1.0 + t);
                        f += c->rhob_cb_w * tmp;
                    }
                    F[m] = f;
                    G[m] = 0.0;
                    double y = 1.0;
                    for (int p = 0; p < g->dim; ++p) y += x[p];
                    y *= (1.0 + t);
                    Y[m] = y;
                } else if (border == ORACLE_BORDER_FIELD) {
                    F[m] = 0.0;
                    double gv = 1.0;
                    for (int p = 0; p < g->dim; ++p) gv += x[p];
                    gv *= (1.0 + t);
                    G[m] = gv;
                    Y[m] = gv;
                } else {
                    /* flux classes: F as on inner nodes; G is the Dirichlet
                     * value on faces with an index 0, else the flux datum
                     * (HeatEqnFDM/Neumann/TimeLinSpaceLin/TimeLinSpaceLin.hpp:102-139,
                     *  HeatEqnFDM/Cauchy/TimeLinSpaceLin/TimeLinSpaceLin.hpp:102-144,
                     *  PennesBioheatEqnFDM/Radiation/TimeLinSpaceLin/TimeLinSpaceLin.hpp:108-158) */
                    double f = 1.0;
                    for (int p = 0; p < g->dim; ++p) f += x[p];
                    f *= 1.0 * 1.0;
                    if (known == ORACLE_KNOWN_PENNES_TIMELIN) {
                        double tmp = 1.0;
                        for (int p = 0; p < g->dim; ++p) tmp += x[p];
                        tmp *= (1.0 + t);
                        f += c->rhob_cb_w * tmp;
                    }
                    F[m] = f;
                    const int idx[3] = {i, j, k};
                    int dir = 0;
                    for (int p = 0; p < g->dim; ++p)
                        if (idx[p] == 0) dir = 1;
                    double gv = 1.0;
                    if (dir) {
                        for (int p = 0; p < g->dim; ++p) gv += x[p];
                        gv *= (1.0 + t);
                    } else if (border == ORACLE_BORDER_NEUMANN) {
                        gv = -1.0 * c->lambda * (1.0 + t);
                    } else if (border == ORACLE_BORDER_CAUCHY) {
                        for (int p = 0; p < g->dim; ++p) gv += x[p];
                        gv *= c->alpha * (1.0 + t);
                        gv += c->lambda * (1.0 + t);
                    } else {
                        for (int p = 0; p < g->dim; ++p) gv += x[p];
                        gv *= (1.0 + t);
                        gv = gv * gv * gv * gv;
                        gv *= c->sigma * c->epsilon;
                        gv += c->lambda * (1.0 + t);
                    }
                    G[m] = gv;
                    double y = 1.0;
                    for (int p = 0; p < g->dim; ++p) y += x[p];
                    y *= (1.0 + t);
                    Y[m] = y;
                }
            }
        }
    }
}

void oracle_known_timelin(const oracle_grid* g, int known, int border, double t,
                          const oracle_heat7_coef* c, double* F, double* G, double* Y)
{
    eval_known_timelin(g, known, border, t, c, F, G, Y);
}

/* compErrLinf over all partition nodes (DataField.hpp:1439-1530); b is a
 * stencil-0 field when bnodes != 0. */
static double linf_vs_nodes(const oracle_grid* g, const double* a, const double* b)
{
    double e = 0.0;
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                double d = fabs(a[pos(g, i, j, k)] - b[npos(g, i, j, k)]);
                if (d > e) e = d;
            }
    return e;
}

double oracle_linf_nodes(const oracle_grid* g, const double* a, const double* b)
{
    double e = 0.0;
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                const long m = pos(g, i, j, k);
                double d = fabs(a[m] - b[m]);
                if (d > e) e = d;
            }
    return e;
}

/* Restated Problem::iterateOverTime (Problem.hpp:2558-2662) for one unknown
 * field U and known fields F, G (and exact Y when computing errors):
 *   n = 0: reset (defaultValue 0 incl. ghosts, :2998-3073), init hooks
 *          (evalInitPbPhase :2754-2826; TimeLinSpaceLin initInner :126-144),
 *          old <- new (setValuesAtMemAll);
 *   n > 0: evalKnownDfs(n) into the "new" known buffers at t = t0 + n*tau
 *          (:3433-3485, time() :2482-2486), update phase with F_old and G_new;
 *   each n: compErrLinf vs Y_new, swap pointers, times advance by tau. */
double oracle_heat7_run(const oracle_grid* g, double* u, int nsteps, double t0,
                        const oracle_heat7_coef* c, int known, int border, double* linf)
{
    const long ne = oracle_nelems(g), nn = oracle_nnodes(g);
    double* uo = (double*)malloc(sizeof(double) * (size_t)ne);
    double* un = (double*)malloc(sizeof(double) * (size_t)ne);
    double* Fo = (double*)calloc((size_t)nn, sizeof(double));
    double* Fn = (double*)calloc((size_t)nn, sizeof(double));
    double* Go = (double*)calloc((size_t)nn, sizeof(double));
    double* Gn = (double*)calloc((size_t)nn, sizeof(double));
    double* Yo = (double*)calloc((size_t)nn, sizeof(double));
    double* Yn = (double*)calloc((size_t)nn, sizeof(double));
    const int timelin = (known == ORACLE_KNOWN_HEAT_TIMELIN || known == ORACLE_KNOWN_PENNES_TIMELIN);
    double t_unknown_new = t0;

    for (int n = 0; n <= nsteps; ++n) {
        if (timelin) eval_known_timelin(g, known, border, t0 + n * c->tau, c, Fn, Gn, Yn);
        if (n == 0) {
            if (timelin) {
                for (long m = 0; m < ne; ++m) un[m] = 0.0;
                for (int k = 0; k < g->n[2]; ++k)
                    for (int j = 0; j < g->n[1]; ++j)
                        for (int i = 0; i < g->n[0]; ++i) {
                            double x[3] = {0, 0, 0};
                            coords(g, i, j, k, x);
                            double y = 1.0;
                            for (int p = 0; p < g->dim; ++p) y += x[p];
                            y *= (1.0 + t0);
                            un[pos(g, i, j, k)] = y;
                        }
            } else {
                memcpy(un, u, sizeof(double) * (size_t)ne);
            }
            memcpy(uo, un, sizeof(double) * (size_t)ne);
            memcpy(Fo, Fn, sizeof(double) * (size_t)nn);
            memcpy(Go, Gn, sizeof(double) * (size_t)nn);
            memcpy(Yo, Yn, sizeof(double) * (size_t)nn);
            t_unknown_new = t0;
        } else {
            oracle_heat7_step(g, uo, un, timelin ? Fo : NULL, Go, Gn,
                              timelin ? border : ORACLE_BORDER_VALUE, 0.0, c);
            t_unknown_new = t_unknown_new + c->tau;
        }
        if (linf && timelin) linf[n] = linf_vs_nodes(g, un, Yn);
        /* swapPointersOf{Known,Unknown}Dfs (:3238-3316) */
        double* tmp;
        tmp = uo; uo = un; un = tmp;
        tmp = Fo; Fo = Fn; Fn = tmp;
        tmp = Go; Go = Gn; Gn = tmp;
        tmp = Yo; Yo = Yn; Yn = tmp;
    }
    memcpy(u, uo, sizeof(double) * (size_t)ne);
    free(uo); free(un); free(Fo); free(Fn); free(Go); free(Gn); free(Yo); free(Yn);
    return t_unknown_new;
}

/*---------------------------------------------------------------------------*/
/* Variable-coefficient conservative 7-point heat.  The reference has no
 * variable-kappa divergence form; this follows the structure of its
 * node-centred hooks (per-axis accumulation into vNew after the copy,
 * HeatEqnFDM.hpp:209-215; coefficient read as a known field like
 * examples/MRIDataPBHEqnFDM/MRIDataPBHEqnFDM.hpp:143-168) with face
 * diffusivity 0.5*(kappa_C + kappa_nb):
 *   u' = u + sum_p tau*(0.5(kE+kC)(uE-uC) - 0.5(kC+kW)(uC-uW)) / h_p^2. */
void oracle_heat7var_step(const oracle_grid* g, const double* uo, double* un,
                          const double* kap, double tau, int border_mode,
                          double border_value)
{
    /* copy-old-to-new before the hooks (Problem.hpp:6226-6256) */
    memcpy(un, uo, sizeof(double) * (size_t)oracle_nelems(g));
    double hh[3];
    long st[3];
    for (int p = 0; p < 3; ++p) {
        hh[p] = p < g->dim ? oracle_gridsize(g, p) * oracle_gridsize(g, p) : 1.0;
        st[p] = stride(g, p);
    }
    const int n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
#pragma omp parallel for collapse(2) schedule(static) num_threads(NTHR)
    for (int k = 0; k < n2; ++k) {
        for (int j = 0; j < n1; ++j) {
            for (int i = 0; i < n0; ++i) {
                const long m = pos(g, i, j, k);
                if (!is_border(g, i, j, k)) {
                    const double uc = uo[m], kc = kap[m];
                    double v = uc;
                    for (int p = 0; p < g->dim; ++p) {
                        const double uw = uo[m - st[p]], ue = uo[m + st[p]];
                        const double kw = kap[m - st[p]], ke = kap[m + st[p]];
                        v += tau * (0.5 * (ke + kc) * (ue - uc) - 0.5 * (kc + kw) * (uc - uw)) / hh[p];
                    }
                    un[m] = v;
                } else if (border_mode == ORACLE_BORDER_VALUE) {
                    un[m] = border_value;
                } else {
                    un[m] = uo[m];
                }
            }
        }
    }
}

void oracle_heat7var_run(const oracle_grid* g, double* u, const double* kap, double tau,
                         int nsteps, int border_mode, double border_value, double* scratch)
{
    const long ne = oracle_nelems(g);
    double* a = u;
    double* b = scratch;
    memcpy(b, a, sizeof(double) * (size_t)ne); /* ghosts identical in both buffers */
    for (int n = 1; n <= nsteps; ++n) {
        oracle_heat7var_step(g, a, b, kap, tau, border_mode, border_value);
        double* t = a; a = b; b = t;
    }
    if (a != u) memcpy(u, a, sizeof(double) * (size_t)ne);
}

/* Gradient of J = 0.5 sum_inner (u^N - target)^2 w.r.t. the kappa field of
 * the variable-kappa step, by the reverse-time discrete adjoint.  The
 * reference differentiates its hooks with ADOL-C (Problem.hpp:5841-5967,
 * 6632-6747); this restates the derivative of the step above:
 *   u'_i = u_i + sum_p (tau/h_p^2) [a_iE (u_E - u_i) - a_iW (u_i - u_W)],
 *   a_ij = (kappa_i + kappa_j)/2, on inner nodes; border nodes depend on
 *   neither u nor kappa (kept or set to the border value).
 * lam^N = u^N - target on inner nodes, 0 on border nodes; per step n = N..1
 *   grad_k += 0.5 * sum_p (tau/h_p^2) sum_{nb = k -+ e_p in the grid}
 *                 (lam_k - lam_nb) (u_nb - u_k),       u = u^{n-1}, lam = lam^n
 *   lam^{n-1} = the step's operator applied to lam^n (it is symmetric), 0 on
 *   border nodes.
 * states holds u^0..u^N (nsteps+1 memAll arrays). */
double oracle_heat7var_grad(const oracle_grid* g, const double* states, const double* kap,
                            const double* target, double tau, int nsteps, double* grad)
{
    const long ne = oracle_nelems(g);
    double* lam = (double*)calloc((size_t)ne, sizeof(double));
    double* lam2 = (double*)calloc((size_t)ne, sizeof(double));
    double hh[3];
    long st[3];
    for (int p = 0; p < 3; ++p) {
        hh[p] = p < g->dim ? oracle_gridsize(g, p) * oracle_gridsize(g, p) : 1.0;
        st[p] = stride(g, p);
    }
    const int n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
    const double* uN = states + (size_t)nsteps * ne;
    double J = 0.0;
    for (int k = 0; k < n2; ++k)
        for (int j = 0; j < n1; ++j)
            for (int i = 0; i < n0; ++i) {
                const long m = pos(g, i, j, k);
                if (is_border(g, i, j, k)) continue;
                const double r = uN[m] - target[m];
                lam[m] = r;
                J += r * r;
            }
    J *= 0.5;
    memset(grad, 0, sizeof(double) * (size_t)ne);
    for (int n = nsteps; n >= 1; --n) {
        const double* u = states + (size_t)(n - 1) * ne;
#pragma omp parallel for collapse(2) schedule(static) num_threads(NTHR)
        for (int k = 0; k < n2; ++k) {
            for (int j = 0; j < n1; ++j) {
                for (int i = 0; i < n0; ++i) {
                    const long m = pos(g, i, j, k);
                    const int idx[3] = {i, j, k};
                    const double lc = lam[m], uc = u[m];
                    double t = 0.0;
                    for (int p = 0; p < g->dim; ++p) {
                        double sp = 0.0;
                        if (idx[p] > 0) sp += (lc - lam[m - st[p]]) * (u[m - st[p]] - uc);
                        if (idx[p] < g->n[p] - 1) sp += (lc - lam[m + st[p]]) * (u[m + st[p]] - uc);
                        t += (tau / hh[p]) * sp;
                    }
                    grad[m] += 0.5 * t;
                    double v = 0.0;
                    if (!is_border(g, i, j, k)) {
                        const double kc = kap[m];
                        v = lc;
                        for (int p = 0; p < g->dim; ++p) {
                            const double lw = lam[m - st[p]], le = lam[m + st[p]];
                            const double kw = kap[m - st[p]], ke = kap[m + st[p]];
                            v += tau * (0.5 * (ke + kc) * (le - lc) - 0.5 * (kc + kw) * (lc - lw)) / hh[p];
                        }
                    }
                    lam2[m] = v;
                }
            }
        }
        double* t = lam;
        lam = lam2;
        lam2 = t;
    }
    free(lam);
    free(lam2);
    return J;
}

/*---------------------------------------------------------------------------*/
/* Pennes bioheat, TumorHeatEqnFDM::updateInner form
 * (examples/TumorHeatEqnFDM/TumorHeatEqnFDM.hpp:233-269) with the
 * piecewise-constant omega/Q replaced by node fields w, Q:
 *   T' = T + sum_p tau*cdiff*(W+E-2T)/h_p^2 + tau*cperf*w*(T_a - T) + tau*csrc*Q.
 * Border nodes keep their copied value (Dirichlet at the initial state). */
void oracle_pennes_step(const oracle_grid* g, const double* to, double* tn, const double* w,
                        const double* q, const oracle_pennes_coef* c)
{
    memcpy(tn, to, sizeof(double) * (size_t)oracle_nelems(g));
    double hh[3];
    long st[3];
    for (int p = 0; p < 3; ++p) {
        hh[p] = p < g->dim ? oracle_gridsize(g, p) * oracle_gridsize(g, p) : 1.0;
        st[p] = stride(g, p);
    }
    const int n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
#pragma omp parallel for collapse(2) schedule(static) num_threads(NTHR)
    for (int k = 0; k < n2; ++k) {
        for (int j = 0; j < n1; ++j) {
            for (int i = 0; i < n0; ++i) {
                const long m = pos(g, i, j, k);
                if (!is_border(g, i, j, k)) {
                    double v = to[m];
                    for (int p = 0; p < g->dim; ++p)
                        v += c->tau * c->cdiff * (to[m - st[p]] + to[m + st[p]] - 2.0 * to[m]) / hh[p];
                    v += c->tau * c->cperf * w[m] * (c->t_a - to[m]);
                    v += c->tau * c->csrc * q[m];
                    tn[m] = v;
                } else {
                    tn[m] = to[m];
                }
            }
        }
    }
}

/* Tumor selection of the hooks (TumorHeatEqnFDM.hpp:239-254 / :282-297):
 * squared distance to the center summed over the axes in order. */
static int tumor_inside(const oracle_grid* g, const oracle_tumor_coef* c, int i, int j, int k)
{
    double x[3] = {0, 0, 0};
    coords(g, i, j, k, x);
    double distance = 0.0;
    for (int p = 0; p < g->dim; ++p) distance += (x[p] - c->center[p]) * (x[p] - c->center[p]);
    return distance <= (c->radius * c->radius);
}

void oracle_tumor_fields(const oracle_grid* g, const oracle_tumor_coef* c, double* omega, double* q)
{
    memset(omega, 0, sizeof(double) * (size_t)oracle_nelems(g));
    memset(q, 0, sizeof(double) * (size_t)oracle_nelems(g));
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                const int in = tumor_inside(g, c, i, j, k);
                omega[pos(g, i, j, k)] = in ? c->omega_tumor : c->omega_brain;
                q[pos(g, i, j, k)] = in ? c->q_tumor : c->q_brain;
            }
}

/* TumorHeatEqnFDM::updateInner (:233-269) and ::updateBorder (:272-349):
 * on border nodes, per axis, the neighbour beyond a face with index n-1 is
 * the inner one minus the Cauchy (last axis) or Neumann flux term; beyond a
 * face with index 0 the inner one plus the Neumann term. */
void oracle_tumor_step(const oracle_grid* g, const double* to, double* tn, const oracle_tumor_coef* c)
{
    memcpy(tn, to, sizeof(double) * (size_t)oracle_nelems(g));
    double h[3];
    long st[3];
    for (int p = 0; p < 3; ++p) {
        h[p] = p < g->dim ? oracle_gridsize(g, p) : 1.0;
        st[p] = stride(g, p);
    }
    const int n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
#pragma omp parallel for collapse(2) schedule(static) num_threads(NTHR)
    for (int k = 0; k < n2; ++k) {
        for (int j = 0; j < n1; ++j) {
            for (int i = 0; i < n0; ++i) {
                const long m = pos(g, i, j, k);
                const int in = tumor_inside(g, c, i, j, k);
                const double omega = in ? c->omega_tumor : c->omega_brain;
                const double Q = in ? c->q_tumor : c->q_brain;
                const int idx[3] = {i, j, k};
                const double u = to[m];
                double v = u;
                const int border = is_border(g, i, j, k);
                for (int p = 0; p < g->dim; ++p) {
                    const double W = to[m - st[p]], E = to[m + st[p]];
                    double t;
                    if (border && idx[p] == g->n[p] - 1) {
                        if (p == g->dim - 1)
                            t = W + W - ((2.0 * h[p] / c->k) * c->h * (u - c->t_inf)) - 2.0 * u;
                        else
                            t = W + W - ((2.0 * h[p] / c->k) * (-1.0 * c->q_bc)) - 2.0 * u;
                    } else if (border && idx[p] == 0) {
                        t = E + ((2.0 * h[p] / c->k) * c->q_bc) + E - 2.0 * u;
                    } else {
                        t = W + E - 2.0 * u;
                    }
                    v += c->tau * (c->k / (c->rho * c->c)) * t / (h[p] * h[p]);
                }
                v += c->tau * ((c->rho_b * c->c_pb) / (c->rho * c->c)) * omega * (c->t_a - u);
                v += c->tau * (1.0 / (c->rho * c->c)) * Q;
                tn[m] = v;
            }
        }
    }
}

/* MRIDataPBHEqnFDM::updateInner (:143-168) and ::updateBorder (:176-260).
 * The open-skull radiation term takes the node's value as accumulated so
 * far in the new field (vNew after the axes before the last one, :200),
 * not the old value; restated as written. */
void oracle_mri_step(const oracle_grid* g, const double* to, double* tn, const double* const* mat,
                     const oracle_mri_coef* c)
{
    memcpy(tn, to, sizeof(double) * (size_t)oracle_nelems(g));
    double h[3];
    long st[3];
    for (int p = 0; p < 3; ++p) {
        h[p] = p < g->dim ? oracle_gridsize(g, p) : 1.0;
        st[p] = stride(g, p);
    }
    const int n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
#pragma omp parallel for collapse(2) schedule(static) num_threads(NTHR)
    for (int k = 0; k < n2; ++k) {
        for (int j = 0; j < n1; ++j) {
            for (int i = 0; i < n0; ++i) {
                const long m = pos(g, i, j, k);
                const long q0 = npos(g, i, j, k);
                const double rho = mat[0][q0], cc = mat[1][q0], kk = mat[2][q0];
                const double rho_blood = mat[3][q0], c_blood = mat[4][q0], omega = mat[5][q0];
                const double t_blood = mat[6][q0], q = mat[7][q0];
                const int trepanation = (int)mat[8][q0];
                const int idx[3] = {i, j, k};
                const double u = to[m];
                double v = u;
                const int border = is_border(g, i, j, k);
                for (int p = 0; p < g->dim; ++p) {
                    const double W = to[m - st[p]], E = to[m + st[p]];
                    double t;
                    if (border && idx[p] == g->n[p] - 1) {
                        if (p == g->dim - 1) {
                            if (trepanation == 1) {
                                const double tempOld = v + 273.15;
                                const double tempOldPow4 = tempOld * tempOld * tempOld * tempOld;
                                const double tempAmb = c->t_amb + 273.15;
                                const double tempAmbPow4 = tempAmb * tempAmb * tempAmb * tempAmb;
                                t = W + W - ((2.0 * h[p] / kk) * c->h * (u - c->t_amb)) -
                                    ((2.0 * h[p] / kk) * c->epsilon * c->sigma * (tempOldPow4 - tempAmbPow4)) - 2.0 * u;
                            } else {
                                t = W + W - ((2.0 * h[p] / kk) * (-1.0 * c->q_skull)) - 2.0 * u;
                            }
                        } else {
                            t = W + W - ((2.0 * h[p] / kk) * (-1.0 * c->q_bc)) - 2.0 * u;
                        }
                    } else if (border && idx[p] == 0) {
                        t = E + ((2.0 * h[p] / kk) * c->q_bc) + E - 2.0 * u;
                    } else {
                        t = W + E - 2.0 * u;
                    }
                    v += c->tau * (kk / (rho * cc)) * t / (h[p] * h[p]);
                }
                v += c->tau * ((rho_blood * c_blood) / (rho * cc)) * omega * (t_blood - u);
                v += c->tau * (1.0 / (rho * cc)) * q;
                tn[m] = v;
            }
        }
    }
}

void oracle_pennes_forward_all(const oracle_grid* g, double* states, const double* w,
                               const double* q, const oracle_pennes_coef* c, int nsteps)
{
    const long ne = oracle_nelems(g);
    for (int n = 1; n <= nsteps; ++n)
        oracle_pennes_step(g, states + (size_t)(n - 1) * ne, states + (size_t)n * ne, w, q, c);
}

/* Discrete adjoint of oracle_pennes_step (inner nodes only; border values
 * are constants, so their adjoint is dropped):
 *   lam^{n-1}_i = lam^n_i + sum_p tau*cdiff*(lam_W + lam_E - 2 lam_i)/h_p^2
 *                 - tau*cperf*w_i*lam^n_i           (lam = 0 on border nodes)
 *   dJ/dw_i    += lam^n_i * tau*cperf*(T_a - T^{n-1}_i). */
double oracle_pennes_grad(const oracle_grid* g, const double* states, const double* w,
                          const double* target, const oracle_pennes_coef* c, int nsteps,
                          double* grad)
{
    const long ne = oracle_nelems(g);
    double* lam = (double*)calloc((size_t)ne, sizeof(double));
    double* lam2 = (double*)calloc((size_t)ne, sizeof(double));
    const double* tN = states + (size_t)nsteps * ne;
    double hh[3];
    long st[3];
    for (int p = 0; p < 3; ++p) {
        hh[p] = p < g->dim ? oracle_gridsize(g, p) * oracle_gridsize(g, p) : 1.0;
        st[p] = stride(g, p);
    }
    double J = 0.0;
    for (long m = 0; m < ne; ++m) grad[m] = 0.0;
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                if (is_border(g, i, j, k)) continue;
                const long m = pos(g, i, j, k);
                const double r = tN[m] - target[m];
                J += 0.5 * r * r;
                lam[m] = r;
            }
    const int n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
    for (int n = nsteps; n >= 1; --n) {
        const double* tp = states + (size_t)(n - 1) * ne;
#pragma omp parallel for collapse(2) schedule(static) num_threads(NTHR)
        for (int k = 0; k < n2; ++k)
            for (int j = 0; j < n1; ++j)
                for (int i = 0; i < n0; ++i) {
                    const long m = pos(g, i, j, k);
                    if (is_border(g, i, j, k)) { lam2[m] = 0.0; continue; }
                    grad[m] += lam[m] * (c->tau * c->cperf * (c->t_a - tp[m]));
                    double v = lam[m];
                    for (int p = 0; p < g->dim; ++p)
                        v += c->tau * c->cdiff * (lam[m - st[p]] + lam[m + st[p]] - 2.0 * lam[m]) / hh[p];
                    v -= c->tau * c->cperf * w[m] * lam[m];
                    lam2[m] = v;
                }
        double* t = lam; lam = lam2; lam2 = t;
    }
    free(lam);
    free(lam2);
    return J;
}

/* oracle_pennes_step on node fields w, Q with the Tumor example's border
 * rule (TumorHeatEqnFDM::updateBorder, TumorHeatEqnFDM.hpp:272-349): every
 * node is updated; on each axis a node with index n_p - 1 replaces its
 * missing outer neighbour by the inner one and the Cauchy (last axis) or
 * Neumann flux term, a node with index 0 by the inner one and the Neumann
 * term.  The operations and their order are those of oracle_tumor_step, so
 * the fields of oracle_tumor_fields give its result bit for bit. */
void oracle_pennes_tumor_step(const oracle_grid* g, const double* to, double* tn, const double* w,
                              const double* q, const oracle_pennes_coef* c, const oracle_tumor_border* b)
{
    memcpy(tn, to, sizeof(double) * (size_t)oracle_nelems(g));
    double h[3];
    long st[3];
    for (int p = 0; p < 3; ++p) {
        h[p] = p < g->dim ? oracle_gridsize(g, p) : 1.0;
        st[p] = stride(g, p);
    }
    const int n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
#pragma omp parallel for collapse(2) schedule(static) num_threads(NTHR)
    for (int k = 0; k < n2; ++k) {
        for (int j = 0; j < n1; ++j) {
            for (int i = 0; i < n0; ++i) {
                const long m = pos(g, i, j, k);
                const int idx[3] = {i, j, k};
                const double u = to[m];
                double v = u;
                for (int p = 0; p < g->dim; ++p) {
                    const double W = to[m - st[p]], E = to[m + st[p]];
                    double t;
                    if (idx[p] == g->n[p] - 1) {
                        if (p == g->dim - 1)
                            t = W + W - ((2.0 * h[p] / b->k) * b->h * (u - b->t_inf)) - 2.0 * u;
                        else
                            t = W + W - ((2.0 * h[p] / b->k) * (-1.0 * b->q_bc)) - 2.0 * u;
                    } else if (idx[p] == 0) {
                        t = E + ((2.0 * h[p] / b->k) * b->q_bc) + E - 2.0 * u;
                    } else {
                        t = W + E - 2.0 * u;
                    }
                    v += c->tau * c->cdiff * t / (h[p] * h[p]);
                }
                v += c->tau * c->cperf * w[m] * (c->t_a - u);
                v += c->tau * c->csrc * q[m];
                tn[m] = v;
            }
        }
    }
}

/* Transpose of the linear part of oracle_pennes_tumor_step.  Node m's row
 * weights its neighbour on axis p by 1, by 2 where m lies on a face (the
 * inner neighbour stands in for the missing one) and by 0 beyond the grid,
 * so the column of node i takes lam of the neighbour at index j-1 twice
 * when that neighbour has index 0, lam of the one at j+1 twice when it has
 * index n_p - 1, and no term beyond the grid; the diagonal gets the Cauchy
 * face's -(2 h/K) H besides -2:
 *   lam'_i = lam_i + sum_p tau*cdiff*t_p/h_p^2 - tau*cperf*w_i*lam_i,
 *   t_p = (lW + lE) [- (2 h_p/K) H lam_i on the Cauchy face] - 2 lam_i. */
void oracle_pennes_tumor_adjoint_step(const oracle_grid* g, const double* lam, double* lam_new, const double* w,
                                      const oracle_pennes_coef* c, const oracle_tumor_border* b)
{
    double h[3];
    long st[3];
    for (int p = 0; p < 3; ++p) {
        h[p] = p < g->dim ? oracle_gridsize(g, p) : 1.0;
        st[p] = stride(g, p);
    }
    const int n0 = g->n[0], n1 = g->n[1], n2 = g->n[2];
#pragma omp parallel for collapse(2) schedule(static) num_threads(NTHR)
    for (int k = 0; k < n2; ++k) {
        for (int j = 0; j < n1; ++j) {
            for (int i = 0; i < n0; ++i) {
                const long m = pos(g, i, j, k);
                const int idx[3] = {i, j, k};
                const double l = lam[m];
                double v = l;
                for (int p = 0; p < g->dim; ++p) {
                    const int x = idx[p], n = g->n[p];
                    double lw = x >= 1 ? lam[m - st[p]] : 0.0;
                    double le = x <= n - 2 ? lam[m + st[p]] : 0.0;
                    if (x == 1) lw = lw + lw;
                    if (x == n - 2) le = le + le;
                    double t;
                    if (x == n - 1 && p == g->dim - 1)
                        t = (lw + le) - ((2.0 * h[p] / b->k) * b->h) * l - 2.0 * l;
                    else
                        t = (lw + le) - 2.0 * l;
                    v += c->tau * c->cdiff * t / (h[p] * h[p]);
                }
                v -= c->tau * c->cperf * w[m] * l;
                lam_new[m] = v;
            }
        }
    }
}

void oracle_pennes_tumor_forward_all(const oracle_grid* g, double* states, const double* w, const double* q,
                                     const oracle_pennes_coef* c, const oracle_tumor_border* b, int nsteps)
{
    const long ne = oracle_nelems(g);
    for (int n = 1; n <= nsteps; ++n)
        oracle_pennes_tumor_step(g, states + (size_t)(n - 1) * ne, states + (size_t)n * ne, w, q, c, b);
}

/* J = 0.5 sum over all nodes (T^N - target)^2 (the Tumor border updates
 * every node) and dJ/dw by the reverse-time discrete adjoint:
 *   lam^N = T^N - target;  for n = N..1:
 *   dJ/dw_i += lam^n_i * tau*cperf*(T_a - T^{n-1}_i);  lam^{n-1} = A^T lam^n. */
double oracle_pennes_tumor_grad(const oracle_grid* g, const double* states, const double* w, const double* target,
                                const oracle_pennes_coef* c, const oracle_tumor_border* b, int nsteps,
                                double* grad)
{
    const long ne = oracle_nelems(g);
    double* lam = (double*)calloc((size_t)ne, sizeof(double));
    double* lam2 = (double*)calloc((size_t)ne, sizeof(double));
    const double* tN = states + (size_t)nsteps * ne;
    double J = 0.0;
    for (long m = 0; m < ne; ++m) grad[m] = 0.0;
    for (int k = 0; k < g->n[2]; ++k)
        for (int j = 0; j < g->n[1]; ++j)
            for (int i = 0; i < g->n[0]; ++i) {
                const long m = pos(g, i, j, k);
                const double r = tN[m] - target[m];
                J += 0.5 * r * r;
                lam[m] = r;
            }
    for (int n = nsteps; n >= 1; --n) {
        const double* tp = states + (size_t)(n - 1) * ne;
        for (int k = 0; k < g->n[2]; ++k)
            for (int j = 0; j < g->n[1]; ++j)
                for (int i = 0; i < g->n[0]; ++i) {
                    const long m = pos(g, i, j, k);
                    grad[m] += lam[m] * (c->tau * c->cperf * (c->t_a - tp[m]));
                }
        oracle_pennes_tumor_adjoint_step(g, lam, lam2, w, c, b);
        double* t = lam; lam = lam2; lam2 = t;
    }
    free(lam);
    free(lam2);
    return J;
}

/*---------------------------------------------------------------------------*/
/* Yee FDTD in a PEC box (normalised units).  The reference's leapfrog mode
 * runs updateInner (H from E) then swaps and runs updateInner2 (E from H)
 * (Problem.hpp:2603-2619); its curl stencil is
 * dest_old + constant*(dfdl - dfdr) with forward differences for H and
 * backward differences for E (examples/EMFDTD/stencils/CompleteCPMLCurlStencil.hpp:38-80,
 * constants +cfln / -cfln at examples/EMFDTD/EMFDTD.hpp:207-212).  Vacuum
 * (no CPML, sigma = 0) reduces it to the plain Yee curl below. */
#define YI(i, j, k) ((i) + (long)M * ((j) + (long)M * (k)))
void oracle_yee_half_h(int N, double s, double* f)
{
    const long M = N + 1, nf = M * M * M;
    double *Ex = f, *Ey = f + nf, *Ez = f + 2 * nf, *Hx = f + 3 * nf, *Hy = f + 4 * nf,
           *Hz = f + 5 * nf;
#pragma omp parallel for schedule(static) num_threads(NTHR)
    for (int k = 0; k <= N; ++k)
        for (int j = 0; j <= N; ++j)
            for (int i = 0; i <= N; ++i) {
                const long m = YI(i, j, k);
                if (j < N && k < N)
                    Hx[m] = Hx[m] + s * ((Ey[YI(i, j, k + 1)] - Ey[m]) - (Ez[YI(i, j + 1, k)] - Ez[m]));
                if (i < N && k < N)
                    Hy[m] = Hy[m] + s * ((Ez[YI(i + 1, j, k)] - Ez[m]) - (Ex[YI(i, j, k + 1)] - Ex[m]));
                if (i < N && j < N)
                    Hz[m] = Hz[m] + s * ((Ex[YI(i, j + 1, k)] - Ex[m]) - (Ey[YI(i + 1, j, k)] - Ey[m]));
            }
}

void oracle_yee_half_e(int N, double s, double* f)
{
    const long M = N + 1, nf = M * M * M;
    double *Ex = f, *Ey = f + nf, *Ez = f + 2 * nf, *Hx = f + 3 * nf, *Hy = f + 4 * nf,
           *Hz = f + 5 * nf;
#pragma omp parallel for schedule(static) num_threads(NTHR)
    for (int k = 0; k <= N; ++k)
        for (int j = 0; j <= N; ++j)
            for (int i = 0; i <= N; ++i) {
                const long m = YI(i, j, k);
                /* tangential E on the PEC walls stays zero */
                if (i < N && j >= 1 && j < N && k >= 1 && k < N)
                    Ex[m] = Ex[m] + s * ((Hz[m] - Hz[YI(i, j - 1, k)]) - (Hy[m] - Hy[YI(i, j, k - 1)]));
                if (j < N && i >= 1 && i < N && k >= 1 && k < N)
                    Ey[m] = Ey[m] + s * ((Hx[m] - Hx[YI(i, j, k - 1)]) - (Hz[m] - Hz[YI(i - 1, j, k)]));
                if (k < N && i >= 1 && i < N && j >= 1 && j < N)
                    Ez[m] = Ez[m] + s * ((Hy[m] - Hy[YI(i - 1, j, k)]) - (Hx[m] - Hx[YI(i, j - 1, k)]));
            }
}

void oracle_yee_step(int N, double s, double* f)
{
    oracle_yee_half_h(N, s, f);
    oracle_yee_half_e(N, s, f);
}

double oracle_yee_energy(int N, const double* f, const double* h_prev)
{
    const long M = N + 1, nf = M * M * M;
    double e = 0.0, h = 0.0;
    for (long m = 0; m < 3 * nf; ++m) e += f[m] * f[m];
    for (long m = 0; m < 3 * nf; ++m) h += f[3 * nf + m] * h_prev[m];
    return 0.5 * e + 0.5 * h;
}

/* Yee with a relative permittivity field eps (one value per node index m,
 * shared by Ex(m), Ey(m), Ez(m)): D += s * (backward differences of H),
 * E = D / eps, i.e. E += (s / eps) * (backward differences of H).  The
 * reference's EMFDTD scales D to E per node by its material coefficient
 * ga = 1 / (1 + sigma dt / eps0) (examples/EMFDTD/EMFDTD.hpp:351-356);
 * 1 / eps takes that role here.  ce = s / eps is formed once per node. */
static int yee_emask(int N, int c, int i, int j, int k)
{
    if (c == 0) return i < N && j >= 1 && j < N && k >= 1 && k < N;
    if (c == 1) return j < N && i >= 1 && i < N && k >= 1 && k < N;
    return k < N && i >= 1 && i < N && j >= 1 && j < N;
}

void oracle_yee_eps_half_e(int N, double s, const double* eps, double* f)
{
    const long M = N + 1, nf = M * M * M;
    double *Ex = f, *Ey = f + nf, *Ez = f + 2 * nf, *Hx = f + 3 * nf, *Hy = f + 4 * nf,
           *Hz = f + 5 * nf;
#pragma omp parallel for schedule(static) num_threads(NTHR)
    for (int k = 0; k <= N; ++k)
        for (int j = 0; j <= N; ++j)
            for (int i = 0; i <= N; ++i) {
                const long m = YI(i, j, k);
                const double ce = s / eps[m];
                if (yee_emask(N, 0, i, j, k))
                    Ex[m] = Ex[m] + ce * ((Hz[m] - Hz[YI(i, j - 1, k)]) - (Hy[m] - Hy[YI(i, j, k - 1)]));
                if (yee_emask(N, 1, i, j, k))
                    Ey[m] = Ey[m] + ce * ((Hx[m] - Hx[YI(i, j, k - 1)]) - (Hz[m] - Hz[YI(i - 1, j, k)]));
                if (yee_emask(N, 2, i, j, k))
                    Ez[m] = Ez[m] + ce * ((Hy[m] - Hy[YI(i - 1, j, k)]) - (Hx[m] - Hx[YI(i, j - 1, k)]));
            }
}

void oracle_yee_eps_step(int N, double s, const double* eps, double* f)
{
    oracle_yee_half_h(N, s, f);
    oracle_yee_eps_half_e(N, s, eps, f);
}

/* Gradient of J = 0.5 sum_{updated E entries} (E^N - target)^2 w.r.t. eps.
 * The step is linear in (E, H): H' = H + s A E, E' = E + C B H' with
 * A = -curl_f, B = curl_b, C = diag(ce) on the updated E entries; as
 * curl_f^T = curl_b on the masked entries, its adjoint is again a pair of
 * Yee half steps run backwards:
 *   mu_H  = lam_H' - Hform(ce * lam_E')     (H-step stencil, coefficient -1)
 *   lam_E = lam_E' - s * Eform(mu_H)        (E-step stencil, coefficient -s)
 * with lam_E kept zero on entries the E step does not update, and
 *   dJ/deps(m) += -(ce/eps)(m) * sum_c lam_E'_c(m) * Eform_c(H')(m),
 * H' = H^{n-1} + s A E^{n-1} recomputed from the stored state.
 * states: (nsteps+1) consecutive states (6 (N+1)^3 values each); target:
 * three E arrays; grad: (N+1)^3 values.  Returns J. */
double oracle_yee_eps_grad(int N, double s, const double* eps, const double* states, const double* target,
                           int nsteps, double* grad)
{
    const long M = N + 1, nf = M * M * M;
    double* lam = (double*)calloc((size_t)(6 * nf), sizeof(double));
    double* hp = (double*)malloc((size_t)(3 * nf) * sizeof(double));
    const double* fN = states + (long)nsteps * 6 * nf;
    double J = 0.0;
    for (int c = 0; c < 3; ++c)
        for (int k = 0; k <= N; ++k)
            for (int j = 0; j <= N; ++j)
                for (int i = 0; i <= N; ++i) {
                    const long m = YI(i, j, k);
                    if (!yee_emask(N, c, i, j, k)) continue;
                    const double r = fN[c * nf + m] - target[c * nf + m];
                    lam[c * nf + m] = r;
                    J += r * r;
                }
    for (long m = 0; m < nf; ++m) grad[m] = 0.0;
    double *lEx = lam, *lEy = lam + nf, *lEz = lam + 2 * nf, *lHx = lam + 3 * nf, *lHy = lam + 4 * nf,
           *lHz = lam + 5 * nf;
    double *Px = hp, *Py = hp + nf, *Pz = hp + 2 * nf;
    for (int n = nsteps; n >= 1; --n) {
        const double* f = states + (long)(n - 1) * 6 * nf;
        const double *Ex = f, *Ey = f + nf, *Ez = f + 2 * nf, *Hx = f + 3 * nf, *Hy = f + 4 * nf,
                     *Hz = f + 5 * nf;
        /* H' and mu_H (in place in lam_H; read only at its own entry) */
#pragma omp parallel for schedule(static) num_threads(NTHR)
        for (int k = 0; k <= N; ++k)
            for (int j = 0; j <= N; ++j)
                for (int i = 0; i <= N; ++i) {
                    const long m = YI(i, j, k);
#define W(F, ii, jj, kk) ((s / eps[YI(ii, jj, kk)]) * F[YI(ii, jj, kk)])
                    Px[m] = Hx[m];
                    Py[m] = Hy[m];
                    Pz[m] = Hz[m];
                    if (j < N && k < N) {
                        Px[m] = Hx[m] + s * ((Ey[YI(i, j, k + 1)] - Ey[m]) - (Ez[YI(i, j + 1, k)] - Ez[m]));
                        lHx[m] = lHx[m] - ((W(lEy, i, j, k + 1) - W(lEy, i, j, k)) -
                                           (W(lEz, i, j + 1, k) - W(lEz, i, j, k)));
                    }
                    if (i < N && k < N) {
                        Py[m] = Hy[m] + s * ((Ez[YI(i + 1, j, k)] - Ez[m]) - (Ex[YI(i, j, k + 1)] - Ex[m]));
                        lHy[m] = lHy[m] - ((W(lEz, i + 1, j, k) - W(lEz, i, j, k)) -
                                           (W(lEx, i, j, k + 1) - W(lEx, i, j, k)));
                    }
                    if (i < N && j < N) {
                        Pz[m] = Hz[m] + s * ((Ex[YI(i, j + 1, k)] - Ex[m]) - (Ey[YI(i + 1, j, k)] - Ey[m]));
                        lHz[m] = lHz[m] - ((W(lEx, i, j + 1, k) - W(lEx, i, j, k)) -
                                           (W(lEy, i + 1, j, k) - W(lEy, i, j, k)));
                    }
#undef W
                }
        /* lam_E and the gradient (lam_E read and written only at its own
         * entry; mu_H and H' at m and the -x/-y/-z neighbours) */
#pragma omp parallel for schedule(static) num_threads(NTHR)
        for (int k = 0; k <= N; ++k)
            for (int j = 0; j <= N; ++j)
                for (int i = 0; i <= N; ++i) {
                    const long m = YI(i, j, k);
                    double acc = 0.0;
                    int any = 0;
                    if (yee_emask(N, 0, i, j, k)) {
                        const double l = lEx[m];
                        acc = acc + l * ((Pz[m] - Pz[YI(i, j - 1, k)]) - (Py[m] - Py[YI(i, j, k - 1)]));
                        lEx[m] = l - s * ((lHz[m] - lHz[YI(i, j - 1, k)]) - (lHy[m] - lHy[YI(i, j, k - 1)]));
                        any = 1;
                    }
                    if (yee_emask(N, 1, i, j, k)) {
                        const double l = lEy[m];
                        acc = acc + l * ((Px[m] - Px[YI(i, j, k - 1)]) - (Pz[m] - Pz[YI(i - 1, j, k)]));
                        lEy[m] = l - s * ((lHx[m] - lHx[YI(i, j, k - 1)]) - (lHz[m] - lHz[YI(i - 1, j, k)]));
                        any = 1;
                    }
                    if (yee_emask(N, 2, i, j, k)) {
                        const double l = lEz[m];
                        acc = acc + l * ((Py[m] - Py[YI(i - 1, j, k)]) - (Px[m] - Px[YI(i, j - 1, k)]));
                        lEz[m] = l - s * ((lHy[m] - lHy[YI(i - 1, j, k)]) - (lHx[m] - lHx[YI(i, j - 1, k)]));
                        any = 1;
                    }
                    if (any) {
                        const double ce = s / eps[m];
                        grad[m] = grad[m] - (ce / eps[m]) * acc;
                    }
                }
    }
    free(lam);
    free(hp);
    return 0.5 * J;
}
#undef YI

/*---------------------------------------------------------------------------*/
/* 2-D heat, scalar kappa, zero Dirichlet border (ensemble member). */
static void heat2d_step(int n, double hh, double ck, const double* u, double* v)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const long m = i + (long)n * j;
            if (i == 0 || j == 0 || i == n - 1 || j == n - 1) { v[m] = u[m]; continue; }
            double x = u[m];
            x += ck * (u[m - 1] + u[m + 1] - 2.0 * u[m]) / hh;
            x += ck * (u[m - n] + u[m + n] - 2.0 * u[m]) / hh;
            v[m] = x;
        }
}

static void lap2d(int n, double hh, const double* u, double* out)
{
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) {
            const long m = i + (long)n * j;
            if (i == 0 || j == 0 || i == n - 1 || j == n - 1) { out[m] = 0.0; continue; }
            out[m] = (u[m - 1] + u[m + 1] - 2.0 * u[m]) / hh + (u[m - n] + u[m + n] - 2.0 * u[m]) / hh;
        }
}

void oracle_heat2d_run(int n, double h, double kappa, double tau, int nsteps, double* u,
                       double* u_prev)
{
    const long nn = (long)n * n;
    double* a = (double*)malloc(sizeof(double) * (size_t)nn);
    double* b = (double*)malloc(sizeof(double) * (size_t)nn);
    memcpy(a, u, sizeof(double) * (size_t)nn);
    memcpy(b, u, sizeof(double) * (size_t)nn);
    const double hh = h * h, ck = tau * kappa;
    for (int s = 0; s < nsteps; ++s) {
        heat2d_step(n, hh, ck, a, b);
        double* t = a; a = b; b = t;
    }
    memcpy(u, a, sizeof(double) * (size_t)nn);
    if (u_prev) memcpy(u_prev, b, sizeof(double) * (size_t)nn);
    free(a);
    free(b);
}

double oracle_heat2d_grad(int n, double h, double kappa, double tau, int nsteps,
                          const double* u0, const double* target, double* dj)
{
    const long nn = (long)n * n;
    const double hh = h * h, ck = tau * kappa;
    double* st = (double*)malloc(sizeof(double) * (size_t)nn * (size_t)(nsteps + 1));
    double* lam = (double*)calloc((size_t)nn, sizeof(double));
    double* lam2 = (double*)calloc((size_t)nn, sizeof(double));
    double* lu = (double*)calloc((size_t)nn, sizeof(double));
    memcpy(st, u0, sizeof(double) * (size_t)nn);
    for (int s = 1; s <= nsteps; ++s) heat2d_step(n, hh, ck, st + (size_t)(s - 1) * nn, st + (size_t)s * nn);
    const double* uN = st + (size_t)nsteps * nn;
    double J = 0.0;
    for (long m = 0; m < nn; ++m) {
        const int i = (int)(m % n), j = (int)(m / n);
        if (i == 0 || j == 0 || i == n - 1 || j == n - 1) continue;
        const double r = uN[m] - target[m];
        J += 0.5 * r * r;
        lam[m] = r;
    }
    double g = 0.0;
    for (int s = nsteps; s >= 1; --s) {
        lap2d(n, hh, st + (size_t)(s - 1) * nn, lu);
        for (long m = 0; m < nn; ++m) g += lam[m] * tau * lu[m];
        lap2d(n, hh, lam, lu);
        for (long m = 0; m < nn; ++m) {
            const int i = (int)(m % n), j = (int)(m / n);
            lam2[m] = (i == 0 || j == 0 || i == n - 1 || j == n - 1) ? 0.0 : lam[m] + ck * lu[m];
        }
        double* t = lam; lam = lam2; lam2 = t;
    }
    *dj = g;
    free(st); free(lam); free(lam2); free(lu);
    return J;
}
