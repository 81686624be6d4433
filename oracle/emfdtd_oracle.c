/* emfdtd_oracle.c -- CPU restatement of the reference's EMFDTD example (the
 * Yee leapfrog with a complete CPML and a conductivity-driven D -> E
 * relation).  TEST INFRASTRUCTURE ONLY: tests, smoke() and bench.py's CPU
 * baseline call it as the checker; the product never does.
 *
 * Reference files followed:
 *   examples/EMFDTD/EMFDTD.hpp:295-313 (updateInner: H from E, three CPML
 *     curl stencils), :330-386 (updateInner2: D from H, then E = ga (D - I),
 *     I += gb E, hard source on Ez, sink sample), :167-172 and :208-213
 *     (stencil offsets, ranges and constants), :191-193 (cfln = 0.5, ds = 1,
 *     dt = cfln ds / C_0)
 *   examples/EMFDTD/stencils/CompleteCPMLCurlStencil.hpp:38-80 (one curl
 *     component with its two auxiliary psi fields)
 *   examples/EMFDTD/CPMLParams.hpp:23-49 (b, c, 1/kappa tables)
 *   examples/EMFDTD/sources/ModulatedGaussianSource.hpp:25-34
 *   examples/EMFDTD/physical_constants.hpp:12-14 (EPSILON_0 is a long double
 *     literal, so every expression that divides by it is evaluated in long
 *     double and rounded to double on assignment; restated the same way)
 *   src/ScaFES_Problem.hpp:2603-2619 (leapfrog: update phase, swap, update2
 *     phase; both phases read only values the phase does not write except at
 *     the node itself, so each runs in place here)
 *
 * Layout: every field is an n^3 node block, x fastest, no ghosts.  The only
 * read outside the node block is the Dx stencil's Hz(i, -1, k) at j = 0 (its
 * range starts at j = 0, EMFDTD.hpp:211); the reference reads the never
 * written ghost layer there, which is 0.
 */
#define _DEFAULT_SOURCE /* M_PI from math.h, the constant the reference's source uses */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "scafes_oracle.h"

#define EPS0_L 8.85418781762e-12L
#define C0_D 299792458.0

void oracle_cpml_tables(int d, double m_a, double m, double k_max, double s_max, double a_max, double dt,
                        double* b, double* c, double* ook)
{
    /* CPMLParams ctor (CPMLParams.hpp:23-49): sizes d+1, init b = c = 0, ook = 1; fill w < d; then
     * reverse the first d entries (index 0 = outermost layer). */
    for (int w = 0; w <= d; ++w) {
        b[w] = 0.0;
        c[w] = 0.0;
        ook[w] = 1.0;
    }
    const double dd = (double)d;
    for (int w = 0; w < d; ++w) {
        const double wd = pow((double)w / (dd - 1.), m);
        const double dw_a = pow((dd - 1. - (double)w) / (dd - 1.), m_a);
        const double s = wd * s_max;
        const double k = 1. + ((k_max - 1.) * wd);
        const double a_ = dw_a * a_max;
        b[w] = (double)expl(-dt * (((s / k) + a_) / EPS0_L));
        c[w] = (s * (b[w] - 1.)) / ((s * k) + (k * k * a_));
        if (b[w] == 1.) c[w] = 0.;
        ook[w] = 1. / k;
    }
    for (int lo = 0, hi = d - 1; lo < hi; ++lo, --hi) {
        double t;
        t = b[lo]; b[lo] = b[hi]; b[hi] = t;
        t = c[lo]; c[lo] = c[hi]; c[hi] = t;
        t = ook[lo]; ook[lo] = ook[hi]; ook[hi] = t;
    }
}

double oracle_emfdtd_dt(double cfln)
{
    return (cfln * 1.0) / C0_D;
}

void oracle_emfdtd_ga_gb(long count, const double* sigma, double dt, double* ga, double* gb)
{
    for (long m = 0; m < count; ++m) {
        ga[m] = (double)(1. / (1. + (sigma[m]) * dt / EPS0_L));
        gb[m] = (double)((sigma[m]) * dt / EPS0_L);
    }
}

double oracle_mod_gauss(double t, double dt, double f_mod, double t0, double spread)
{
    double v = (t0 - t) / spread;
    v *= v;
    v *= -.5;
    v = exp(v);
    const double mod = cos(2. * M_PI * f_mod * t * dt);
    return v * mod;
}

typedef struct {
    int ol[3], or_[3];  /* offsets of the two differenced fields */
    int lo[3], hi[3];   /* valid node range (inclusive) */
    int end[3];         /* I_, J_, K_ of the stencil */
    double constant;
} curl_spec;

/* One CompleteCurlCPMLStencil::operator() evaluation at node (i,j,k). */
static double curl_cpml(const oracle_emfdtd* p, const curl_spec* s, int i, int j, int k, double dest_old,
                        const double* from_l, double* psi_l, const double* from_r, double* psi_r)
{
    const int n = p->n;
    const int idx[3] = {i, j, k};
    for (int q = 0; q < 3; ++q)
        if (s->lo[q] > idx[q] || idx[q] > s->hi[q]) return 0.;
    const long m = (long)i + (long)n * ((long)j + (long)n * k);
    int dlo = 0, dro = 0;
    for (int q = 0; q < 3; ++q) {
        const int md = idx[q] < s->end[q] - idx[q] ? idx[q] : s->end[q] - idx[q];
        dlo += md * abs(s->ol[q]);
        dro += md * abs(s->or_[q]);
    }
    const int dl = dlo < p->d ? dlo : p->d;
    const int dr = dro < p->d ? dro : p->d;
    /* neighbour reads; a read at index -1 is the zero ghost layer */
    const int li = i + s->ol[0], lj = j + s->ol[1], lk = k + s->ol[2];
    const int ri = i + s->or_[0], rj = j + s->or_[1], rk = k + s->or_[2];
    const double fl = (li < 0 || lj < 0 || lk < 0) ? 0.0 : from_l[li + (long)n * (lj + (long)n * lk)];
    const double fr = (ri < 0 || rj < 0 || rk < 0) ? 0.0 : from_r[ri + (long)n * (rj + (long)n * rk)];
    double dfdl = fl - from_l[m];
    const double psi_l_n = (p->b[dl] * psi_l[m]) + (p->c[dl] * dfdl);
    dfdl = (p->ook[dl] * dfdl) + psi_l_n;
    psi_l[m] = psi_l_n;
    double dfdr = fr - from_r[m];
    const double psi_r_n = (p->b[dr] * psi_r[m]) + (p->c[dr] * dfdr);
    dfdr = (p->ook[dr] * dfdr) + psi_r_n;
    psi_r[m] = psi_r_n;
    return dest_old + (s->constant * (dfdl - dfdr));
}

static curl_spec make_spec(int n, int l0, int l1, int l2, int r0, int r1, int r2, int s0, int s1, int s2, int e0,
                           int e1, int e2, double constant)
{
    curl_spec s;
    const int ol[3] = {l0, l1, l2}, or_[3] = {r0, r1, r2}, st[3] = {s0, s1, s2}, en[3] = {e0, e1, e2};
    for (int q = 0; q < 3; ++q) {
        s.ol[q] = ol[q];
        s.or_[q] = or_[q];
        s.lo[q] = st[q];
        s.end[q] = (n - 1) + en[q];
        s.hi[q] = s.end[q];
    }
    s.constant = constant;
    return s;
}

void oracle_emfdtd_step(const oracle_emfdtd* p, double* const* f, double src_value, double* sink)
{
    const int n = p->n;
    const double S = p->cfln;
    /* EMFDTD.hpp:167-172 stencil offsets and :208-213 ranges / constants */
    const curl_spec hx = make_spec(n, 0, 0, 1, 0, 1, 0, 1, 0, 0, 0, -1, -1, S);
    const curl_spec hy = make_spec(n, 1, 0, 0, 0, 0, 1, 0, 1, 0, -1, 0, -1, S);
    const curl_spec hz = make_spec(n, 0, 1, 0, 1, 0, 0, 0, 0, 1, -1, -1, 0, S);
    const curl_spec dx = make_spec(n, 0, -1, 0, 0, 0, -1, 0, 0, 1, -1, 0, 0, -S);
    const curl_spec dy = make_spec(n, 0, 0, -1, -1, 0, 0, 1, 0, 1, 0, -1, 0, -S);
    const curl_spec dz = make_spec(n, -1, 0, 0, 0, -1, 0, 1, 1, 0, 0, 0, -1, -S);
    /* update phase: H from E (updateInner) */
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const long m = (long)i + (long)n * ((long)j + (long)n * k);
                f[OEM_HX][m] = curl_cpml(p, &hx, i, j, k, f[OEM_HX][m], f[OEM_EY], f[OEM_PSI_EYDZ], f[OEM_EZ],
                                         f[OEM_PSI_EZDY]);
                f[OEM_HY][m] = curl_cpml(p, &hy, i, j, k, f[OEM_HY][m], f[OEM_EZ], f[OEM_PSI_EZDX], f[OEM_EX],
                                         f[OEM_PSI_EXDZ]);
                f[OEM_HZ][m] = curl_cpml(p, &hz, i, j, k, f[OEM_HZ][m], f[OEM_EX], f[OEM_PSI_EXDY], f[OEM_EY],
                                         f[OEM_PSI_EYDX]);
            }
    /* update2 phase: D from H, then E and I (updateInner2) */
    const long src = (long)p->src[0] + (long)n * ((long)p->src[1] + (long)n * p->src[2]);
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j)
            for (int i = 0; i < n; ++i) {
                const long m = (long)i + (long)n * ((long)j + (long)n * k);
                const double Dx = curl_cpml(p, &dx, i, j, k, f[OEM_DX][m], f[OEM_HZ], f[OEM_PSI_HZDY], f[OEM_HY],
                                            f[OEM_PSI_HYDZ]);
                const double Dy = curl_cpml(p, &dy, i, j, k, f[OEM_DY][m], f[OEM_HX], f[OEM_PSI_HXDZ], f[OEM_HZ],
                                            f[OEM_PSI_HZDX]);
                const double Dz = curl_cpml(p, &dz, i, j, k, f[OEM_DZ][m], f[OEM_HY], f[OEM_PSI_HYDX], f[OEM_HX],
                                            f[OEM_PSI_HXDY]);
                const double ga = f[OEM_GA][m], gb = f[OEM_GB][m];
                const double Ex = ga * (Dx - f[OEM_IX][m]);
                const double Ey = ga * (Dy - f[OEM_IY][m]);
                const double Ez = ga * (Dz - f[OEM_IZ][m]);
                f[OEM_IX][m] = f[OEM_IX][m] + (gb * Ex);
                f[OEM_IY][m] = f[OEM_IY][m] + (gb * Ey);
                f[OEM_IZ][m] = f[OEM_IZ][m] + (gb * Ez);
                f[OEM_DX][m] = Dx;
                f[OEM_DY][m] = Dy;
                f[OEM_DZ][m] = Dz;
                f[OEM_EX][m] = Ex;
                f[OEM_EY][m] = Ey;
                f[OEM_EZ][m] = m == src ? src_value : Ez;
            }
    if (sink) {
        const long s = (long)p->sink[0] + (long)n * ((long)p->sink[1] + (long)n * p->sink[2]);
        *sink = f[OEM_EZ][s];
    }
}

void oracle_emfdtd_run(const oracle_emfdtd* p, double* const* f, int nsteps, double t_start, double f_mod,
                       double t0, double spread, double* sink)
{
    const double dt = oracle_emfdtd_dt(p->cfln);
    for (int s = 1; s <= nsteps; ++s) {
        /* the source reads the old field's time, t_start + (s-1) tau, tau = 1 */
        const double t_old = t_start + (double)(s - 1) * 1.0;
        oracle_emfdtd_step(p, f, oracle_mod_gauss(t_old, dt, f_mod, t0, spread), sink ? &sink[s - 1] : NULL);
    }
}

double oracle_ellipsoid(double x, double y, double z, double cx, double cy, double cz, double a, double b_,
                        double c, double inside)
{
    /* IndicatorEllipsoid::operator() (geometry/IndicatorEllipsoid.hpp:48-68), h = 1e-8 */
    const double h = 1.e-8;
    const double fx = x - cx, fy = y - cy, fz = z - cz;
    const double tmp = (fx * fx) + (fy * fy);
    const double theta = atan2(sqrt(tmp), fz);
    const double phi = atan2(fy, fx);
    const double ro = sqrt(tmp + (fz * fz));
    const double x1 = a * cos(theta) * cos(phi);
    const double y1 = b_ * cos(theta) * sin(phi);
    const double z1 = c * sin(theta);
    const double rn = sqrt((x1 * x1) + (y1 * y1) + (z1 * z1));
    return 0.5 * inside * (((ro + rn) / (sqrt((ro + rn) * (ro + rn) + h * h))) -
                           ((ro - rn) / (sqrt((ro - rn) * (ro - rn) + h * h))));
}
