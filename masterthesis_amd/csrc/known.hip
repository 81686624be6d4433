// known.hip -- known data fields of the TimeLinSpaceLin system problems,
// evaluated on the device.
//
// The reference re-evaluates every known field at every node on the host
// each time step (evalKnownDfs, src/ScaFES_Problem.hpp:3433-3485, calling the
// problem's evalInner/evalBorder hooks).  For the TimeLinSpaceLin problems
// those hooks are closed forms of (x, t), so a kernel can produce F, G and the
// exact solution Y in HBM directly and the host never touches them:
//   heat   tests/system/HeatEqnFDM/Dirichlet/TimeLinSpaceLin/TimeLinSpaceLin.hpp:75-123
//   Pennes tests/system/PennesBioheatEqnFDM/Dirichlet/TimeLinSpaceLin/TimeLinSpaceLin.hpp:75-124
//   flux borders: tests/system/HeatEqnFDM/{Neumann,Cauchy}/TimeLinSpaceLin/TimeLinSpaceLin.hpp:102-144,
//                 tests/system/PennesBioheatEqnFDM/Radiation/TimeLinSpaceLin/TimeLinSpaceLin.hpp:108-158
// Every expression keeps the hooks' operation order, so the fields are
// bit-identical to the host evaluation (-ffp-contract=off).
#include <hip/hip_runtime.h>

#include "scafes_common.hpp"

namespace scafes {
namespace {

struct KnownArgs {
    int dim, n0, n1, n2, nl;
    double h[3], x0[3];
    int form, border;
    double rho_c, rhob_cb_w, lambda, alpha, epsilon, sigma;
    double t;
};

// x fastest, one thread per node; F and G are stencil-0 fields (no ghosts),
// Y has the memAll layout ygeo (the unknown field's, for compErr).
__global__ __launch_bounds__(256) void known_timelin_kernel(KnownArgs a, Geo ygeo, double* __restrict__ F,
                                                            double* __restrict__ G, double* __restrict__ Y)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;
    const int k = blockIdx.z;
    if (i >= a.n0) return;
    const int idx[3] = {i, j, k};
    const int nn[3] = {a.n0, a.n1, a.n2};
    double x[3] = {0.0, 0.0, 0.0};
    bool border = false, near = false;
    for (int p = 0; p < a.dim; ++p) {
        x[p] = a.x0[p] + a.h[p] * (double)idx[p];
        border = border || idx[p] < a.nl || idx[p] > nn[p] - 1 - a.nl;
        near = near || idx[p] == 0;
    }
    double s = 1.0;  // 1 + sum_p x_p, summed in axis order
    for (int p = 0; p < a.dim; ++p) s += x[p];
    const double tt = 1.0 + a.t;
    const long m = (long)i + (long)a.n0 * ((long)j + (long)a.n1 * k);
    double f, gv, y = s * tt;
    if (!border || a.border != SCAFES_BORDER_FIELD) {
        // evalInner (the flux classes' evalBorder starts with it)
        f = s * a.rho_c;
        if (a.form == SCAFES_KNOWN_PENNES_TIMELIN) f += a.rhob_cb_w * (s * tt);
        gv = 0.0;
        if (border) {
            if (near) {
                gv = y;
            } else if (a.border == SCAFES_BORDER_NEUMANN) {
                gv = (-1.0 * a.lambda) * tt;
            } else if (a.border == SCAFES_BORDER_CAUCHY) {
                gv = s * (a.alpha * tt);
                gv += a.lambda * tt;
            } else {
                gv = s * tt;
                gv = gv * gv * gv * gv;
                gv *= a.sigma * a.epsilon;
                gv += a.lambda * tt;
            }
        }
    } else {
        // Dirichlet border: F = 0, G = Y
        f = 0.0;
        gv = y;
    }
    if (F) F[m] = f;
    if (G) G[m] = gv;
    if (Y) Y[at(ygeo, i, j, k)] = y;
}

}  // namespace
}  // namespace scafes

using namespace scafes;

extern "C" {

int scafes_known_timelin_f64(scafes_ctx* ctx, const scafes_geom* g, const scafes_timelin_params* p, double t,
                             double* F, double* G, double* Y, int hw_y)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !p) return fail(SCAFES_ERR_ARG, "known_timelin: null pointer");
    Geo geo, ygeo;
    if (make_geo(g, 0, &geo) != SCAFES_OK) return SCAFES_ERR_ARG;
    if (make_geo(g, hw_y, &ygeo) != SCAFES_OK) return SCAFES_ERR_ARG;
    if (p->form != SCAFES_KNOWN_HEAT_TIMELIN && p->form != SCAFES_KNOWN_PENNES_TIMELIN)
        return fail(SCAFES_ERR_ARG, "known_timelin: unknown form");
    if (p->border_mode != SCAFES_BORDER_FIELD && p->border_mode != SCAFES_BORDER_NEUMANN &&
        p->border_mode != SCAFES_BORDER_CAUCHY && p->border_mode != SCAFES_BORDER_RADIATION)
        return fail(SCAFES_ERR_ARG, "known_timelin: border must be FIELD, NEUMANN, CAUCHY or RADIATION");
    if (g->n[1] > 65535 || g->n[2] > 65535) return fail(SCAFES_ERR_ARG, "known_timelin: grid too large");
    if (!F && !G && !Y) return SCAFES_OK;
    KnownArgs a;
    a.dim = g->dim;
    a.n0 = g->n[0];
    a.n1 = g->n[1];
    a.n2 = g->n[2];
    a.nl = g->nlayers;
    for (int q = 0; q < 3; ++q) {
        a.h[q] = g->h[q];
        a.x0[q] = p->x0[q];
    }
    a.form = p->form;
    a.border = p->border_mode;
    a.rho_c = p->rho_c;
    a.rhob_cb_w = p->rhob_cb_w;
    a.lambda = p->lambda;
    a.alpha = p->alpha;
    a.epsilon = p->epsilon;
    a.sigma = p->sigma;
    a.t = t;
    const dim3 grid((unsigned)ceil_div(g->n[0], 256), (unsigned)g->n[1], (unsigned)g->n[2]);
    known_timelin_kernel<<<grid, 256, 0, c->stream>>>(a, ygeo, F, G, Y);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

}  // extern "C"
