/*
 * scafes_oracle.h -- CPU restatement of ScaFES's explicit time-stepping path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  Nothing in masterthesis_amd/ links or calls it.
 *
 * Every routine restates the reference's FP64 arithmetic in the reference's
 * operation order (compiled with -ffp-contract=off so no FMA contraction
 * changes the rounding), citing the reference file:line it follows.
 *
 * Layout (ScaFES DataField, src/ScaFES_DataField.hpp:622-642, :1187-1199;
 * src/ScaFES_Grid.hpp:955-968): Fortran order, x fastest.  A field with
 * stencil width hw has memAll extents n_p + 2*hw on active axes (extent 1
 * and no halo on axes >= dim).  The trailing time slot of the reference
 * lump (DataField.hpp:947-953) is kept by the caller as a scalar.
 *
 * Pinning status of each routine (how it is checked, see tests/):
 *   oracle_heat7_*      pinned: exact discrete solutions of the reference's
 *                       system tests (TimeLinSpaceLin heat/Pennes with
 *                       Dirichlet, Neumann, Cauchy and radiation borders),
 *                       the discrete sine eigenmode closed form, brute force.
 *   oracle_heat7var_*   pinned: reduction to heat7 for constant kappa,
 *                       operator symmetry, face-assembled numpy operator.
 *   oracle_pennes_*     pinned: reduction to heat7 (w=0,Q=0), Tumor-form
 *                       closed-form steady state, adjoint vs central
 *                       finite differences and dot-product test.
 *   oracle_pennes_tumor_* pinned: the Tumor-border step equals
 *                       oracle_tumor_step bit for bit on the example's
 *                       fields; adjoint step by the dot-product test
 *                       against the step's linear part; gradient vs central
 *                       finite differences.
 *   oracle_yee_*        pinned: exact PEC cavity eigenmode with the Yee
 *                       dispersion relation, discrete energy invariant,
 *                       discrete Gauss law.
 *   oracle_yee_eps_*    pinned: eps = 1 reduces to oracle_yee_step bit for
 *                       bit; constant eps = a numpy-assembled step; the
 *                       eps-weighted energy invariant; gradient vs central
 *                       finite differences (every component, tiny grid).
 *   oracle_heat2d_*     pinned: 2-D sine eigenmode closed form, gradient vs
 *                       central finite differences.
 *   ADOL-C forward-mode gradients of the reference: parity unpinned (ADOL-C
 *                       is not installed and no reference test checks them).
 */
#ifndef SCAFES_ORACLE_H_
#define SCAFES_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Geometry of one field on the global (single-partition) grid. */
typedef struct {
    int dim;          /* 1..3 */
    int n[3];         /* nodes per axis; axes >= dim must be 1 */
    int hw;           /* stencil width (ghost layers) of this field */
    double x0[3];     /* coordNodeFirst */
    double x1[3];     /* coordNodeLast */
    int nlayers;      /* nLayersAtBorder (Parameters.hpp:446) */
} oracle_grid;

/* Border rule for the unknown field at GLOBAL_BORDER nodes.  The flux rules
 * follow the reference's system-test border classes: a border node with an
 * index 0 on any axis takes the Dirichlet value G_new; otherwise, on every
 * axis where its index is n_p - 1, the missing outer neighbour is replaced
 * by the inner one plus a flux term built from G_old:
 *   NEUMANN   tests/system/HeatEqnFDM/Neumann/Neumann.hpp:126-168
 *   CAUCHY    tests/system/HeatEqnFDM/Cauchy/Cauchy.hpp:126-170
 *   RADIATION tests/system/PennesBioheatEqnFDM/Radiation/Radiation.hpp:126-175 */
enum {
    ORACLE_BORDER_KEEP = 0,
    ORACLE_BORDER_FIELD = 1,
    ORACLE_BORDER_VALUE = 2,
    ORACLE_BORDER_NEUMANN = 3,
    ORACLE_BORDER_CAUCHY = 4,
    ORACLE_BORDER_RADIATION = 5
};

/* Known-field evaluator used inside the restated time loop. */
enum {
    ORACLE_KNOWN_NONE = 0,       /* F = G = 0 (sine-mode heat)                 */
    ORACLE_KNOWN_HEAT_TIMELIN,   /* tests/system/HeatEqnFDM/Dirichlet/TimeLinSpaceLin */
    ORACLE_KNOWN_PENNES_TIMELIN  /* tests/system/PennesBioheatEqnFDM/Dirichlet/TimeLinSpaceLin */
};

/* Coefficients of the constant-coefficient heat/Pennes node update
 * (HeatEqnFDM.hpp:204-218, PennesBioheatEqnFDM.hpp:153-168). */
typedef struct {
    double tau;
    double coeff_a;     /* LAMBDA/(RHO*C) */
    double rhoc_inv;    /* 1/(RHO*C) multiplying F_old */
    double perf_c;      /* (RHO_BLOOD*C_BLOOD)/(RHO*C)          (if has_perf) */
    double perf_w;      /* W                                     (if has_perf) */
    double rhob_cb_w;   /* RHO_BLOOD*C_BLOOD*W inside the Pennes source F */
    int has_perf;
    int has_source;
    double lambda;      /* LAMBDA of the flux borders (HeatEqnFDM.hpp:98) */
    double alpha;       /* ALPHA (Cauchy) */
    double epsilon;     /* EPSILON (radiation) */
    double sigma;       /* SIGMA (radiation) */
} oracle_heat7_coef;

long oracle_nelems(const oracle_grid* g);
long oracle_nnodes(const oracle_grid* g);
double oracle_gridsize(const oracle_grid* g, int p);

/* One update step (copy-old-to-new, inner stencil, border rule).  f_old,
 * g_old, g_new: stencil-0 known fields (nullable where the rule does not
 * read them). */
void oracle_heat7_step(const oracle_grid* g, const double* u_old, double* u_new,
                       const double* f_old, const double* g_old, const double* g_new,
                       int border_mode, double border_value, const oracle_heat7_coef* c);

/* Full restated iterateOverTime for the heat/Pennes system problems.
 * u: memAll of the unknown field; on entry ignored if known==TIMELIN (the
 * init hook is restated), else the initial state; on exit the final state.
 * border: ORACLE_BORDER_FIELD (Dirichlet system tests) or one of the flux
 * rules; the TIMELIN known fields G then follow that border class's
 * TimeLinSpaceLin test.  linf (nullable, nsteps+1 entries): "Local Linf"
 * of compErrLinf per step against the exact solution (TIMELIN modes only).
 * Returns final time. */
double oracle_heat7_run(const oracle_grid* g, double* u, int nsteps, double t0,
                        const oracle_heat7_coef* c, int known, int border, double* linf);

/* The TimeLinSpaceLin known fields at time t (evalInner/evalBorder of the
 * system problems, restated as used by oracle_heat7_run): F, G and the exact
 * solution Y, stencil-0 arrays of oracle_nnodes(g) doubles. */
void oracle_known_timelin(const oracle_grid* g, int known, int border, double t,
                          const oracle_heat7_coef* c, double* F, double* G, double* Y);

/* Variable-coefficient conservative 7-point heat: one step, FP64.
 * kappa has the same layout (and stencil width) as u. */
void oracle_heat7var_step(const oracle_grid* g, const double* u_old, double* u_new,
                          const double* kappa, double tau, int border_mode,
                          double border_value);
/* Gradient of J = 0.5 sum_inner (u^N - target)^2 w.r.t. kappa (discrete
 * adjoint); states = u^0..u^N.  Returns J. */
double oracle_heat7var_grad(const oracle_grid* g, const double* states, const double* kappa,
                            const double* target, double tau, int nsteps, double* grad);
void oracle_heat7var_run(const oracle_grid* g, double* u, const double* kappa,
                         double tau, int nsteps, int border_mode, double border_value,
                         double* scratch);

/* Pennes bioheat with perfusion field w and source field Q (TumorHeatEqnFDM
 * form, examples/TumorHeatEqnFDM/TumorHeatEqnFDM.hpp:233-269), Dirichlet
 * borders kept at their initial values.  w, q: same layout as T. */
typedef struct {
    double tau;
    double cdiff;   /* K/(RHO*C) */
    double cperf;   /* 1/(RHO*C) (w already carries rho_b*c_b) */
    double csrc;    /* 1/(RHO*C) */
    double t_a;
} oracle_pennes_coef;

void oracle_pennes_step(const oracle_grid* g, const double* t_old, double* t_new,
                        const double* w, const double* q, const oracle_pennes_coef* c);
/* states: (nsteps+1) * nelems doubles, states[0] = initial state. */
void oracle_pennes_forward_all(const oracle_grid* g, double* states, const double* w,
                               const double* q, const oracle_pennes_coef* c, int nsteps);
/* J = 0.5*sum_{inner}(T^N - target)^2 and dJ/dw (nelems, zero on non-inner
 * nodes) by the explicit reverse-time discrete adjoint over stored states. */
double oracle_pennes_grad(const oracle_grid* g, const double* states, const double* w,
                          const double* target, const oracle_pennes_coef* c, int nsteps,
                          double* grad);

/* The Tumor example's border constants (TumorHeatEqnFDM.hpp:272-349):
 * conductivity K, convection coefficient H, ambient T_INF, heat flux Q_BC. */
typedef struct {
    double k, h, t_inf, q_bc;
} oracle_tumor_border;

/* oracle_pennes_step with the Tumor border rule on every border node (all
 * nodes updated); fields w, Q as for oracle_pennes_step. */
void oracle_pennes_tumor_step(const oracle_grid* g, const double* t_old, double* t_new, const double* w,
                              const double* q, const oracle_pennes_coef* c, const oracle_tumor_border* b);
/* lam_new = A^T lam, A the linear part of oracle_pennes_tumor_step. */
void oracle_pennes_tumor_adjoint_step(const oracle_grid* g, const double* lam, double* lam_new, const double* w,
                                      const oracle_pennes_coef* c, const oracle_tumor_border* b);
void oracle_pennes_tumor_forward_all(const oracle_grid* g, double* states, const double* w, const double* q,
                                     const oracle_pennes_coef* c, const oracle_tumor_border* b, int nsteps);
/* J = 0.5*sum_{all nodes}(T^N - target)^2 and dJ/dw (every node). */
double oracle_pennes_tumor_grad(const oracle_grid* g, const double* states, const double* w, const double* target,
                                const oracle_pennes_coef* c, const oracle_tumor_border* b, int nsteps,
                                double* grad);

/* Yee FDTD, vacuum, normalised (c=1, h=1, Courant s=tau/h), PEC box of
 * N cells per axis; fields stored on (N+1)^3 nodes each, no ghost layer.
 * order of f: Ex,Ey,Ez,Hx,Hy,Hz, each (N+1)^3.  One leapfrog step = H
 * half step (updateInner) then E half step (updateInner2), as the
 * reference's leapfrog mode (Problem.hpp:2603-2619, EMFDTD.hpp:295-386). */
void oracle_yee_half_h(int ncell, double s, double* f);
void oracle_yee_half_e(int ncell, double s, double* f);
void oracle_yee_step(int ncell, double s, double* f);
/* Discrete energy 0.5*sum(E^n.E^n) + 0.5*sum(H^{n-1/2}.H^{n+1/2}); f holds
 * E^n and H^{n+1/2} (state right after an H half step), h_prev the three
 * H components before it.  Invariant of the scheme in exact arithmetic. */
double oracle_yee_energy(int ncell, const double* f, const double* h_prev);
/* Yee with a relative permittivity field eps ((N+1)^3 values, one per node
 * index, shared by the three E components there): the E half step becomes
 * E += (s / eps) * (backward differences of H), i.e. D += s curl H with
 * E = D / eps (the per-node D -> E scaling of examples/EMFDTD/EMFDTD.hpp:351-356).
 * eps = 1 gives oracle_yee_step bit for bit. */
void oracle_yee_eps_half_e(int ncell, double s, const double* eps, double* f);
void oracle_yee_eps_step(int ncell, double s, const double* eps, double* f);
/* Discrete adjoint gradient of J = 0.5 sum over updated E entries of
 * (E^N - target)^2 w.r.t. eps; states holds the nsteps+1 states, target the
 * three E arrays.  Writes grad ((N+1)^3), returns J. */
double oracle_yee_eps_grad(int ncell, double s, const double* eps, const double* states, const double* target,
                           int nsteps, double* grad);

/* Batched-ensemble member: 2-D heat, scalar kappa, zero Dirichlet border,
 * n x n nodes (no ghost layer), u_new = u + tau*kappa*lap(u) on inner nodes
 * (lap with 1/h^2, per-axis accumulation as HeatEqnFDM.hpp:209-215). */
void oracle_heat2d_run(int n, double h, double kappa, double tau, int nsteps,
                       double* u, double* u_prev);
/* J = 0.5*sum (u^N - target)^2, dJ/dkappa by explicit reverse sweep. */
double oracle_heat2d_grad(int n, double h, double kappa, double tau, int nsteps,
                          const double* u0, const double* target, double* dj_dkappa);

/* Linf of (a - b) over the nodes of a field (compErrLinf, DataField.hpp:1439-1530). */
double oracle_linf_nodes(const oracle_grid* g, const double* a, const double* b);

/* Threads used by the OpenMP loops (0 = OpenMP default). */
void oracle_set_threads(int n);
int oracle_get_threads(void);

/*---------------------------------------------------------------------------*/
/* TumorHeatEqnFDM example (examples/TumorHeatEqnFDM/TumorHeatEqnFDM.hpp):
 * Pennes bioheat with perfusion and heat source chosen per node by a
 * spherical tumor, Neumann (Q_BC) borders and a Cauchy (H, T_INF) border on
 * the far face of the last axis. */
typedef struct {
    double tau, k, rho, c, rho_b, c_pb;
    double omega_brain, omega_tumor, q_brain, q_tumor, t_a;
    double h, t_inf, q_bc;
    double radius, center[3];
} oracle_tumor_coef;
/* updateInner / updateBorder (TumorHeatEqnFDM.hpp:233-349), one step. */
void oracle_tumor_step(const oracle_grid* g, const double* t_old, double* t_new, const oracle_tumor_coef* c);
/* omega and Q as the hooks select them, at every node of a memAll field
 * (ghosts 0). */
void oracle_tumor_fields(const oracle_grid* g, const oracle_tumor_coef* c, double* omega, double* q);

/*---------------------------------------------------------------------------*/
/* MRIDataPBHEqnFDM example (examples/MRIDataPBHEqnFDM/MRIDataPBHEqnFDM.hpp):
 * Pennes bioheat with per-node material fields mat[0..8] = rho, c, lambda,
 * rho_blood, c_blood, omega, T_blood, q, surface (stencil-0 node blocks),
 * Neumann (Q_BC) borders, and on the far face of the last axis either an
 * open skull (surface == 1: convection H, T_inf and radiation epsilon,
 * sigma in Kelvin) or a closed one (Neumann Q_SKULL). */
typedef struct {
    double tau, h, t_amb, q_bc, q_skull, epsilon, sigma;
} oracle_mri_coef;
/* updateInner / updateBorder (MRIDataPBHEqnFDM.hpp:143-260), one step. */
void oracle_mri_step(const oracle_grid* g, const double* t_old, double* t_new, const double* const* mat,
                     const oracle_mri_coef* c);

/*---------------------------------------------------------------------------*/
/* EMFDTD example: Yee leapfrog with a complete CPML and conductive media
 * (examples/EMFDTD/EMFDTD.hpp, stencils/CompleteCPMLCurlStencil.hpp,
 * CPMLParams.hpp).  Fields are n^3 node blocks, x fastest. */
enum {
    OEM_EX = 0, OEM_EY, OEM_EZ, OEM_HX, OEM_HY, OEM_HZ, OEM_DX, OEM_DY, OEM_DZ, OEM_IX, OEM_IY, OEM_IZ,
    OEM_PSI_EXDY, OEM_PSI_EXDZ, OEM_PSI_EYDX, OEM_PSI_EYDZ, OEM_PSI_EZDX, OEM_PSI_EZDY,
    OEM_PSI_HXDY, OEM_PSI_HXDZ, OEM_PSI_HYDX, OEM_PSI_HYDZ, OEM_PSI_HZDX, OEM_PSI_HZDY,
    OEM_GA, OEM_GB, OEM_NFIELDS
};
typedef struct {
    int n;                       /* nodes per axis (cubic grid) */
    int d;                       /* PML depth: tables have d+1 entries */
    const double *b, *c, *ook;   /* CPML tables, index 0 = outermost layer */
    double cfln;                 /* Courant factor (0.5 in the reference) */
    int src[3], sink[3];         /* node indices of the hard Ez source and the sink */
} oracle_emfdtd;

/* CPMLParams (CPMLParams.hpp:23-49), d >= 2. */
void oracle_cpml_tables(int d, double m_a, double m, double k_max, double s_max, double a_max, double dt,
                        double* b, double* c, double* ook);
/* dt = cfln * ds / C_0 with ds = 1 (EMFDTD.hpp:191-193). */
double oracle_emfdtd_dt(double cfln);
/* ga = 1/(1 + sigma dt/eps0), gb = sigma dt/eps0 per node (EMFDTD.hpp:351-352). */
void oracle_emfdtd_ga_gb(long count, const double* sigma, double dt, double* ga, double* gb);
/* ModulatedGaussianSource (sources/ModulatedGaussianSource.hpp:25-34). */
double oracle_mod_gauss(double t, double dt, double f_mod, double t0, double spread);
/* IndicatorEllipsoid at point (x,y,z) (geometry/IndicatorEllipsoid.hpp:48-68). */
double oracle_ellipsoid(double x, double y, double z, double cx, double cy, double cz, double a, double b,
                        double c, double inside);
/* One leapfrog step in place: H phase, then D/E/I phase with the hard source
 * value at src; *sink (nullable) receives the new Ez at the sink node.
 * f: OEM_NFIELDS arrays (GA/GB read only). */
void oracle_emfdtd_step(const oracle_emfdtd* p, double* const* f, double src_value, double* sink);
/* nsteps steps with the source evaluated at the old field's time
 * t_start + (s-1) (tau = 1); sink: nsteps values (nullable). */
void oracle_emfdtd_run(const oracle_emfdtd* p, double* const* f, int nsteps, double t_start, double f_mod,
                       double t0, double spread, double* sink);

#ifdef __cplusplus
}
#endif
#endif
