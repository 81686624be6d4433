/*
 * scafes_hip.h -- C ABI of the MI355X-native explicit time-stepping engine.
 *
 * The reference (ScaFES 2.4.0) has no C ABI: its hot path is the C++ CRTP
 * engine Problem<OWNPRBLM,CT,DIM> (src/ScaFES_Problem.hpp:264) whose
 * OpenMP node loops call the user hooks.  Its only device code is an empty
 * CUDA stub (src/ScaFES_ProblemCuda.cu:5-37, host wrapper
 * src/ScaFES_ProblemCuda.hpp:84-105).  This header is the slot that stub
 * was meant to fill: plain pointers and sizes, int status codes, no C++
 * types, no exceptions.  The C++11 host API (masterthesis_amd/host/ScaFES)
 * and the Python harness (ctypes) call it.
 *
 * Conventions
 *  - Every function returns SCAFES_OK (0) or a negative code; the text of
 *    the last error of the calling thread is scafes_last_error().
 *  - Device pointers come from scafes_malloc (or any HIP allocation on the
 *    context's device).  All work is ordered on the context's stream;
 *    functions returning host values synchronise that stream.
 *  - One host thread per context.
 *  - Field layout: ScaFES DataField memAll (src/ScaFES_DataField.hpp:622-642,
 *    src/ScaFES_Grid.hpp:955-968): Fortran order (x fastest), extents
 *    n_p + 2*hw on active axes.  The reference's trailing time slot
 *    (DataField.hpp:947-953) is kept by the host, not in device arrays.
 */
#ifndef SCAFES_HIP_H_
#define SCAFES_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCAFES_OK 0
#define SCAFES_ERR_ARG (-1)
#define SCAFES_ERR_HIP (-2)
#define SCAFES_ERR_ALLOC (-3)
#define SCAFES_ERR_UNSUPPORTED (-4)

typedef struct scafes_ctx scafes_ctx;

/* Arithmetic of a field.  F32_KAHAN: FP32 values plus an FP32 compensation
 * field that carries the rounding of the accumulated per-step increments
 * (value ~ hi + lo). */
enum { SCAFES_F32 = 0, SCAFES_F64 = 1, SCAFES_F32_KAHAN = 2 };

/* Border rule at GLOBAL_BORDER nodes (GridGlobal.hpp:138-151 flags):
 * KEEP  = the hook leaves the copied old value (Problem.hpp:6226-6256),
 * FIELD = Dirichlet from a known field G_new (Dirichlet.hpp:124-130),
 * VALUE = Dirichlet from a constant,
 * NEUMANN / CAUCHY / RADIATION = the flux border classes of the system tests
 *   (tests/system/HeatEqnFDM/Neumann/Neumann.hpp:126-168,
 *    tests/system/HeatEqnFDM/Cauchy/Cauchy.hpp:126-170,
 *    tests/system/PennesBioheatEqnFDM/Radiation/Radiation.hpp:126-175):
 *   a border node with index 0 on an active axis takes G_new; otherwise, on
 *   each axis where its index is n_p - 1, the outer neighbour is replaced by
 *   the inner one W and a flux term from G_old (A = 2 h_p / lambda):
 *     NEUMANN   (W + W) - A*G_old - 2u
 *     CAUCHY    ((W + W) - A*alpha*u) + A*G_old - 2u
 *     RADIATION ((W + W) - A*eps*sigma*u*u*u*u) + A*G_old - 2u
 *   (central differences on the other axes, then perfusion and source as
 *   on inner nodes).  heat7 only. */
enum {
    SCAFES_BORDER_KEEP = 0,
    SCAFES_BORDER_FIELD = 1,
    SCAFES_BORDER_VALUE = 2,
    SCAFES_BORDER_NEUMANN = 3,
    SCAFES_BORDER_CAUCHY = 4,
    SCAFES_BORDER_RADIATION = 5,
    SCAFES_BORDER_TUMOR = 6   /* Pennes only: see scafes_pennes_params */
};

/* Discrete domain of one partition covering the whole grid
 * (GridGlobal::discreteDomain(), GridGlobal.hpp:384-422). */
typedef struct {
    int dim;       /* 1..3 */
    int n[3];      /* nodes per axis; 1 on axes >= dim */
    int nlayers;   /* nLayersAtBorder (Parameters.hpp:446, default 1) */
    double h[3];   /* gridsize (Grid.hpp:547-555) */
} scafes_geom;

/* ---- context, errors, memory (Problem owns its lumps: Problem.hpp:1215-1275) */
const char* scafes_last_error(void);
int scafes_device_count(int* n);
int scafes_ctx_create(int device, scafes_ctx** out);
int scafes_ctx_destroy(scafes_ctx* ctx);
/* The gradient entry points (scafes_*_gradient) keep their workspace --
 * the forward and adjoint states and the snapshot pool, up to the largest
 * call's size -- in the context for the next call, since re-allocating
 * freed device memory makes the driver clear it first (about 0.5 s per
 * 137 GB store-all pool).  This frees it (scafes_ctx_destroy does too).  No
 * counterpart in the reference, which keeps its ADOL-C tapes per Problem. */
int scafes_ctx_release_workspace(scafes_ctx* ctx);
int scafes_ctx_sync(scafes_ctx* ctx);
int scafes_ctx_info(scafes_ctx* ctx, int* cu_count, char* name, int name_len);
/* The context's HIP stream (a hipStream_t, default priority, non-blocking)
 * for kernels the caller launches itself in stream order with the
 * library's: the generic device update phase of a hipcc-compiled problem
 * (the role of the reference's per-node hook loop, Problem.hpp:5443-5466). */
int scafes_ctx_stream(scafes_ctx* ctx, void** stream);
/* Device-side wall clock on the context's stream (ScaFES::Timer restart/
 * elapsed, src/ScaFES_Timer.hpp:171-214, for GPU work): record event slot
 * 0..7, then read the elapsed milliseconds between two recorded slots. */
int scafes_event_record(scafes_ctx* ctx, int slot);
int scafes_event_elapsed(scafes_ctx* ctx, int slot_start, int slot_end, float* ms);
int scafes_malloc(scafes_ctx* ctx, size_t bytes, void** dptr);
int scafes_free(scafes_ctx* ctx, void* dptr);
int scafes_memcpy_h2d(scafes_ctx* ctx, void* dst, const void* src, size_t bytes);
int scafes_memcpy_d2h(scafes_ctx* ctx, void* dst, const void* src, size_t bytes);
int scafes_memcpy_d2d(scafes_ctx* ctx, void* dst, const void* src, size_t bytes);
/* Fill count elements of dtype with value (DataField::operator=(CT),
 * DataField.hpp:1137-1145); a kernel, never a memset node. */
int scafes_fill(scafes_ctx* ctx, void* dptr, int dtype, size_t count, double value);
/* Convert between f64 and f32 device arrays (count elements). */
int scafes_convert(scafes_ctx* ctx, void* dst, int dst_dtype, const void* src, int src_dtype,
                   size_t count);

/* ---- constant-coefficient heat / Pennes system update ---------------------
 * Replaces the update phase evalTypeTwoAt (Problem.hpp:5397-5476) for the
 * hooks HeatEqnFDM::updateInner (tests/system/HeatEqnFDM/HeatEqnFDM.hpp:204-218),
 * PennesBioheatEqnFDM::updateInner (PennesBioheatEqnFDM.hpp:153-168) and
 * Dirichlet::updateBorder (Dirichlet.hpp:124-130), FP64 in the reference's
 * operation order (bit-identical to the CPU path).
 *   inner:  u' = u + sum_p (tau*a*(W+E-2u))/h_p^2 [- ((tau*perf_c)*perf_w)*u]
 *                  [+ (tau*rhoc_inv)*F_old]
 * u_old/u_new: memAll with hw ghost layers; f_old/g_old/g_new: stencil-0
 * fields (NULL when unused; g_old is read by the flux border rules only).
 * Ghost layers of u_new are not written.
 * Stability: returns SCAFES_ERR_ARG unless
 *   tau * (sum_p 2*coeff_a/h_p^2 [+ perf_c*perf_w]) <= 1,
 * the positivity bound of the explicit step that the reference's drivers
 * check (tests/system/HeatEqnFDM/Dirichlet/TimeLinSpaceLin/main3D.cpp:52-55);
 * the border flux terms are not part of the bound. */
typedef struct {
    double tau, coeff_a, rhoc_inv, perf_c, perf_w;
    int has_perf, has_source;
    int border_mode;
    double border_value;
    double lambda, alpha, epsilon, sigma;  /* flux border coefficients */
} scafes_heat7_params;
int scafes_heat7_step_f64(scafes_ctx* ctx, const scafes_geom* g, int hw, const double* u_old,
                          double* u_new, const double* f_old, const double* g_old,
                          const double* g_new, const scafes_heat7_params* p);
/* nsteps of the same update without known fields (no source, border KEEP
 * or VALUE and the other rules that read no G), ping-ponging u -> u2 ->
 * u ...: the loop of iterateOverTime (Problem.hpp:2558-2715) over the
 * update phase for problems whose known fields do not enter the update,
 * one launch per step and no host work in between.  *result_in_second =
 * 1 when the final state is in u2.  Bit-identical to nsteps calls of
 * scafes_heat7_step_f64. */
int scafes_heat7_run_f64(scafes_ctx* ctx, const scafes_geom* g, int hw, double* u, double* u2, int nsteps,
                         const scafes_heat7_params* p, int* result_in_second);

/* ---- variable-coefficient conservative 7-point heat -----------------------
 *   inner: u' = u + sum_p tau*(0.5(kE+kC)(uE-uC) - 0.5(kC+kW)(uC-uW))/h_p^2
 * (node-centred hook structure of HeatEqnFDM.hpp:209-215 with a known
 * coefficient field as in examples/MRIDataPBHEqnFDM/MRIDataPBHEqnFDM.hpp:143-168).
 * mode: SCAFES_HEATVAR_F64 (bit-identical to the FP64 CPU path),
 *       SCAFES_HEATVAR_F32 (plain FP32 stream),
 *       SCAFES_HEATVAR_F32_KAHAN (FP32 stream with a persistent FP32
 *       compensation field c: u_true ~ u + c).
 * Runs nsteps steps ping-ponging between (u, u2) and (c, c2); on return
 * *result_in_second = 1 when the final state is in u2/c2.  kappa has the
 * layout of u.  cfg: tile configuration index (-1 = autotuned/default).
 * A multi-step configuration given as cfg runs nsteps / s sweeps of s
 * steps, a remainder of two as one two-step sweep, the rest one step per
 * sweep.  With cfg = -1 the steps left behind three-step sweeps go to one
 * two-step sweep, or, where the mode has a four-step tile of the same cost
 * per step (FP32 + Kahan), to four-step sweeps (7 = 4 + 3, 20 = 2 x 4 +
 * 4 x 3); a tuned four-step tile takes its remainder with the three-, two-
 * and one-step tiles.  Stability: SCAFES_ERR_ARG unless tau * sum_p 2*kappa_max/h_p^2 <= 1
 * (kappa_max = max |kappa| over all nodes, a device reduction kept per
 * pointer until the library writes the field or scafes_coef_changed names
 * it). */
enum { SCAFES_HEATVAR_F64 = 0, SCAFES_HEATVAR_F32 = 1, SCAFES_HEATVAR_F32_KAHAN = 2 };
int scafes_heat7var_run(scafes_ctx* ctx, const scafes_geom* g, int hw, int mode, void* u,
                        void* u2, void* c, void* c2, const void* kappa, double tau, int nsteps,
                        int border_mode, double border_value, int cfg, int* result_in_second);
/* Number of tile configurations and a description of one: workgroup tile,
 * z planes per workgroup, prefetch depth in planes, and time steps per HBM
 * sweep (1, or 2 / 3 / 4 for the temporally blocked kernels). */
int scafes_heat7var_num_cfgs(void);
int scafes_heat7var_cfg_desc(int cfg, int* tx, int* ty, int* zchunk, int* depth,
                             int* steps_per_sweep);
/* The configuration scafes_heat7var_run uses for cfg = -1 on this grid
 * (autotuned choice, else the per-mode default; the two-steps-per-sweep
 * tiles need a 3-D grid). */
int scafes_heat7var_active_cfg(scafes_ctx* ctx, const scafes_geom* g, int hw, int mode, int* cfg);
/* How scafes_heat7var_run with cfg = -1 splits nsteps into sweeps on this
 * grid: the steps of each sweep in launch order (the first min(max, count)
 * written to sweeps), their number in *count.  No counterpart in the
 * reference, which updates one step per kernel launch
 * (src/ScaFES_ProblemCuda.cu); callers that time the sweep kinds
 * separately use it. */
int scafes_heat7var_plan(scafes_ctx* ctx, const scafes_geom* g, int hw, int mode, int nsteps, int* sweeps,
                         int max, int* count);
/* Project autotuner (no counterpart in the reference, whose device stub
 * fixes BLOCK_SIZE 32, src/ScaFES_ProblemCuda.hpp:22): times every
 * configuration on the given buffers over 24*ceil(nsteps/24) steps (a
 * multiple of every configuration's steps per sweep, so no remainder sweep
 * enters the timing) and keeps the fastest for this grid (extents, hw) and
 * mode: later calls with cfg = -1 on the same grid use it (at most eight
 * grids per context).  The buffers are overwritten: u/u2 (and c/c2) are
 * advanced with border KEEP, an even number of sweeps per configuration. */
int scafes_autotune_heat7var(scafes_ctx* ctx, const scafes_geom* g, int hw, int mode, void* u,
                             void* u2, void* c, void* c2, const void* kappa, double tau,
                             int nsteps, int* best_cfg, float* best_ms);
/* Synthetic initial state of the 512^3 benchmark workload:
 * u = prod sin(pi x_p) + amp*(2*U-1), U = splitmix64(seed, node) / 2^53 in
 * [0,1) on inner nodes (border nodes: the sine product), and
 * kappa = 1 + 0.5 sin(2 pi x) sin(2 pi y).  dtype of u and kappa. */
int scafes_heat7var_init(scafes_ctx* ctx, const scafes_geom* g, int hw, int dtype, void* u,
                         void* kappa, uint64_t seed, double amp);

/* ---- known fields on the device ---------------------------------------
 * Replaces evalKnownDfs (Problem.hpp:3433-3485) for the TimeLinSpaceLin
 * system problems, whose evalInner/evalBorder hooks are closed forms of
 * (x, t): heat tests/system/HeatEqnFDM/Dirichlet/TimeLinSpaceLin/
 * TimeLinSpaceLin.hpp:75-123, Pennes tests/system/PennesBioheatEqnFDM/
 * Dirichlet/TimeLinSpaceLin/TimeLinSpaceLin.hpp:75-124, and the flux-border
 * variants .../{Neumann,Cauchy,Radiation}/TimeLinSpaceLin/TimeLinSpaceLin.hpp.
 * With s = 1 + sum_p x_p (x_p = x0_p + h_p i_p, Grid.hpp:719-733):
 *   inner:  F = s*rho_c [+ rhob_cb_w*(s*(1+t)) for the Pennes form], G = 0
 *   border, FIELD (Dirichlet): F = 0, G = s*(1+t)
 *   border, flux modes: F as inner; G = s*(1+t) on nodes with an index 0,
 *     else NEUMANN (-lambda)(1+t); CAUCHY s*(alpha(1+t)) + lambda(1+t);
 *     RADIATION (s(1+t))^4*(sigma*epsilon) + lambda(1+t)
 *   Y (exact solution) = s*(1+t) everywhere.
 * F, G: stencil-0 FP64 device arrays (x fastest); Y: FP64 memAll with hw_y
 * ghost layers (0 for a stencil-0 array; the unknown field's width when Y
 * is its reference solution for compErrLinf); each may be NULL, ghosts are
 * not written.  Bit-identical to evaluating the hooks on the host. */
enum { SCAFES_KNOWN_HEAT_TIMELIN = 1, SCAFES_KNOWN_PENNES_TIMELIN = 2 };
typedef struct {
    int form;          /* SCAFES_KNOWN_HEAT_TIMELIN or SCAFES_KNOWN_PENNES_TIMELIN */
    int border_mode;   /* SCAFES_BORDER_FIELD / _NEUMANN / _CAUCHY / _RADIATION */
    double x0[3];      /* coordNodeFirst */
    double rho_c;      /* RHO*C */
    double rhob_cb_w;  /* RHO_BLOOD*C_BLOOD*W (Pennes form) */
    double lambda, alpha, epsilon, sigma;
} scafes_timelin_params;
int scafes_known_timelin_f64(scafes_ctx* ctx, const scafes_geom* g, const scafes_timelin_params* p,
                             double t, double* F, double* G, double* Y, int hw_y);

/* ---- reductions (DataField::compErrLinf DataField.hpp:1439-1530 and
 *      DataField::checkConv DataField.hpp:1532-1636) ---------------------
 * a has hw_a ghost layers, b has hw_b (0 for a stencil-0 known field).
 * op: 0 = max|a-b| over nodes, 1 = sum (a-b)^2 over nodes, 2 = sum a*b.
 * inner_only: restrict to GLOBAL_INNER nodes.  Result in FP64 on host. */
int scafes_reduce(scafes_ctx* ctx, const scafes_geom* g, int dtype, const void* a, int hw_a,
                  const void* b, int hw_b, int op, int inner_only, double* out);
/* Tells the stability checks that the coefficient field at p was written
 * outside the library (by a kernel of the caller); the next step re-reads
 * its maximum.  Writes through scafes_memcpy_*, scafes_fill, scafes_convert
 * and scafes_free are seen without it. */
int scafes_coef_changed(scafes_ctx* ctx, const void* p);

/* ---- Pennes bioheat with perfusion / source fields ---------------------
 * TumorHeatEqnFDM::updateInner form (examples/TumorHeatEqnFDM/TumorHeatEqnFDM.hpp:233-269)
 * with omega, Q as node fields:
 *   T' = T + sum_p (tau*cdiff*(W+E-2T))/h_p^2 + ((tau*cperf)*w)*(T_a-T) + (tau*csrc)*Q.
 * border_mode KEEP (0): border nodes keep their value.  TUMOR: the
 * example's updateBorder (TumorHeatEqnFDM.hpp:272-349): on each axis p, with
 * A = (2 h_p)/k_cond, a node with index n_p - 1 replaces the missing
 * neighbour by the inner one and a flux term,
 *   last axis (p = dim-1), Cauchy:  ((W + W) - (A*h_conv)*(T - t_inf)) - 2T
 *   other axes, Neumann:            ((W + W) - A*(-q_bc)) - 2T,
 * a node with index 0 (Neumann)      ((E + A*q_bc) + E) - 2T,
 * other axes central; then perfusion and source as on inner nodes.
 * dtype F64: bit-identical to the FP64 CPU path; F32: FP32 stream.  The
 * adjoint and gradient entry points take KEEP and TUMOR. */
typedef struct {
    double tau, cdiff, cperf, csrc, t_a;
    int border_mode;                      /* SCAFES_BORDER_KEEP or SCAFES_BORDER_TUMOR */
    double k_cond, h_conv, t_inf, q_bc;   /* K, H, T_INF, Q_BC of the TUMOR border */
} scafes_pennes_params;
int scafes_pennes_step(scafes_ctx* ctx, const scafes_geom* g, int hw, int dtype,
                       const void* t_old, void* t_new, const void* w, const void* q,
                       const scafes_pennes_params* p);
/* Stability of scafes_pennes_step(_kahan) and scafes_pennes_gradient
 * (examples/TumorHeatEqnFDM/startSimulation.py:238-301): SCAFES_ERR_ARG
 * unless tau * (sum_p 2*cdiff/h_p^2 + cperf*w_max) <= 1, with the
 * convective face's 2*cdiff*h_conv/(k_cond*h_last) added for the TUMOR
 * border; w_max = max |w| (cached per pointer as for heat7var). */
/* FP32 step with Kahan-compensated accumulation: (t_old + c_old) ->
 * (t_new + c_new); the increment is formed from t_old. */
int scafes_pennes_step_kahan(scafes_ctx* ctx, const scafes_geom* g, int hw, const float* t_old,
                             float* t_new, const float* c_old, float* c_new, const float* w,
                             const float* q, const scafes_pennes_params* p);
/* Discrete adjoint step (replaces ADOL-C fov_forward, Problem.hpp:6632-6747):
 *   lam'_i = lam_i + sum_p tau*cdiff*(lam_W+lam_E-2lam_i)/h_p^2 - tau*cperf*w_i*lam_i
 *   grad_i += lam_i * tau*cperf*(T_a - t_prev_i)          (inner nodes; FP64 grad)
 * lam' = 0 on border nodes (KEEP).  TUMOR: every node is updated with the
 * transpose of the Tumor step's linear part: on axis p, lam_W counts twice
 * where the W neighbour has index 0 and lam_E twice where the E neighbour
 * has index n_p - 1, a neighbour beyond the grid counts 0, and a node on
 * the far face of the last axis adds -(2 h_p/k_cond)*h_conv*lam_i (Cauchy);
 * the gradient term is taken on every node. */
int scafes_pennes_adjoint_step(scafes_ctx* ctx, const scafes_geom* g, int hw, int dtype,
                               const void* lam, void* lam_new, const void* w,
                               const void* t_prev, double* grad, const scafes_pennes_params* p);
/* Gradient of J = 0.5 sum_inner (T^N - target)^2 w.r.t. the perfusion field
 * w by the reverse-time adjoint with binomial checkpointing (Griewank's
 * treeverse schedule) using at most max_snapshots stored states.
 * border_mode TUMOR (the Tumor example's Neumann/Cauchy borders): J sums
 * over every node, since every node is updated, and the gradient covers
 * every node.
 * dtype F64, F32, or F32_KAHAN (FP32 fields, compensated forward whose
 * checkpoints store (hi, lo) pairs; the adjoint field is plain FP32).  For
 * FP32 modes t0, w, q, target are FP32 (t0 without compensation).
 * t0 is read only; grad (FP64, layout of T) is overwritten (ghost layers
 * zero).  stats (nullable, 4 entries): forward steps executed, adjoint
 * steps, snapshots used, checkpoint bytes. */
int scafes_pennes_gradient(scafes_ctx* ctx, const scafes_geom* g, int hw, int dtype,
                           const void* t0, const void* w, const void* q, const void* target,
                           const scafes_pennes_params* p, int nsteps, int max_snapshots,
                           double* grad, double* J, double* stats);
/* Number of forward steps the treeverse schedule runs for (nsteps, snaps). */
long scafes_checkpoint_forward_steps(int nsteps, int snaps);
/* Gradient of J = 0.5 sum_inner (u^N - target)^2 w.r.t. the kappa field of
 * the variable-kappa heat (the scafes_heat7var_run step, border KEEP or VALUE)
 * by the reverse-time discrete adjoint with the same checkpointing as
 * scafes_pennes_gradient.  No counterpart in the reference, whose gradients
 * are ADOL-C tapes of its hooks w.r.t. a few scalar parameters
 * (Problem.hpp:5841-5967, 6632-6747); this is the field gradient of the
 * conservative var-kappa update:
 *   lam^N = u^N - target on inner nodes, 0 on border nodes; for n = N..1
 *   grad_k += 0.5 sum_p (tau/h_p^2) sum_{nb = k -+ e_p in the grid}
 *             (lam_k - lam_nb)(u^{n-1}_nb - u^{n-1}_k)     (every node)
 *   lam^{n-1} = the step's operator applied to lam^n, 0 on border nodes.
 * mode SCAFES_HEATVAR_F64: bit-identical to the FP64 CPU restatement;
 * F32 and F32_KAHAN: u0, kappa, target FP32 (u0 without compensation), the
 * adjoint field FP32, the gradient FP64 (the compensated forward's lo parts
 * do not enter the adjoint).  grad has the layout of u (layout hw) and is
 * overwritten; stats as for scafes_pennes_gradient. */
int scafes_heat7var_gradient(scafes_ctx* ctx, const scafes_geom* g, int hw, int mode, const void* u0,
                             const void* kappa, const void* target, double tau, int border_mode,
                             double border_value, int nsteps, int max_snapshots, double* grad, double* J,
                             double* stats);

/* ---- Pennes bioheat with per-node materials (MRI data) ------------------
 * Replaces the update phase of examples/MRIDataPBHEqnFDM (MRIDataPBHEqnFDM.hpp:
 * 143-273, fields of main3D.cpp:17-31): nine known material fields (stencil
 * width 0, node layout hw = 0) and the temperature T (layout hw).  Per node,
 * with rc = rho*c:
 *   T' = T + sum_p ((tau*(lambda/rc)) * d_p) / h_p^2
 *          + ((tau*((rho_b*c_b)/rc))*omega) * (T_blood - T) + (tau*(1/rc))*q,
 * d_p central on inner axes.  On border nodes, with A = (2 h_p)/lambda:
 *   index 0 (Neumann)                 d_p = ((E + A*q_bc) + E) - 2T
 *   index n_p-1, p < dim-1 (Neumann)  d_p = ((W + W) - A*(-q_bc)) - 2T
 *   index n_p-1, p = dim-1, (int)surface == 1 (open skull):
 *     d_p = (((W + W) - (A*h_conv)*(T - t_amb)) - ((A*epsilon)*sigma)*(Tr^4 - Ta^4)) - 2T
 *     with Tr = (T + the axes before p's increments) + 273.15 (the hook reads
 *     vNew there, MRIDataPBHEqnFDM.hpp:200) and Ta = t_amb + 273.15;
 *   otherwise (closed skull)          d_p = ((W + W) - A*(-q_skull)) - 2T.
 * FP64, bit-identical to the reference's CPU path. */
enum {
    SCAFES_MRI_RHO = 0,
    SCAFES_MRI_C = 1,
    SCAFES_MRI_LAMBDA = 2,
    SCAFES_MRI_RHO_BLOOD = 3,
    SCAFES_MRI_C_BLOOD = 4,
    SCAFES_MRI_OMEGA = 5,
    SCAFES_MRI_T_BLOOD = 6,
    SCAFES_MRI_Q = 7,
    SCAFES_MRI_SURFACE = 8,
    SCAFES_MRI_NFIELDS = 9,
    SCAFES_MRI_NCOEF = 4
};
typedef struct {
    double tau;
    double h_conv, t_amb;     /* H, T_AMB: open-skull convection */
    double q_bc, q_skull;     /* Q_BC (brain faces), Q_SKULL (closed skull) */
    double epsilon, sigma;    /* EPSILON, SIGMA (5.670373e-8): open-skull radiation */
} scafes_mri_params;
/* Folds the materials mat[0..7] (device pointers, node layout) for p->tau
 * into SCAFES_MRI_NCOEF per-node coefficient fields, each laid out like T
 * (layout hw), stored back to back in coef (4 * memAll(hw) doubles).  The
 * materials are time-independent, so this runs once per run (again if tau or
 * a material changes). */
int scafes_mri_coef_f64(scafes_ctx* ctx, const scafes_geom* g, int hw, const double* const* mat,
                        const scafes_mri_params* p, double* coef);
/* One step t_old -> t_new (out of place).  mat is the host array of the nine
 * material field pointers (the step reads lambda and surface on border
 * nodes); coef comes from scafes_mri_coef_f64 with the same tau. */
int scafes_mri_step_f64(scafes_ctx* ctx, const scafes_geom* g, int hw, const double* t_old, double* t_new,
                        const double* const* mat, const double* coef, const scafes_mri_params* p);

/* ---- Yee FDTD, PEC box, vacuum ------------------------------------------
 * Leapfrog of the reference's two-phase mode (updateInner then
 * updateInner2, Problem.hpp:2603-2619; curl stencil of
 * examples/EMFDTD/stencils/CompleteCPMLCurlStencil.hpp:38-80 without CPML):
 *   H += s*(forward differences of E);  E += s*(backward differences of H);
 * tangential E stays zero on the walls.  f: six fields Ex,Ey,Ez,Hx,Hy,Hz of
 * dtype (F32 stream, or F64 in the CPU path's operation order), each
 * (ncell+1)^3 nodes, x fastest, updated in place over nsteps steps.
 * s = Courant number tau/h in (0, 1/sqrt(3)].
 * f2: NULL, or a second state buffer of the same size; with f2 each step is
 * one out-of-place sweep doing both half steps (6 field reads + 6 writes
 * per cell instead of 9 + 6 for the two in-place half-step sweeps) and the
 * state ping-pongs between f and f2: on return *result_in_second (nullable)
 * is 1 when the final state is in f2.  Both schedules give identical bits.
 * energy (host FP64 array of ceil(nsteps/energy_every) entries, used when
 * energy_every > 0): discrete energy 0.5 sum E^n.E^n + 0.5 sum
 * H^{n-1/2}.H^{n+1/2} of steps 0, energy_every, ..., accumulated in FP64
 * inside the H half step. */
int scafes_yee_run(scafes_ctx* ctx, int ncell, int dtype, double s, void* f, void* f2, int nsteps,
                   int energy_every, double* energy, int* result_in_second);
/* scafes_yee_run without any allocation, copy or wait: no counterpart in
 * the reference; for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels and results as scafes_yee_run (bit for bit), except that the
 * energy records go to the caller's DEVICE array energy_dev
 * (ceil(nsteps/energy_every) FP64 entries, written in stream order, not
 * waited for).  The per-workgroup energy partials live in a static 1 MiB
 * device array per GPU (2^17 doubles, two per workgroup of a launch: boxes
 * up to about 1150 cells per axis); a larger box is refused with
 * SCAFES_ERR_ARG before anything runs.  Nothing in the array needs clearing:
 * every workgroup writes its slot before the slots are summed.  All calls on
 * one GPU share that array: calls that overlap on one GPU (two contexts or
 * streams at once) race and must be ordered by the caller. */
int scafes_yee_run_noalloc(scafes_ctx* ctx, int ncell, int dtype, double s, void* f, void* f2, int nsteps,
                           int energy_every, double* energy_dev, int* result_in_second);
/* Initial state of the Yee workload: all fields zero except
 * Ez = exp(-r^2/(2 sigma^2)) at inner Ez nodes, r from the box centre in
 * cell units (Ez(i,j,k) sits at (i, j, k+1/2)). */
int scafes_yee_init_pulse(scafes_ctx* ctx, int ncell, int dtype, void* f, double sigma);
/* Yee leapfrog with a relative permittivity field eps ((N+1)^3 values in
 * the state dtype, one per node index m, shared by Ex(m), Ey(m), Ez(m);
 * eps > 0): the E half step becomes D += s curl H, E = D / eps, i.e.
 * E += (s / eps) * (backward differences of H) -- the per-node D -> E
 * scaling of the reference's EMFDTD update (examples/EMFDTD/EMFDTD.hpp:351-356,
 * there 1/(1 + sigma dt/eps0)).  eps = 1 gives scafes_yee_run's bits.
 * f (six arrays) is advanced nsteps in place; work is a second state
 * buffer (out-of-place sweeps: two steps per sweep, a single fused step for
 * an odd remainder).  F32 or F64; 2 <= ncell <= 1400. */
int scafes_yee_eps_run(scafes_ctx* ctx, int ncell, int dtype, double s, const void* eps, void* f, void* work,
                       int nsteps);
/* Gradient of J = 0.5 sum over the E entries the step updates of
 * (E^N - target)^2 (target: three E arrays) w.r.t. eps, by the reverse-time
 * discrete adjoint of scafes_yee_eps_run's step with the checkpointing of
 * scafes_pennes_gradient (max_snapshots stored states of six arrays).  No
 * counterpart in the reference (its EMFDTD optimisation tapes a few
 * geometry parameters with ADOL-C, Problem.hpp:5841-5967).  The adjoint is
 * the pair of half steps run backwards:
 *   mu_H = lam_H - Hform(ce lam_E),  lam_E <- lam_E - s Eform(mu_H),
 *   grad(m) -= (ce/eps)(m) sum_c lam_E_c(m) Eform_c(H^{n-1/2})(m),  ce = s/eps.
 * F64: bit-identical to the FP64 CPU restatement; F32: f0, eps, target and
 * the adjoint fields FP32, grad FP64.  grad ((N+1)^3 FP64) is overwritten;
 * stats as for scafes_pennes_gradient. */
int scafes_yee_eps_gradient(scafes_ctx* ctx, int ncell, int dtype, double s, const void* f0, const void* eps,
                            const void* target, int nsteps, int max_snapshots, double* grad, double* J,
                            double* stats);

/* ---- EMFDTD example: Yee leapfrog + complete CPML + conductive media -----
 * Replaces the update phases of examples/EMFDTD/EMFDTD.hpp (EMFDTDng:
 * updateInner = H from E, :295-313; updateInner2 = D from H, E = ga (D - I),
 * I += gb E, hard Ez source, Ez sink, :330-386) with the CPML curl stencil
 * examples/EMFDTD/stencils/CompleteCPMLCurlStencil.hpp:38-80.  FP64, in the
 * reference's operation order (bit-identical to its CPU path).
 * State: SCAFES_EM_NFIELDS arrays of n^3 doubles (x fastest, no ghosts) in
 * the order below, one allocation; ga/gb are per-node coefficients from
 * sigma (scafes_emfdtd_ga_gb) and are read only.  The reference requires a
 * cubic grid and tau = 1 (time counts steps). */
#define SCAFES_EMFDTD_MAX_D 64
enum {
    SCAFES_EM_EX = 0, SCAFES_EM_EY, SCAFES_EM_EZ, SCAFES_EM_HX, SCAFES_EM_HY, SCAFES_EM_HZ,
    SCAFES_EM_DX, SCAFES_EM_DY, SCAFES_EM_DZ, SCAFES_EM_IX, SCAFES_EM_IY, SCAFES_EM_IZ,
    SCAFES_EM_PSI_EXDY, SCAFES_EM_PSI_EXDZ, SCAFES_EM_PSI_EYDX, SCAFES_EM_PSI_EYDZ,
    SCAFES_EM_PSI_EZDX, SCAFES_EM_PSI_EZDY, SCAFES_EM_PSI_HXDY, SCAFES_EM_PSI_HXDZ,
    SCAFES_EM_PSI_HYDX, SCAFES_EM_PSI_HYDZ, SCAFES_EM_PSI_HZDX, SCAFES_EM_PSI_HZDY,
    SCAFES_EM_GA, SCAFES_EM_GB, SCAFES_EM_NFIELDS
};
typedef struct {
    int n;          /* nodes per axis */
    int d;          /* PML layers (CPMLParams D), 2..SCAFES_EMFDTD_MAX_D */
    double cfln;    /* Courant factor of the curl updates (EMFDTD.hpp:191: 0.5) */
    double dt;      /* seconds per step, scafes_emfdtd_dt(cfln) */
    double b[SCAFES_EMFDTD_MAX_D + 1], c[SCAFES_EMFDTD_MAX_D + 1], ook[SCAFES_EMFDTD_MAX_D + 1];
    int src[3], sink[3];         /* node indices (i, j, k) of the source and the sink */
    double f_mod, t0, spread;    /* ModulatedGaussianSource parameters */
    int flags;                   /* SCAFES_EMFDTD_* below, 0 for a state of unknown origin */
} scafes_emfdtd_params;
/* flags: PSI_SETTLED = the state comes from a previous run with the same
 * tables, so every interior psi is +-0 and the first step needs no full
 * psi pass; VACUUM_OK = scafes_emfdtd_vacuum_ok reported 1 for this state
 * (every gb = 0 node has ga = 1 and I = 0), so the check is skipped.  Both
 * are exact shortcuts; the caller vouches for them. */
#define SCAFES_EMFDTD_PSI_SETTLED 1
#define SCAFES_EMFDTD_VACUUM_OK 2
/* CPMLParams (examples/EMFDTD/CPMLParams.hpp:23-49): d+1 entries, index 0 =
 * outermost layer, index d = interior (b = c = 0, 1/kappa = 1).  Host only. */
int scafes_emfdtd_tables(int d, double m_a, double m, double k_max, double s_max, double a_max,
                         double dt, double* b, double* c, double* ook);
/* dt = cfln * ds / C_0, ds = 1 (EMFDTD.hpp:191-193). */
double scafes_emfdtd_dt(double cfln);
/* ga = 1/(1 + sigma dt/eps0), gb = sigma dt/eps0 on host arrays, evaluated
 * like EMFDTD.hpp:351-352 (long double, EPSILON_0 being a long double). */
int scafes_emfdtd_ga_gb(size_t count, const double* sigma, double dt, double* ga, double* gb);
/* ModulatedGaussianSource (sources/ModulatedGaussianSource.hpp:25-34). */
double scafes_emfdtd_source(double t, double dt, double f_mod, double t0, double spread);
/* nsteps leapfrog steps in place.  Step s sets Ez at the source node to the
 * source at the old field's time t_start + (s - 1) (EMFDTD.hpp:370-371) and,
 * when sink != NULL (device, nsteps doubles), stores the new Ez at the sink
 * node in sink[s - 1] (EMFDTD.hpp:376-380). */
int scafes_emfdtd_run(scafes_ctx* ctx, const scafes_emfdtd_params* p, double* state, int nsteps,
                      double t_start, double* sink);
/* *ok = 1 when every node with gb = 0 has ga = 1 and Ix = Iy = Iz = 0 (the
 * condition for the vacuum shortcut; stays true under stepping). */
int scafes_emfdtd_vacuum_ok(scafes_ctx* ctx, const scafes_emfdtd_params* p, const double* state, int* ok);

/* ---- batched ensemble of 2-D heat problems -----------------------------
 * The reference runs one problem per process; this runs nprob independent
 * n_i x n_i problems (unit square, h_i = 1/(n_i-1), zero Dirichlet border,
 * scalar kappa_i and tau_i, the 2-D HeatEqnFDM update) for nsteps explicit
 * FP32 steps each in ONE persistent kernel launch.  Problem i's state is
 * n_i^2 floats (x fastest) at u + off[i], updated in place; 2 <= n_i <= 256
 * (scafes_batch2d_lds_limit).  order: processing order, best sorted by
 * descending ceil(n_i/32)^2 (workgroups take it in snake order).
 * With target != NULL (same packing as u) also, in FP64,
 * J_i = 0.5 sum_inner (u_i^N - target_i)^2 and dJ_i/dkappa_i
 * = N tau_i <u_i^N - target_i, L u_i^{N-1}> (exact discrete adjoint for a
 * scalar kappa with fixed tau: the step operator commutes with L).
 * n, off, kappa, tau, order, u, target, J, grad: device arrays.
 * use_graph != 0 captures the launch into a hipGraph the first time it sees
 * an argument set and replays the instantiated graph on later calls with the
 * same arguments (the context keeps the last one).  Both paths return
 * without waiting (stream order); the sizes are checked on the host first. */
int scafes_batch2d_run(scafes_ctx* ctx, int nprob, const int* n, const long* off,
                       const float* kappa, const float* tau, const int* order, float* u,
                       const float* target, double* J, double* grad, int nsteps, int use_graph);
int scafes_batch2d_lds_limit(int* max_n);

#ifdef __cplusplus
}
#endif
#endif
