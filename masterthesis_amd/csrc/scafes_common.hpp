// scafes_common.hpp -- shared device/host helpers of the HIP stencil library.
//
// Field layout (ScaFES DataField, src/ScaFES_DataField.hpp:622-642 and
// src/ScaFES_Grid.hpp:955-968): Fortran order, x fastest, memAll extents
// n_p + 2*hw on active axes.  Kernels address node (i,j,k) as
//   base + i + e0*j + e01*k,   base = hw + e0*(hw + e1*hw)  (active axes).
// Stencil-0 fields (known dfs) use the same formula with hw = 0.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "../../include/scafes_hip.h"

namespace scafes {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int check_hip(hipError_t e, const char* what);

#define SCAFES_HIP(call)                                        \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) return ::scafes::check_hip(e_, #call); \
    } while (0)

// ---------------------------------------------------------------- context
struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    double* red_partial = nullptr;  // per-block reduction partials (device)
    double* red_result = nullptr;   // final reduction values (device, 8 slots)
    double* host_result = nullptr;  // pinned host mirror of red_result
    int red_capacity = 0;           // number of partial slots
    int cu_count = 0;
    // autotuned heat7var tile per grid and mode (scafes_autotune_heat7var)
    struct Tuned {
        int mode, nx, ny, nz;
        long e0, e01;  // memAll strides: the ghost width of the tuned layout
        int cfg;
    };
    Tuned tuned[8] = {};
    int ntuned = 0;
    // largest |value| of coefficient fields read by the explicit-step
    // stability checks, per device pointer (invalidated by every write the
    // library makes into the field, by scafes_free and by scafes_coef_changed)
    struct CoefMax {
        const void* p;
        long count;   // nodes of the grid it was reduced over
        size_t span;  // bytes of its memAll layout from p
        int dtype, hw;
        double v;
    };
    CoefMax coef[8] = {};
    int ncoef = 0;
    hipEvent_t events[8] = {};
    // gradient workspace (states, adjoint fields, snapshot pool) kept between
    // gradient calls; grows on demand, freed by scafes_ctx_release_workspace
    // and scafes_ctx_destroy
    void* grad_ws = nullptr;
    size_t grad_ws_bytes = 0;
    // batched 2-D ensemble: the instantiated graph of the last captured
    // launch and the launch arguments it was captured with
    hipGraphExec_t batch_exec = nullptr;
    unsigned char batch_key[128] = {};
};

Ctx* as_ctx(scafes_ctx* c);

// ---------------------------------------------------------------- geometry
struct Geo {
    int nx, ny, nz;     // nodes per axis (1 on inactive axes)
    int dim;
    int nl0, nl1, nl2;  // nLayersAtBorder on active axes, 0 on inactive ones
    long e0, e01;       // memAll strides
    long base;          // offset of node (0,0,0)
};

// Builds the node->memory map of a field with hw ghost layers.
int make_geo(const scafes_geom* g, int hw, Geo* out);

__host__ __device__ inline long at(const Geo& g, int i, int j, int k)
{
    return g.base + i + g.e0 * j + g.e01 * k;
}

__device__ inline bool is_border(const Geo& g, int i, int j, int k)
{
    return i < g.nl0 || i > g.nx - 1 - g.nl0 || j < g.nl1 || j > g.ny - 1 - g.nl1 ||
           k < g.nl2 || k > g.nz - 1 - g.nl2;
}

// Largest plane range in bytes of the buffer-resource kernels.  They issue
// the same loads and stores on every iteration and send unneeded ones to the
// out-of-range offset `range` (a load there returns 0 without a memory
// access, a store is dropped), so the compiler can count outstanding
// accesses and wait only for the ones a plane consumes.
constexpr long kOobLimit = 0xFFFFFF00L;

// Maps the hardware dispatch order (round-robin over the 8 XCDs of an
// MI355X) to a logical tile order in which each XCD owns a contiguous run
// of tiles, so y-neighbouring tiles share that XCD's L2.
__device__ inline int xcd_remap(int b, int nb)
{
    constexpr int kXcd = 8;
    const int per = nb / kXcd;
    const int rem = nb % kXcd;
    const int x = b % kXcd;
    const int q = b / kXcd;
    // XCD x gets per (+1 for x < rem) logical blocks starting at its offset
    const int off = x * per + (x < rem ? x : rem);
    if (q < per + (x < rem ? 1 : 0)) return off + q;
    return b;  // unreachable for a full grid
}

inline int ceil_div(long a, long b) { return (int)((a + b - 1) / b); }

// Pennes forward / adjoint launches (defined in stencil3d.hip).
// dtype SCAFES_F32_KAHAN: compensated FP32 forward (cin/cout compensation fields);
// tprev_lo != null: the adjoint reads the forward state as a (hi, lo) pair.
int pennes_step_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, const void* t_old,
                       void* t_new, const void* w, const void* q, const scafes_pennes_params* p,
                       const void* cin, void* cout);
int pennes_adj_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, const void* lam,
                      void* lam_new, const void* w, const void* tprev, double* grad,
                      const scafes_pennes_params* p, const void* tprev_lo);

// One variable-kappa step with the one-step tile of the mode (defined in
// stencil3d.hip); mode is a SCAFES_HEATVAR_* value.
int heat7var_step_one(Ctx* c, const Geo& geo, const scafes_geom* g, int mode, const void* u, void* un,
                      const void* cin, void* cout, const void* kap, double tau, int border_mode, double border_value);
// Steps of the next sweep of a run with `left` steps to go on this grid
// (the mode's default or autotuned tile, the remainders of its multi-step
// sweeps taken by the two-, three- or four-step tile of the mode), and
// that sweep of k steps (stencil3d.hip); bit-identical to k one-step
// launches.
int heat7var_next_sweep(const Ctx* c, const Geo& geo, int mode, long left);
int heat7var_sweep_k(Ctx* c, const Geo& geo, const scafes_geom* g, int mode, int k, const void* u, void* un,
                     const void* cin, void* cout, const void* kap, double tau, int border_mode, double border_value);
// k = 3 or 4 steps u -> un (cin -> cout) in one sweep that also writes the
// states after 1..k-1 steps to mid[0..k-2] (value parts only in the
// compensated mode): k = 3 with the grid's three-step tile where it has such
// a variant (heat7var_states_sweep_ok), k = 4 with the mode's four-step tile
// behind it (heat7var_states4_ok); SCAFES_ERR_UNSUPPORTED (nothing
// launched) otherwise.
bool heat7var_states_sweep_ok(const Ctx* c, const Geo& geo, int mode);
bool heat7var_states4_ok(const Ctx* c, const Geo& geo, int mode);
int heat7var_sweep_states(Ctx* c, const Geo& geo, const scafes_geom* g, int mode, int k, const void* u, void* un,
                          const void* cin, void* cout, void* const* mid, const void* kap, double tau,
                          int border_mode, double border_value);
// Adjoint step of the variable-kappa heat w.r.t. kappa (defined in
// stencil3d.hip): lam^n -> lam^{n-1}, grad += the step's kappa terms with
// u^{n-1} = uprev; dtype SCAFES_F64 or SCAFES_F32 (fields), grad FP64.
int heatvar_adj_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, const void* lam, void* lam_new,
                       const void* kap, const void* uprev, double* grad, double tau);

// Two reverse steps of the variable-kappa adjoint per sweep (defined in
// heatvar_tb2.hip): lam^n -> lam^{n-2} with u1 = u^{n-1}, u2 = u^{n-2};
// bit-identical to two heatvar_adj_launch calls.
bool heatvar_adj2_supported(const Geo& geo);
int heatvar_adj2_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, const void* lam, void* lamo,
                        const void* kap, const void* u1, const void* u2, double* grad, double tau);

// K = 2..kAdjMaxK reverse Pennes adjoint steps per sweep (defined in
// pennes_adjk.hip): lam^n -> lam^{n-K} with t[l] = T^{n-1-l} (+ tlo[l] for
// compensated states, else tlo = nullptr), l = 0..K-1.
constexpr int kAdjMaxK = 4;
// K = 2..kAdjMaxK reverse variable-kappa adjoint steps per sweep (defined in
// heatvar_adjk.hip): lam^n -> lam^{n-K} with u[l] = u^{n-1-l} (value parts),
// bit-identical to K heatvar_adj_launch calls.
bool heatvar_adjk_supported(const Geo& geo);
int heatvar_adjk_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, int K, const void* lam,
                        void* lamo, const void* kap, const void* const* u, double* grad, double tau);
bool pennes_adjk_supported(const Geo& geo);
int pennes_adjk_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, int K, const void* lam,
                       void* lam_out, const void* w, const void* const* t, const void* const* tlo, double* grad,
                       const scafes_pennes_params* p);

// Multi-step (two, three or four steps per sweep) variable-kappa heat
// (defined in heatvar_tb2.hip); heatvar_tb2_launch advances the
// configuration's steps.
int heatvar_tb2_num_cfgs();
void heatvar_tb2_cfg(int cfg, int* tx, int* ty, int* zc, int* d, int* steps);
bool heatvar_tb2_supported(const Geo& geo, int mode);  // 3-D grids with 32-bit in-plane offsets
// mid = nullptr, or the steps - 1 intermediate state outputs of a
// configuration with a state-writing variant (heatvar_tb_states_supported).
int heatvar_tb2_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int mode, const void* u, void* uo,
                       const void* cin, void* co, const void* kap, double tau, int border_mode, double border_value,
                       int cfg, void* const* mid = nullptr);
bool heatvar_tb_states_supported(int cfg);

// Yee box of N cells per axis; fields hold N+1 nodes per axis, x fastest.
struct YeeDims {
    int N;       // cells per axis
    long M, M2;  // N+1, (N+1)^2
    long nf;     // (N+1)^3
};
// Yee with a relative permittivity field and its adjoint (defined in
// yee_eps.hip).  ce = s / eps per node (state arithmetic); states are the six
// Yee arrays; adjoint states hold lam_E then lam_H.
bool yee_dims(int ncell, YeeDims* d);
long yee_eps_partials();
int yee_eps_coef(Ctx* c, const YeeDims& d, int dtype, double s, const void* eps, void* ce);
int yee_eps_step(Ctx* c, const YeeDims& d, int dtype, double s, const void* ce, const void* a, void* b);
// Two steps a -> b in one sweep (the two-step kernel with ce); only where
// yee_eps_two_step_supported.
bool yee_eps_two_step_supported(const YeeDims& d, int dtype);
int yee_eps_step2(Ctx* c, const YeeDims& d, int dtype, double s, const void* ce, const void* a, void* b);
// The same sweep also writing the state after one step to mid.
int yee_eps_step2_states(Ctx* c, const YeeDims& d, int dtype, double s, const void* ce, const void* a, void* mid,
                         void* b);
int yee_eps_lam_init(Ctx* c, const YeeDims& d, int dtype, const void* fN, const void* target, void* lam,
                     double* partial, double* J);
// One adjoint step: st = state n-1, la = (lam_E', lam_H') -> lb, grad +=.
int yee_eps_adjoint(Ctx* c, const YeeDims& d, int dtype, double s, const void* ce, const void* eps, const void* st,
                    const void* la, void* lb, double* grad);
// Looks up, once per device and with that device current, the device
// address of the scratch array of scafes_yee_run_noalloc (yee.hip); called
// by scafes_ctx_create.
int yee_noalloc_setup(int device);
// One fused leapfrog step a -> b with the E coefficient field ce (yee.hip).
template <typename T>
int yee_fused_eps_launch(Ctx* c, const YeeDims& d, T s, const T* a, T* b, const T* ce);
// Two leapfrog steps per sweep, out of place (defined in yee_tb2.hip); ep0 /
// ep1 (nullable): per-workgroup FP64 energy partials of the two steps.
struct Tb2Plan {
    int tiles_x, tiles_y, zc;  // xy tiles, z planes per workgroup
    long nb;                   // workgroups
};
bool yee_tb2_supported(const YeeDims& d, int elem_bytes);
Tb2Plan yee_tb2_plan(const YeeDims& d, int elem_bytes, int cu_count);
// mid (nullable, with ce and without energy partials): the state after one
// step is written there too.
template <typename T>
int yee_tb2_launch(Ctx* c, const YeeDims& d, const Tb2Plan& pl, T s, const T* a, T* b, double* ep0, double* ep1,
                   const T* ce = nullptr, T* mid = nullptr);

// Reductions (defined in scafes_capi.hip).
enum RedOp { RED_MAX_ABS_DIFF = 0, RED_SUM_SQ_DIFF = 1, RED_SUM_PROD = 2 };
// Sum of the first nblocks entries of c->red_partial (per-block FP64 partial
// sums a kernel wrote there, nblocks <= c->red_capacity), waited for.
int reduce_partials(Ctx* c, int nblocks, double* out);
int reduce_nodes(Ctx* c, const Geo& ga, const Geo& gb, int dtype, const void* a, const void* b,
                 int op, int inner_only, double* out);

// Largest |value| over all nodes of a coefficient field with hw ghost
// layers (FP64 device reduction).  use_cache: reuse the value of an earlier
// call for the same pointer, grid and dtype unless the field was written
// through the library since.
int coef_abs_max(Ctx* c, const Geo& geo, int hw, int dtype, const void* p, bool use_cache, double* out);
// Drops cached maxima of fields overlapping [p, p + bytes); bytes = 0 drops all.
void coef_invalidate(Ctx* c, const void* p, size_t bytes);

// Explicit-Euler stability of the diffusion updates, in the positivity form
// the reference's drivers check (tests/system/HeatEqnFDM/Dirichlet/
// TimeLinSpaceLin/main3D.cpp:52-55, examples/TumorHeatEqnFDM/
// startSimulation.py:238-301): the new value is a nonnegative combination of
// the old ones when tau * rate <= 1, rate = sum_p 2 d_max / h_p^2 (+ the
// largest reaction coefficient).  Returns SCAFES_ERR_ARG with both numbers
// otherwise.
double diffusion_rate(const scafes_geom* g, double dmax);
int check_explicit_tau(double tau, double rate, const char* who);
// The Pennes bound (stencil3d.hip): diffusion + cperf * max|w| (+ the
// Tumor border's convective face); w has the layout of T (hw).
int pennes_check_tau(Ctx* c, const Geo& geo, const scafes_geom* g, int hw, int dtype, const void* w,
                     const scafes_pennes_params* p, bool use_cache, const char* who);

}  // namespace scafes
