// stencil3d.hip -- 2.5D z-streaming 7-point stencil kernels for gfx950.
//
// One workgroup owns a TX x TY tile of an xy plane and walks a chunk of z
// planes.  The centre column of the stencil field lives in a 3-deep
// register queue (z-1, z, z+1); the current xy plane plus a one-node halo
// is staged in LDS (double buffered, one barrier per plane), so every
// field element is fetched from HBM once per sweep apart from the tile
// halo, which neighbouring tiles on the same XCD bring into L2.  Tiles
// are mapped XCD-contiguously (xcd_remap).
//
// Kernels never read ghost layers: every coordinate is clamped to the node
// range, and only GLOBAL_INNER nodes (whose neighbours are nodes) use
// neighbour values.  Ghost layers of the output are never written.
//
// Replaces the OpenMP node loop evalTypeTwoAt (src/ScaFES_Problem.hpp:5397-5476),
// the copy-old-to-new (Problem.hpp:6226-6256) and the per-node index math
// (DataField.hpp:962-969 -> Grid.hpp:955-968) of the reference.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "heatvar_ops.hpp"
#include "scafes_common.hpp"

namespace scafes {

template <typename T>
struct St7 {
    T c, w, e, s, n, d, u;  // centre, x-, x+, y-, y+, z-, z+
};

// ---------------------------------------------------------------- skeleton
// D = prefetch depth in planes (even, >= 2).  Every value a thread needs for
// plane k (its centre, the halo value it stages for the tile, and the
// centre-only fields of the op) is loaded D planes ahead into a register
// ring, so each plane's barrier waits on loads issued D iterations before.
// Op interface: kTwo (second stencil field b, e.g. kappa), kThree (third
// stencil field c, with kTwo), kNX centre-only fields x[0..kNX-1] of type T,
// apply(g, i, j, k, a7, b7, c7, xv).
template <class Op, typename T, int TX, int TY, int D>
__global__ __launch_bounds__(TX* TY) void stream_kernel(Op op, Geo g, int zc, int tiles_x,
                                                         int tiles_y)
{
    static_assert(D >= 2 && D % 2 == 0, "prefetch depth must be even");
    constexpr bool TWO = Op::kTwo;
    constexpr bool THREE = Op::kThree;
    static_assert(!THREE || TWO, "a third stencil field needs the second");
    constexpr int NX = Op::kNX;
    constexpr int NXA = NX > 0 ? NX : 1;
    __shared__ T sa[2][TY + 2][TX + 2];
    __shared__ T sb[TWO ? 2 : 1][TWO ? TY + 2 : 1][TWO ? TX + 2 : 1];
    __shared__ T sc[THREE ? 2 : 1][THREE ? TY + 2 : 1][THREE ? TX + 2 : 1];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per_plane = tiles_x * tiles_y;
    const int tz = lb / per_plane;
    const int r = lb - tz * per_plane;
    const int tyi = r / tiles_x;
    const int txi = r - tyi * tiles_x;
    const int i = txi * TX + tx, j = tyi * TY + ty;
    const int ic = min(i, g.nx - 1), jc = min(j, g.ny - 1);
    const int k0 = tz * zc, k1 = min(k0 + zc, g.nz);
    const int kcmax = min(k1, g.nz - 1);  // last plane whose centre value is needed
    const bool active = (i < g.nx) && (j < g.ny);
    // halo duty of this thread: one x-halo column value and/or one y-halo row value
    const bool hx = (tx == 0) || (tx == TX - 1);
    const bool hy = (ty == 0) || (ty == TY - 1);
    const int hxi = (tx == 0) ? max(ic - 1, 0) : min(ic + 1, g.nx - 1);
    const int hyj = (ty == 0) ? max(jc - 1, 0) : min(jc + 1, g.ny - 1);
    const int sxh = (tx == 0) ? 0 : TX + 1;
    const int syh = (ty == 0) ? 0 : TY + 1;
    const long oc = at(g, ic, jc, 0), ohx = at(g, hxi, jc, 0), ohy = at(g, ic, hyj, 0);
    const T* __restrict__ A = op.a;
    const T* __restrict__ B = op.b;
    const T* __restrict__ Cf = op.c3;

    T qa[D], qb[D], qax[D], qbx[D], qay[D], qby[D], qc[D], qcx[D], qcy[D], qx[NXA][D];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        qa[s] = qb[s] = qax[s] = qbx[s] = qay[s] = qby[s] = qc[s] = qcx[s] = qcy[s] = T(0);
#pragma unroll
        for (int x = 0; x < NXA; ++x) qx[x][s] = T(0);
    }
    auto fetch = [&](const int slot, const int kk) {
        if (kk <= kcmax) {
            const long o = g.e01 * kk;
            qa[slot] = A[oc + o];
            if constexpr (TWO) qb[slot] = B[oc + o];
            if constexpr (THREE) qc[slot] = Cf[oc + o];
            if (kk < k1) {
                if (hx) {
                    qax[slot] = A[ohx + o];
                    if constexpr (TWO) qbx[slot] = B[ohx + o];
                    if constexpr (THREE) qcx[slot] = Cf[ohx + o];
                }
                if (hy) {
                    qay[slot] = A[ohy + o];
                    if constexpr (TWO) qby[slot] = B[ohy + o];
                    if constexpr (THREE) qcy[slot] = Cf[ohy + o];
                }
#pragma unroll
                for (int x = 0; x < NX; ++x) qx[x][slot] = op.x[x][oc + o];
            }
        }
    };

    T am = A[oc + g.e01 * max(k0 - 1, 0)];
    T bm = T(0), cm = T(0);
    if constexpr (TWO) bm = B[oc + g.e01 * max(k0 - 1, 0)];
    if constexpr (THREE) cm = Cf[oc + g.e01 * max(k0 - 1, 0)];
#pragma unroll
    for (int s = 0; s < D; ++s) fetch(s, k0 + s);

    for (int kb = k0; kb < k1; kb += D) {
#pragma unroll
        for (int s = 0; s < D; ++s) {
            const int k = kb + s;
            if (k >= k1) break;  // uniform over the workgroup
            const int buf = s & 1;
            const T ac = qa[s], ap = qa[(s + 1) % D];
            const T bc = qb[s], bp = qb[(s + 1) % D];
            const T cc = qc[s], cp = qc[(s + 1) % D];
            T xv[NXA];
#pragma unroll
            for (int x = 0; x < NXA; ++x) xv[x] = qx[x][s];
            sa[buf][ty + 1][tx + 1] = ac;
            if (hx) sa[buf][ty + 1][sxh] = qax[s];
            if (hy) sa[buf][syh][tx + 1] = qay[s];
            if constexpr (TWO) {
                sb[buf][ty + 1][tx + 1] = bc;
                if (hx) sb[buf][ty + 1][sxh] = qbx[s];
                if (hy) sb[buf][syh][tx + 1] = qby[s];
            }
            if constexpr (THREE) {
                sc[buf][ty + 1][tx + 1] = cc;
                if (hx) sc[buf][ty + 1][sxh] = qcx[s];
                if (hy) sc[buf][syh][tx + 1] = qcy[s];
            }
            fetch(s, k + D);
            __syncthreads();
            if (active) {
                St7<T> sA{ac, sa[buf][ty + 1][tx], sa[buf][ty + 1][tx + 2], sa[buf][ty][tx + 1],
                          sa[buf][ty + 2][tx + 1], am, ap};
                St7<T> sB{};
                if constexpr (TWO)
                    sB = St7<T>{bc, sb[buf][ty + 1][tx], sb[buf][ty + 1][tx + 2], sb[buf][ty][tx + 1],
                                sb[buf][ty + 2][tx + 1], bm, bp};
                St7<T> sC{};
                if constexpr (THREE)
                    sC = St7<T>{cc, sc[buf][ty + 1][tx], sc[buf][ty + 1][tx + 2], sc[buf][ty][tx + 1],
                                sc[buf][ty + 2][tx + 1], cm, cp};
                op.apply(g, i, j, k, sA, sB, sC, xv);
            }
            am = ac;
            bm = bc;
            cm = cc;
        }
    }
}

// ---------------------------------------------------------------- ops
// Constant-coefficient heat / Pennes system update in the reference's FP64
// operation order (HeatEqnFDM.hpp:209-217, PennesBioheatEqnFDM.hpp:158-167,
// Dirichlet.hpp:124-130).  The library is built with -ffp-contract=off.
struct HeatConstOp {
    static constexpr bool kTwo = false;
    static constexpr bool kThree = false;
    static constexpr int kNX = 0;
    const double* a;
    const double* b;
    const double* c3;
    const double* x[1];
    double* out;
    const double* f_old;
    const double* g_old;
    const double* g_new;
    Geo gk;  // layout of the stencil-0 known fields
    double ctau_a, ctw, ctf, hh0, hh1, hh2, border_value;
    double a2[3], alpha, epsilon, sigma;  // flux borders: a2[p] = 2 h_p / lambda
    int has_perf, has_source, border_mode;
    // Second difference along one axis at a flux border node: the outer
    // neighbour replaced by the inner one w plus the flux term.
    __device__ double flux_diff(double w, double u, double a, double gold) const
    {
        double t = w + w;
        if (border_mode == SCAFES_BORDER_NEUMANN) {
            t = t - a * gold;
        } else if (border_mode == SCAFES_BORDER_CAUCHY) {
            t = t - a * alpha * u;
            t = t + a * gold;
        } else {
            t = t - a * epsilon * sigma * u * u * u * u;
            t = t + a * gold;
        }
        return t - 2.0 * u;
    }
    __device__ void apply(const Geo& g, int i, int j, int k, const St7<double>& s,
                          const St7<double>&, const St7<double>&, const double*) const
    {
        const long m = at(g, i, j, k);
        if (!is_border(g, i, j, k)) {
            double v = s.c;
            v += (ctau_a * ((s.w + s.e) - 2.0 * s.c)) / hh0;
            if (g.dim > 1) v += (ctau_a * ((s.s + s.n) - 2.0 * s.c)) / hh1;
            if (g.dim > 2) v += (ctau_a * ((s.d + s.u) - 2.0 * s.c)) / hh2;
            if (has_perf) v -= ctw * s.c;
            if (has_source) v += ctf * f_old[at(gk, i, j, k)];
            out[m] = v;
        } else if (border_mode == SCAFES_BORDER_FIELD) {
            out[m] = g_new[at(gk, i, j, k)];
        } else if (border_mode == SCAFES_BORDER_VALUE) {
            out[m] = border_value;
        } else if (border_mode >= SCAFES_BORDER_NEUMANN) {
            const long q = at(gk, i, j, k);
            if (i == 0 || (g.dim > 1 && j == 0) || (g.dim > 2 && k == 0)) {
                out[m] = g_new[q];
                return;
            }
            const double gold = g_old[q];
            double v = s.c;
            v += (ctau_a * (i == g.nx - 1 ? flux_diff(s.w, s.c, a2[0], gold) : (s.w + s.e) - 2.0 * s.c)) / hh0;
            if (g.dim > 1)
                v += (ctau_a * (j == g.ny - 1 ? flux_diff(s.s, s.c, a2[1], gold) : (s.s + s.n) - 2.0 * s.c)) / hh1;
            if (g.dim > 2)
                v += (ctau_a * (k == g.nz - 1 ? flux_diff(s.d, s.c, a2[2], gold) : (s.d + s.u) - 2.0 * s.c)) / hh2;
            if (has_perf) v -= ctw * s.c;
            if (has_source) v += ctf * f_old[q];
            out[m] = v;
        } else {
            out[m] = s.c;
        }
    }
};

// Variable-coefficient conservative heat.  MODE 0: FP64 in the CPU path's
// order; 1: FP32; 2: FP32 with Kahan-compensated accumulation of the
// per-step increment (u_true ~ u + c).
template <typename T, int MODE>
struct HeatVarOp {
    static constexpr bool kTwo = true;
    static constexpr bool kThree = false;
    static constexpr int kNX = MODE == 2 ? 1 : 0;
    const T* a;  // u
    const T* b;  // kappa
    const T* c3;
    const T* x[1];  // x[0] = incoming compensation (Kahan mode)
    T* out;
    T* cout;
    double tau, hh0, hh1, hh2;  // FP64 path
    float c0, c1, c2;           // 0.5*tau/h_p^2 for the FP32 paths
    T border_value;
    int border_mode;
    __device__ void apply(const Geo& g, int i, int j, int k, const St7<T>& s,
                          const St7<T>& q, const St7<T>&, const T* xv) const
    {
        const long m = at(g, i, j, k);
        if (!is_border(g, i, j, k)) {
            if constexpr (MODE == 0) {
                double v = s.c;
                v += (tau * (0.5 * (q.e + q.c) * (s.e - s.c) - 0.5 * (q.c + q.w) * (s.c - s.w))) / hh0;
                if (g.dim > 1)
                    v += (tau * (0.5 * (q.n + q.c) * (s.n - s.c) - 0.5 * (q.c + q.s) * (s.c - s.s))) / hh1;
                if (g.dim > 2)
                    v += (tau * (0.5 * (q.u + q.c) * (s.u - s.c) - 0.5 * (q.c + q.d) * (s.c - s.d))) / hh2;
                out[m] = v;
            } else {
                float d = c0 * __builtin_fmaf(q.e + q.c, s.e - s.c, -((q.c + q.w) * (s.c - s.w)));
                if (g.dim > 1)
                    d = __builtin_fmaf(c1, __builtin_fmaf(q.n + q.c, s.n - s.c, -((q.c + q.s) * (s.c - s.s))), d);
                if (g.dim > 2)
                    d = __builtin_fmaf(c2, __builtin_fmaf(q.u + q.c, s.u - s.c, -((q.c + q.d) * (s.c - s.d))), d);
                if constexpr (MODE == 2) {
                    const float y = d + xv[0];
                    const float t = s.c + y;
                    cout[m] = y - (t - s.c);
                    out[m] = t;
                } else {
                    out[m] = s.c + d;
                }
            }
        } else {
            out[m] = border_mode == SCAFES_BORDER_VALUE ? border_value : s.c;
            if constexpr (MODE == 2) cout[m] = border_mode == SCAFES_BORDER_VALUE ? 0.0f : xv[0];
        }
    }
};

// Pennes bioheat, TumorHeatEqnFDM::updateInner form with node fields w, Q
// (examples/TumorHeatEqnFDM/TumorHeatEqnFDM.hpp:258-267); border nodes keep.
// KAHAN (FP32 only): the per-step increment is accumulated with a
// persistent FP32 compensation field (T_true ~ out + cout).
template <typename T, bool KAHAN = false>
struct PennesOp {
    static constexpr bool kTwo = false;
    static constexpr bool kThree = false;
    static constexpr int kNX = KAHAN ? 3 : 2;
    const T* a;  // T_old
    const T* b;
    const T* c3;
    const T* x[3];  // perfusion w, source Q, incoming compensation (KAHAN)
    T* out;
    T* cout;
    double ctd, ctp, ctq, ta, hh0, hh1, hh2;  // FP64 path
    float fd0, fd1, fd2, fp, fq, fta;         // FP32 path
    int tumor;                                // border rule of the Tumor example
    double a2h[3], hconv, tinf, qbc;          // (2 h_p)/K, H, T_INF, Q_BC
    // One axis of TumorHeatEqnFDM::updateBorder (TumorHeatEqnFDM.hpp:300-345):
    // the missing neighbour replaced by the inner one plus a flux term.
    __device__ T tumor_term(int idx, int n, bool last, int p, T wv, T ev, T c) const
    {
        const T a = (T)a2h[p];
        if (idx == n - 1) {
            if (last) return ((wv + wv) - ((a * (T)hconv) * (c - (T)tinf))) - (T)2 * c;
            return ((wv + wv) - (a * ((T)-1 * (T)qbc))) - (T)2 * c;
        }
        if (idx == 0) return ((ev + (a * (T)qbc)) + ev) - (T)2 * c;
        return (wv + ev) - (T)2 * c;
    }
    __device__ void apply(const Geo& g, int i, int j, int k, const St7<T>& s, const St7<T>&,
                          const St7<T>&, const T* xv) const
    {
        const long m = at(g, i, j, k);
        // Tumor border: every node takes the border form, a branch uniform
        // over the launch (per node only selects); away from the faces it is
        // the interior form term for term, so inner nodes get the same value.
        if (tumor || is_border(g, i, j, k)) {
            if (!tumor) {
                out[m] = s.c;
                if constexpr (KAHAN) cout[m] = xv[2];
                return;
            }
            const T t0 = tumor_term(i, g.nx, g.dim == 1, 0, s.w, s.e, s.c);
            const T t1 = g.dim > 1 ? tumor_term(j, g.ny, g.dim == 2, 1, s.s, s.n, s.c) : (T)0;
            const T t2 = g.dim > 2 ? tumor_term(k, g.nz, true, 2, s.d, s.u, s.c) : (T)0;
            if constexpr (sizeof(T) == 8) {
                double v = s.c;
                v += (ctd * t0) / hh0;
                if (g.dim > 1) v += (ctd * t1) / hh1;
                if (g.dim > 2) v += (ctd * t2) / hh2;
                v += (ctp * xv[0]) * (ta - s.c);
                v += ctq * xv[1];
                out[m] = v;
            } else {
                float d = fd0 * t0;
                if (g.dim > 1) d = __builtin_fmaf(fd1, t1, d);
                if (g.dim > 2) d = __builtin_fmaf(fd2, t2, d);
                d = __builtin_fmaf(fp * xv[0], fta - s.c, d);
                d = __builtin_fmaf(fq, xv[1], d);
                if constexpr (KAHAN) {
                    const float y = d + xv[2];
                    const float t = s.c + y;
                    cout[m] = y - (t - s.c);
                    out[m] = t;
                } else {
                    out[m] = s.c + d;
                }
            }
            return;
        }
        if constexpr (sizeof(T) == 8) {
            double v = s.c;
            v += (ctd * ((s.w + s.e) - 2.0 * s.c)) / hh0;
            if (g.dim > 1) v += (ctd * ((s.s + s.n) - 2.0 * s.c)) / hh1;
            if (g.dim > 2) v += (ctd * ((s.d + s.u) - 2.0 * s.c)) / hh2;
            v += (ctp * xv[0]) * (ta - s.c);
            v += ctq * xv[1];
            out[m] = v;
        } else {
            float d = fd0 * ((s.w + s.e) - 2.0f * s.c);
            if (g.dim > 1) d = __builtin_fmaf(fd1, (s.s + s.n) - 2.0f * s.c, d);
            if (g.dim > 2) d = __builtin_fmaf(fd2, (s.d + s.u) - 2.0f * s.c, d);
            d = __builtin_fmaf(fp * xv[0], fta - s.c, d);
            d = __builtin_fmaf(fq, xv[1], d);
            if constexpr (KAHAN) {
                const float y = d + xv[2];
                const float t = s.c + y;
                cout[m] = y - (t - s.c);
                out[m] = t;
            } else {
                out[m] = s.c + d;
            }
        }
    }
};

// Pennes bioheat with per-node materials, MRIDataPBHEqnFDM::updateInner /
// updateBorder (examples/MRIDataPBHEqnFDM/MRIDataPBHEqnFDM.hpp:143-273), FP64
// in the reference's operation order.  The materials do not change over a
// run, so mri_coef_kernel folds them once into four per-node coefficients
// with the same rounding as the hook's left-to-right products:
//   x[0] = tau * (lambda/(rho c))                  (diffusion, every axis)
//   x[1] = (tau * ((rho_b c_b)/(rho c))) * omega   (perfusion)
//   x[2] = T_blood
//   x[3] = (tau * (1/(rho c))) * q                 (metabolic source)
// and a step streams T plus these four instead of nine material fields.
// Border nodes also read lambda (for 2 h_p / lambda) and, on the far face of
// the last axis, the surface flag, both in the stencil-0 node layout gk.
struct MriOp {
    static constexpr bool kTwo = false;
    static constexpr bool kThree = false;
    static constexpr int kNX = 4;
    const double* a;  // T_old
    const double* b;
    const double* c3;
    const double* x[4];
    double* out;
    const double* lam;   // lambda, layout gk
    const double* surf;  // surface flag, layout gk
    Geo gk;
    double hh0, hh1, hh2, h[3];
    double hconv, tamb, qbc, qskull, epsilon, sigma, tamb4;  // tamb4 = (T_AMB + 273.15)^4
    // One axis of updateBorder (MRIDataPBHEqnFDM.hpp:193-262): the missing
    // neighbour replaced by the inner one plus the face's flux term.  On the
    // far face of the last axis an open skull (surface == 1) has convection
    // and radiation, where the radiated T^4 is taken at the value the update
    // has accumulated so far (the hook reads vNew there, :200); a closed
    // skull has the flux Q_SKULL.
    __device__ double term(int idx, int n, bool last, double hp, double kk, double wv, double ev, double c,
                           double acc, long q) const
    {
        const double a = (2.0 * hp) / kk;
        if (idx == n - 1) {
            if (last) {
                if ((int)surf[q] == 1) {
                    const double t = acc + 273.15;
                    const double t4 = ((t * t) * t) * t;
                    return (((wv + wv) - ((a * hconv) * (c - tamb))) - (((a * epsilon) * sigma) * (t4 - tamb4))) -
                           2.0 * c;
                }
                return ((wv + wv) - (a * (-1.0 * qskull))) - 2.0 * c;
            }
            return ((wv + wv) - (a * (-1.0 * qbc))) - 2.0 * c;
        }
        if (idx == 0) return ((ev + (a * qbc)) + ev) - 2.0 * c;
        return (wv + ev) - 2.0 * c;
    }
    __device__ void apply(const Geo& g, int i, int j, int k, const St7<double>& s, const St7<double>&,
                          const St7<double>&, const double* xv) const
    {
        const long m = at(g, i, j, k);
        double v = s.c;
        if (!is_border(g, i, j, k)) {
            v += (xv[0] * ((s.w + s.e) - 2.0 * s.c)) / hh0;
            if (g.dim > 1) v += (xv[0] * ((s.s + s.n) - 2.0 * s.c)) / hh1;
            if (g.dim > 2) v += (xv[0] * ((s.d + s.u) - 2.0 * s.c)) / hh2;
        } else {
            const long q = at(gk, i, j, k);
            const double kk = lam[q];
            v += (xv[0] * term(i, g.nx, g.dim == 1, h[0], kk, s.w, s.e, s.c, v, q)) / hh0;
            if (g.dim > 1) v += (xv[0] * term(j, g.ny, g.dim == 2, h[1], kk, s.s, s.n, s.c, v, q)) / hh1;
            if (g.dim > 2) v += (xv[0] * term(k, g.nz, true, h[2], kk, s.d, s.u, s.c, v, q)) / hh2;
        }
        v += xv[1] * (xv[2] - s.c);
        v += xv[3];
        out[m] = v;
    }
};

// Per-node coefficients of MriOp from the nine material fields (layout gk)
// into the temperature's layout g; ghost entries are left untouched.
__global__ void mri_coef_kernel(Geo g, Geo gk, const double* __restrict__ rho, const double* __restrict__ cp,
                                const double* __restrict__ lam, const double* __restrict__ rho_b,
                                const double* __restrict__ c_b, const double* __restrict__ omega,
                                const double* __restrict__ t_b, const double* __restrict__ qm, double tau,
                                double* __restrict__ coef, long stride)
{
    const long nn = (long)g.nx * g.ny * g.nz;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nn; t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.nx);
        const long r = t / g.nx;
        const int j = (int)(r % g.ny);
        const int k = (int)(r / g.ny);
        const long q = at(gk, i, j, k), m = at(g, i, j, k);
        const double rc = rho[q] * cp[q];
        coef[m] = tau * (lam[q] / rc);
        coef[stride + m] = (tau * ((rho_b[q] * c_b[q]) / rc)) * omega[q];
        coef[2 * stride + m] = t_b[q];
        coef[3 * stride + m] = (tau * (1.0 / rc)) * qm[q];
    }
}

// Discrete adjoint of HeatVarOp w.r.t. the kappa field (a = lam^n, b =
// kappa, c3 = u^{n-1}; FP64 gradient):
//   grad_k += 0.5 * sum_p th_p sum_{nb = k -+ e_p in the grid} (lam_k - lam_nb)(u_nb - u_k)
// on every node (kappa at a border node enters the faces of its inner
// neighbours), th_p = tau/h_p^2; lam^{n-1} = HeatVarOp's update applied to
// lam on inner nodes (the operator is symmetric), 0 on border nodes.  FP64
// fields: the oracle's operation order; FP32 fields: the FP32 update of
// HeatVarOp and each step's gradient term in FP32, added to the FP64
// gradient.
template <typename T>
struct HeatVarAdjOp {
    static constexpr bool kTwo = true;
    static constexpr bool kThree = true;
    static constexpr int kNX = 0;
    const T* a;   // lam^n
    const T* b;   // kappa
    const T* c3;  // u^{n-1}
    const T* x[1];
    T* out;       // lam^{n-1}
    double* grad;
    double tau, hh0, hh1, hh2, th0, th1, th2;  // th_p = tau / hh_p
    float c0, c1, c2;                          // 0.5*tau/h_p^2 (FP32 update)
    float f0, f1, f2;                          // 0.5 th_p rounded to FP32 (FP32 gradient terms)
    __device__ static double face_sum(bool lo, bool hi, double lc, double lm, double lp, double uc, double um,
                                      double up)
    {
        double sp = 0.0;
        if (lo) sp += (lc - lm) * (um - uc);
        if (hi) sp += (lc - lp) * (up - uc);
        return sp;
    }
    __device__ void apply(const Geo& g, int i, int j, int k, const St7<T>& l, const St7<T>& q, const St7<T>& u,
                          const T*) const
    {
        const long m = at(g, i, j, k);
        double t = 0.0;  // the gradient increment 0.5 * (the step's term)
        if constexpr (sizeof(T) == 8) {
            t += th0 * face_sum(i > 0, i < g.nx - 1, (double)l.c, (double)l.w, (double)l.e, (double)u.c,
                                (double)u.w, (double)u.e);
            if (g.dim > 1)
                t += th1 * face_sum(j > 0, j < g.ny - 1, (double)l.c, (double)l.s, (double)l.n, (double)u.c,
                                    (double)u.s, (double)u.n);
            if (g.dim > 2)
                t += th2 * face_sum(k > 0, k < g.nz - 1, (double)l.c, (double)l.d, (double)l.u, (double)u.c,
                                    (double)u.d, (double)u.u);
            t = 0.5 * t;
        } else {
            // FP32 fields: half the term in FP32 (adj_term3's arithmetic in heatvar_ops.hpp)
            float tf = f0 * adj_face_sum_f(i > 0, i < g.nx - 1, l.c, l.w, l.e, u.c, u.w, u.e);
            if (g.dim > 1) tf = __builtin_fmaf(f1, adj_face_sum_f(j > 0, j < g.ny - 1, l.c, l.s, l.n, u.c, u.s, u.n), tf);
            if (g.dim > 2) tf = __builtin_fmaf(f2, adj_face_sum_f(k > 0, k < g.nz - 1, l.c, l.d, l.u, u.c, u.d, u.u), tf);
            t = (double)tf;
        }
        grad[m] += t;
        if (is_border(g, i, j, k)) {
            out[m] = T(0);
            return;
        }
        if constexpr (sizeof(T) == 8) {
            double v = l.c;
            v += (tau * (0.5 * (q.e + q.c) * (l.e - l.c) - 0.5 * (q.c + q.w) * (l.c - l.w))) / hh0;
            if (g.dim > 1) v += (tau * (0.5 * (q.n + q.c) * (l.n - l.c) - 0.5 * (q.c + q.s) * (l.c - l.s))) / hh1;
            if (g.dim > 2) v += (tau * (0.5 * (q.u + q.c) * (l.u - l.c) - 0.5 * (q.c + q.d) * (l.c - l.d))) / hh2;
            out[m] = v;
        } else {
            float d = c0 * __builtin_fmaf(q.e + q.c, l.e - l.c, -((q.c + q.w) * (l.c - l.w)));
            if (g.dim > 1) d = __builtin_fmaf(c1, __builtin_fmaf(q.n + q.c, l.n - l.c, -((q.c + q.s) * (l.c - l.s))), d);
            if (g.dim > 2) d = __builtin_fmaf(c2, __builtin_fmaf(q.u + q.c, l.u - l.c, -((q.c + q.d) * (l.c - l.d))), d);
            out[m] = l.c + d;
        }
    }
};

// Discrete adjoint of PennesOp (inner nodes), FP64 gradient accumulation.
// LO: the forward state is a (hi, lo) pair from the compensated forward.
template <typename T, bool LO = false>
struct PennesAdjOp {
    static constexpr bool kTwo = false;
    static constexpr bool kThree = false;
    static constexpr int kNX = LO ? 3 : 2;
    const T* a;  // lambda^n
    const T* b;
    const T* c3;
    const T* x[3];  // perfusion w, previous forward state T^{n-1} (hi, lo)
    T* out;         // lambda^{n-1}
    double* grad;
    double ctd, ctp, ta, hh0, hh1, hh2;
    float fd0, fd1, fd2, fp;
    int tumor;                   // the Tumor border rule (every node updated)
    double a2h[3], hconv;        // (2 h_p)/K, H
    // Axis p of the transposed Tumor step at a node with index idx of n:
    // the row of the neighbour at idx-1 weighs this node twice when that
    // neighbour lies on the face (index 0), the row of the one at idx+1
    // twice when it has index n-1, and there is no neighbour beyond the
    // grid (selected away, so ghost values never enter); on the Cauchy face
    // (last axis, idx = n-1) the diagonal gets -(2 h/K) H besides -2.
    __device__ T tumor_adj(int idx, int n, bool last, int p, T lw, T le, T l) const
    {
        lw = idx >= 1 ? lw : T(0);
        le = idx <= n - 2 ? le : T(0);
        if (idx == 1) lw = lw + lw;
        if (idx == n - 2) le = le + le;
        if (last && idx == n - 1) return ((lw + le) - ((T)a2h[p] * (T)hconv) * l) - (T)2 * l;
        return (lw + le) - (T)2 * l;
    }
    __device__ void apply(const Geo& g, int i, int j, int k, const St7<T>& s, const St7<T>&,
                          const St7<T>&, const T* xv) const
    {
        const long m = at(g, i, j, k);
        if (tumor) {
            apply_tumor(g, i, j, k, m, s, xv);
            return;
        }
        if (is_border(g, i, j, k)) {
            out[m] = T(0);
            return;
        }
        double tp = (double)xv[1];
        if constexpr (LO) tp += (double)xv[2];
        grad[m] += (double)s.c * (ctp * (ta - tp));
        if constexpr (sizeof(T) == 8) {
            double v = s.c;
            v += (ctd * ((s.w + s.e) - 2.0 * s.c)) / hh0;
            if (g.dim > 1) v += (ctd * ((s.s + s.n) - 2.0 * s.c)) / hh1;
            if (g.dim > 2) v += (ctd * ((s.d + s.u) - 2.0 * s.c)) / hh2;
            v -= (ctp * xv[0]) * s.c;
            out[m] = v;
        } else {
            float d = fd0 * ((s.w + s.e) - 2.0f * s.c);
            if (g.dim > 1) d = __builtin_fmaf(fd1, (s.s + s.n) - 2.0f * s.c, d);
            if (g.dim > 2) d = __builtin_fmaf(fd2, (s.d + s.u) - 2.0f * s.c, d);
            d = __builtin_fmaf(-fp * xv[0], s.c, d);
            out[m] = s.c + d;
        }
    }
    __device__ void apply_tumor(const Geo& g, int i, int j, int k, long m, const St7<T>& s, const T* xv) const
    {
        double tp = (double)xv[1];
        if constexpr (LO) tp += (double)xv[2];
        grad[m] += (double)s.c * (ctp * (ta - tp));
        const T t0 = tumor_adj(i, g.nx, g.dim == 1, 0, s.w, s.e, s.c);
        const T t1 = g.dim > 1 ? tumor_adj(j, g.ny, g.dim == 2, 1, s.s, s.n, s.c) : T(0);
        const T t2 = g.dim > 2 ? tumor_adj(k, g.nz, true, 2, s.d, s.u, s.c) : T(0);
        if constexpr (sizeof(T) == 8) {
            double v = s.c;
            v += (ctd * t0) / hh0;
            if (g.dim > 1) v += (ctd * t1) / hh1;
            if (g.dim > 2) v += (ctd * t2) / hh2;
            v -= (ctp * xv[0]) * s.c;
            out[m] = v;
        } else {
            float d = fd0 * t0;
            if (g.dim > 1) d = __builtin_fmaf(fd1, t1, d);
            if (g.dim > 2) d = __builtin_fmaf(fd2, t2, d);
            d = __builtin_fmaf(-fp * xv[0], s.c, d);
            out[m] = s.c + d;
        }
    }
};

// ---------------------------------------------------------------- launch
struct TileCfg {
    int tx, ty, zc, d;  // tile edge x, y; z planes per workgroup; prefetch depth
};

static const TileCfg kCfgs[] = {
    {64, 4, 64, 4},   {64, 4, 128, 4}, {64, 8, 64, 4},  {64, 8, 128, 4}, {128, 2, 64, 4},
    {128, 4, 64, 4},  {32, 8, 64, 4},  {64, 4, 64, 2},  {64, 8, 128, 2}, {64, 16, 128, 4},
    {64, 8, 256, 4},  {64, 4, 128, 6}, {64, 8, 128, 6}, {128, 4, 128, 6}, {64, 4, 256, 8},
};
static constexpr int kNumCfgs = sizeof(kCfgs) / sizeof(kCfgs[0]);
static constexpr int kDefaultCfg = 3;
// Fastest one-step tile per heat7var mode, measured on an MI355X at 512^3
// (F64 64x8 zc128 d6, F32 128x2 zc64 d4, F32+Kahan 64x16 zc128 d4); used
// for the odd last step behind two-steps-per-sweep launches.
static constexpr int kOneStepCfgByMode[3] = {12, 4, 9};
// Default per mode: the fastest multi-step tile at 512^3 on an MI355X
// (indices into heatvar_tb2.hip's table), all three steps per sweep: F64
// 64x16 zc64 d2 (10; 620 us/step against 658 for the best two-step tile),
// F32 and F32+Kahan 64x12 zc64 d2 (13; 234 and 280 us/step).
static constexpr int kDefaultCfgByMode[3] = {kNumCfgs + 10, kNumCfgs + 13, kNumCfgs + 13};
// Two-step tile per mode for a remainder of two steps behind a three-step
// configuration (F64 64x8 zc64 d4, F32 and F32+Kahan 64x8 zc64 d2): one
// sweep instead of two one-step sweeps.
static constexpr int kTwoStepCfgByMode[3] = {kNumCfgs + 6, kNumCfgs + 2, kNumCfgs + 2};
// Four-step tile per mode, or -1: F32 + Kahan 64x16 zc128 d2.  Measured on
// two MI355X boxes at 512^3 (us per sweep): 1084-1185 for four steps
// against 800-813 for three with the default tile (271-290 against 267-271
// us per step) and 776-780 for the two-step tile, so 20 steps as 2 x 4 +
// 4 x 3 took 5.42-5.61 ms against 5.64-5.66 as 6 x 3 + 2.  FP32 (963
// against 638 for three steps: 241 and 213 us per step) and FP64 (2660
// against 1790: 665 and 597) lose more per step with four steps and keep
// the two-step remainder.
static constexpr int kFourStepCfgByMode[3] = {-1, -1, kNumCfgs + 24};

template <class Op, typename T, int TX, int TY, int D>
static int launch_t(Ctx* c, const Op& op, const Geo& g, int zc)
{
    const int tiles_x = ceil_div(g.nx, TX), tiles_y = ceil_div(g.ny, TY);
    const int tiles_z = ceil_div(g.nz, zc);
    const long nb = (long)tiles_x * tiles_y * tiles_z;
    if (nb > 0x7fffffffL) return fail(SCAFES_ERR_ARG, "grid too large");
    stream_kernel<Op, T, TX, TY, D><<<dim3((unsigned)nb), dim3(TX, TY), 0, c->stream>>>(op, g, zc, tiles_x,
                                                                                       tiles_y);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

// Tile of the one-step launches that take no configuration (heat7, Pennes
// forward and adjoint, MRI): 128 x 2 threads, 8 z planes per workgroup.
// Short z chunks give enough workgroups to fill 256 CUs on small grids and
// cost one extra halo plane per 4 streamed ones, served from L2.  Measured on
// an MI355X (us per step, this tile vs the best of 15 tiles x 5 chunk
// lengths): Pennes FP32 256^3 50.5 (best), 512^3 482 (best); Pennes FP64
// 256^3 112 vs 109; MRI FP64 256^3 200 vs 192, 120x120x50 16.0 (best); the
// previous 64 x 8, 128-plane default took 87, 515, 127, 263 and 102.
// Rows of at most 64 nodes take 64 x 4 threads instead (a 128-wide row
// would leave half its lanes idle), and where 8-plane chunks would give
// fewer workgroups than CUs the chunks shrink towards four workgroups per
// CU (heat 64^3 FP64, 100 steps: 7.2 us per step with 128 x 2 and 8-plane
// chunks, 256 workgroups).
constexpr int kAutoZc = 8;

template <class Op, typename T>
static int launch_auto(Ctx* c, const Op& op, const Geo& g)
{
    if (g.nx > 64) return launch_t<Op, T, 128, 2, 4>(c, op, g, kAutoZc);
    const long per_plane = ceil_div(g.ny, 4);
    const long cus = c->cu_count > 0 ? c->cu_count : 256;
    int zc = kAutoZc;
    if (per_plane * ceil_div(g.nz, zc) < cus)
        zc = (int)std::max(1L, std::min((long)kAutoZc, per_plane * g.nz / (4 * cus)));
    return launch_t<Op, T, 64, 4, 4>(c, op, g, zc);
}

template <class Op, typename T>
static int launch(Ctx* c, const Op& op, const Geo& g, int cfg)
{
    if (cfg < 0) return launch_auto<Op, T>(c, op, g);
    if (cfg >= kNumCfgs) cfg = kDefaultCfg;
    const TileCfg t = kCfgs[cfg];
    switch (t.tx * 100000 + t.ty * 100 + t.d) {
    case 6400404: return launch_t<Op, T, 64, 4, 4>(c, op, g, t.zc);
    case 6400804: return launch_t<Op, T, 64, 8, 4>(c, op, g, t.zc);
    case 12800204: return launch_t<Op, T, 128, 2, 4>(c, op, g, t.zc);
    case 12800404: return launch_t<Op, T, 128, 4, 4>(c, op, g, t.zc);
    case 3200804: return launch_t<Op, T, 32, 8, 4>(c, op, g, t.zc);
    case 6400402: return launch_t<Op, T, 64, 4, 2>(c, op, g, t.zc);
    case 6400802: return launch_t<Op, T, 64, 8, 2>(c, op, g, t.zc);
    case 6401604: return launch_t<Op, T, 64, 16, 4>(c, op, g, t.zc);
    case 6400406: return launch_t<Op, T, 64, 4, 6>(c, op, g, t.zc);
    case 6400806: return launch_t<Op, T, 64, 8, 6>(c, op, g, t.zc);
    case 12800406: return launch_t<Op, T, 128, 4, 6>(c, op, g, t.zc);
    case 6400408: return launch_t<Op, T, 64, 4, 8>(c, op, g, t.zc);
    default: return fail(SCAFES_ERR_ARG, "bad tile configuration");
    }
}

static void hh_of(const scafes_geom* g, double hh[3])
{
    for (int p = 0; p < 3; ++p) hh[p] = p < g->dim ? g->h[p] * g->h[p] : 1.0;
}

static int check_stencil_geo(const scafes_geom* g, int hw)
{
    if (hw < 1) return fail(SCAFES_ERR_ARG, "unknown field needs stencil width >= 1");
    if (g->nlayers < 1) return fail(SCAFES_ERR_ARG, "nLayersAtBorder must be >= 1 for a 7-point update");
    return SCAFES_OK;
}

int heat7var_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int mode, const void* u, void* un,
                    const void* cin, void* cout, const void* kap, double tau, int border_mode,
                    double border_value, int cfg)
{
    double hh[3];
    hh_of(g, hh);
    if (mode == SCAFES_HEATVAR_F64) {
        HeatVarOp<double, 0> op{};
        op.a = (const double*)u;
        op.b = (const double*)kap;
        op.out = (double*)un;
        op.tau = tau;
        op.hh0 = hh[0];
        op.hh1 = hh[1];
        op.hh2 = hh[2];
        op.border_value = border_value;
        op.border_mode = border_mode;
        return launch<HeatVarOp<double, 0>, double>(c, op, geo, cfg);
    }
    const float c0 = (float)(0.5 * tau / hh[0]), c1 = (float)(0.5 * tau / hh[1]),
                c2 = (float)(0.5 * tau / hh[2]);
    if (mode == SCAFES_HEATVAR_F32) {
        HeatVarOp<float, 1> op{};
        op.a = (const float*)u;
        op.b = (const float*)kap;
        op.out = (float*)un;
        op.c0 = c0;
        op.c1 = c1;
        op.c2 = c2;
        op.border_value = (float)border_value;
        op.border_mode = border_mode;
        return launch<HeatVarOp<float, 1>, float>(c, op, geo, cfg);
    }
    HeatVarOp<float, 2> op{};
    op.a = (const float*)u;
    op.b = (const float*)kap;
    op.out = (float*)un;
    op.x[0] = (const float*)cin;
    op.cout = (float*)cout;
    op.c0 = c0;
    op.c1 = c1;
    op.c2 = c2;
    op.border_value = (float)border_value;
    op.border_mode = border_mode;
    return launch<HeatVarOp<float, 2>, float>(c, op, geo, cfg);
}

int pennes_step_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, const void* t_old,
                       void* t_new, const void* w, const void* q, const scafes_pennes_params* p,
                       const void* cin, void* cout)
{
    double hh[3];
    hh_of(g, hh);
    if (p->border_mode != SCAFES_BORDER_KEEP && p->border_mode != SCAFES_BORDER_TUMOR)
        return fail(SCAFES_ERR_ARG, "pennes: border_mode must be KEEP or TUMOR");
    if (p->border_mode == SCAFES_BORDER_TUMOR && p->k_cond == 0.0)
        return fail(SCAFES_ERR_ARG, "pennes: the TUMOR border needs k_cond != 0");
    auto tumor = [&](auto& op) {
        op.tumor = p->border_mode == SCAFES_BORDER_TUMOR ? 1 : 0;
        for (int q = 0; q < 3; ++q) op.a2h[q] = (2.0 * g->h[q]) / p->k_cond;
        op.hconv = p->h_conv;
        op.tinf = p->t_inf;
        op.qbc = p->q_bc;
    };
    if (dtype == SCAFES_F64) {
        PennesOp<double> op{};
        op.a = (const double*)t_old;
        op.out = (double*)t_new;
        op.x[0] = (const double*)w;
        op.x[1] = (const double*)q;
        op.ctd = p->tau * p->cdiff;
        op.ctp = p->tau * p->cperf;
        op.ctq = p->tau * p->csrc;
        op.ta = p->t_a;
        op.hh0 = hh[0];
        op.hh1 = hh[1];
        op.hh2 = hh[2];
        if (p->border_mode == SCAFES_BORDER_TUMOR) tumor(op);
        return launch<PennesOp<double>, double>(c, op, geo, -1);
    }
    auto fill32 = [&](auto& op) {
        op.a = (const float*)t_old;
        op.out = (float*)t_new;
        op.x[0] = (const float*)w;
        op.x[1] = (const float*)q;
        op.fd0 = (float)(p->tau * p->cdiff / hh[0]);
        op.fd1 = (float)(p->tau * p->cdiff / hh[1]);
        op.fd2 = (float)(p->tau * p->cdiff / hh[2]);
        op.fp = (float)(p->tau * p->cperf);
        op.fq = (float)(p->tau * p->csrc);
        op.fta = (float)p->t_a;
        op.tumor = 0;
        if (p->border_mode == SCAFES_BORDER_TUMOR) tumor(op);
    };
    if (dtype == SCAFES_F32_KAHAN) {
        if (!cin || !cout) return fail(SCAFES_ERR_ARG, "pennes: compensated step needs both compensation fields");
        PennesOp<float, true> op{};
        fill32(op);
        op.x[2] = (const float*)cin;
        op.cout = (float*)cout;
        return launch<PennesOp<float, true>, float>(c, op, geo, -1);
    }
    PennesOp<float> op{};
    fill32(op);
    return launch<PennesOp<float>, float>(c, op, geo, -1);
}

int heat7var_step_one(Ctx* c, const Geo& geo, const scafes_geom* g, int mode, const void* u, void* un,
                      const void* cin, void* cout, const void* kap, double tau, int border_mode, double border_value)
{
    return heat7var_launch(c, geo, g, mode, u, un, cin, cout, kap, tau, border_mode, border_value,
                           kOneStepCfgByMode[mode]);
}

int heatvar_adj_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, const void* lam, void* lam_new,
                       const void* kap, const void* uprev, double* grad, double tau)
{
    double hh[3];
    hh_of(g, hh);
    auto fill = [&](auto& op) {
        op.tau = tau;
        op.hh0 = hh[0];
        op.hh1 = hh[1];
        op.hh2 = hh[2];
        op.th0 = tau / hh[0];
        op.th1 = tau / hh[1];
        op.th2 = tau / hh[2];
        op.f0 = (float)(0.5 * op.th0);
        op.f1 = (float)(0.5 * op.th1);
        op.f2 = (float)(0.5 * op.th2);
        op.c0 = (float)(0.5 * tau / hh[0]);
        op.c1 = (float)(0.5 * tau / hh[1]);
        op.c2 = (float)(0.5 * tau / hh[2]);
        op.grad = grad;
    };
    if (dtype == SCAFES_F64) {
        HeatVarAdjOp<double> op{};
        op.a = (const double*)lam;
        op.b = (const double*)kap;
        op.c3 = (const double*)uprev;
        op.out = (double*)lam_new;
        fill(op);
        return launch<HeatVarAdjOp<double>, double>(c, op, geo, -1);
    }
    HeatVarAdjOp<float> op{};
    op.a = (const float*)lam;
    op.b = (const float*)kap;
    op.c3 = (const float*)uprev;
    op.out = (float*)lam_new;
    fill(op);
    return launch<HeatVarAdjOp<float>, float>(c, op, geo, -1);
}

int pennes_adj_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, const void* lam,
                      void* lam_new, const void* w, const void* tprev, double* grad,
                      const scafes_pennes_params* p, const void* tprev_lo)
{
    if (p->border_mode != SCAFES_BORDER_KEEP && p->border_mode != SCAFES_BORDER_TUMOR)
        return fail(SCAFES_ERR_ARG, "pennes adjoint: border_mode must be KEEP or TUMOR");
    if (p->border_mode == SCAFES_BORDER_TUMOR && p->k_cond == 0.0)
        return fail(SCAFES_ERR_ARG, "pennes adjoint: the TUMOR border needs k_cond != 0");
    double hh[3];
    hh_of(g, hh);
    auto tumor = [&](auto& op) {
        op.tumor = p->border_mode == SCAFES_BORDER_TUMOR ? 1 : 0;
        for (int q = 0; q < 3; ++q) op.a2h[q] = (2.0 * g->h[q]) / p->k_cond;
        op.hconv = p->h_conv;
    };
    if (dtype == SCAFES_F64) {
        PennesAdjOp<double> op{};
        op.a = (const double*)lam;
        op.out = (double*)lam_new;
        op.x[0] = (const double*)w;
        op.x[1] = (const double*)tprev;
        op.grad = grad;
        op.ctd = p->tau * p->cdiff;
        op.ctp = p->tau * p->cperf;
        op.ta = p->t_a;
        op.hh0 = hh[0];
        op.hh1 = hh[1];
        op.hh2 = hh[2];
        tumor(op);
        return launch<PennesAdjOp<double>, double>(c, op, geo, -1);
    }
    auto fill32 = [&](auto& op) {
        op.a = (const float*)lam;
        op.out = (float*)lam_new;
        op.x[0] = (const float*)w;
        op.x[1] = (const float*)tprev;
        op.grad = grad;
        op.ctp = p->tau * p->cperf;
        op.ta = p->t_a;
        op.fd0 = (float)(p->tau * p->cdiff / hh[0]);
        op.fd1 = (float)(p->tau * p->cdiff / hh[1]);
        op.fd2 = (float)(p->tau * p->cdiff / hh[2]);
        op.fp = (float)(p->tau * p->cperf);
        tumor(op);
    };
    if (tprev_lo) {
        PennesAdjOp<float, true> op{};
        fill32(op);
        op.x[2] = (const float*)tprev_lo;
        return launch<PennesAdjOp<float, true>, float>(c, op, geo, -1);
    }
    PennesAdjOp<float> op{};
    fill32(op);
    return launch<PennesAdjOp<float>, float>(c, op, geo, -1);
}

// ---------------------------------------------------------------- init
__device__ inline uint64_t splitmix64(uint64_t seed, uint64_t idx)
{
    uint64_t z = seed + (idx + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

template <typename T>
__global__ void heatvar_init_kernel(Geo g, double hx, double hy, double hz, T* u, T* kap,
                                    uint64_t seed, double amp)
{
    const long total = (long)g.nx * g.ny * g.nz;
    const double pi = 3.14159265358979323846;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int i = (int)(t % g.nx);
        const long r = t / g.nx;
        const int j = (int)(r % g.ny);
        const int k = (int)(r / g.ny);
        const double x = hx * i, y = hy * j, z = hz * k;
        double v = sin(pi * x);
        if (g.dim > 1) v *= sin(pi * y);
        if (g.dim > 2) v *= sin(pi * z);
        if (!is_border(g, i, j, k)) {
            const double un = (double)(splitmix64(seed, (uint64_t)t) >> 11) * (1.0 / 9007199254740992.0);
            v += amp * (2.0 * un - 1.0);
        }
        const long m = at(g, i, j, k);
        u[m] = (T)v;
        kap[m] = (T)(1.0 + 0.5 * sin(2.0 * pi * x) * sin(2.0 * pi * y));
    }
}

}  // namespace scafes

using namespace scafes;

namespace scafes {

int pennes_check_tau(Ctx* c, const Geo& geo, const scafes_geom* g, int hw, int dtype, const void* w,
                     const scafes_pennes_params* p, bool use_cache, const char* who)
{
    double wmax = 0.0;
    if (p->cperf != 0.0) {
        const int rc = coef_abs_max(c, geo, hw, dtype == SCAFES_F64 ? SCAFES_F64 : SCAFES_F32, w, use_cache, &wmax);
        if (rc) return rc;
    }
    double rate = diffusion_rate(g, fabs(p->cdiff)) + fabs(p->cperf) * wmax;
    // the Cauchy face of the Tumor border (startSimulation.py:262-270)
    if (p->border_mode == SCAFES_BORDER_TUMOR && p->k_cond != 0.0)
        rate += 2.0 * fabs(p->cdiff * p->h_conv / p->k_cond) / g->h[g->dim - 1];
    return check_explicit_tau(p->tau, rate, who);
}

}  // namespace scafes

extern "C" {

// Argument checks and the operator of one constant-coefficient heat step
// (shared by the step and the run entry points).
static int heat7_const_op(Ctx* c, const scafes_geom* g, int hw, const double* u_old, double* u_new,
                          const double* f_old, const double* g_old, const double* g_new,
                          const scafes_heat7_params* p, Geo* geo, HeatConstOp* op)
{
    if (!c || !g || !u_old || !u_new || !p) return fail(SCAFES_ERR_ARG, "heat7: null pointer");
    if (u_old == u_new) return fail(SCAFES_ERR_ARG, "heat7: in-place update not allowed");
    int rc = check_stencil_geo(g, hw);
    if (rc) return rc;
    if (p->has_source && !f_old) return fail(SCAFES_ERR_ARG, "heat7: source requested without F");
    if (p->border_mode < SCAFES_BORDER_KEEP || p->border_mode > SCAFES_BORDER_RADIATION)
        return fail(SCAFES_ERR_ARG, "heat7: bad border mode");
    if (p->border_mode == SCAFES_BORDER_FIELD && !g_new)
        return fail(SCAFES_ERR_ARG, "heat7: border FIELD without G");
    if (p->border_mode >= SCAFES_BORDER_NEUMANN && (!g_new || !g_old))
        return fail(SCAFES_ERR_ARG, "heat7: flux border needs G_old and G_new");
    if (p->border_mode >= SCAFES_BORDER_NEUMANN && !(p->lambda != 0.0))
        return fail(SCAFES_ERR_ARG, "heat7: flux border needs lambda != 0");
    Geo gk;
    rc = make_geo(g, hw, geo);
    if (rc) return rc;
    rc = make_geo(g, 0, &gk);
    if (rc) return rc;
    rc = check_explicit_tau(p->tau, diffusion_rate(g, fabs(p->coeff_a)) +
                                        (p->has_perf ? fabs(p->perf_c * p->perf_w) : 0.0),
                            "heat7");
    if (rc) return rc;
    double hh[3];
    hh_of(g, hh);
    *op = HeatConstOp{};
    op->a = u_old;
    op->out = u_new;
    op->f_old = f_old;
    op->g_old = g_old;
    op->g_new = g_new;
    op->gk = gk;
    for (int q = 0; q < 3; ++q) op->a2[q] = q < g->dim ? 2.0 * g->h[q] / p->lambda : 0.0;
    op->alpha = p->alpha;
    op->epsilon = p->epsilon;
    op->sigma = p->sigma;
    op->ctau_a = p->tau * p->coeff_a;
    op->ctw = (p->tau * p->perf_c) * p->perf_w;
    op->ctf = p->tau * p->rhoc_inv;
    op->hh0 = hh[0];
    op->hh1 = hh[1];
    op->hh2 = hh[2];
    op->has_perf = p->has_perf;
    op->has_source = p->has_source;
    op->border_mode = p->border_mode;
    op->border_value = p->border_value;
    return SCAFES_OK;
}

int scafes_heat7_step_f64(scafes_ctx* ctx, const scafes_geom* g, int hw, const double* u_old,
                          double* u_new, const double* f_old, const double* g_old, const double* g_new,
                          const scafes_heat7_params* p)
{
    Ctx* c = as_ctx(ctx);
    Geo geo;
    HeatConstOp op;
    const int rc = heat7_const_op(c, g, hw, u_old, u_new, f_old, g_old, g_new, p, &geo, &op);
    if (rc) return rc;
    return launch<HeatConstOp, double>(c, op, geo, -1);
}

int scafes_heat7_run_f64(scafes_ctx* ctx, const scafes_geom* g, int hw, double* u, double* u2, int nsteps,
                         const scafes_heat7_params* p, int* result_in_second)
{
    Ctx* c = as_ctx(ctx);
    if (nsteps < 0) return fail(SCAFES_ERR_ARG, "heat7 run: negative step count");
    if (p && (p->has_source || p->border_mode == SCAFES_BORDER_FIELD || p->border_mode >= SCAFES_BORDER_NEUMANN))
        return fail(SCAFES_ERR_ARG, "heat7 run: known fields change per step; call scafes_heat7_step_f64");
    Geo geo;
    HeatConstOp op;
    int rc = heat7_const_op(c, g, hw, u, u2, nullptr, nullptr, nullptr, p, &geo, &op);
    if (rc) return rc;
    // One launch per step; a HIP graph of the launches (captured once,
    // replayed) was measured at heat 64^3 FP64, 100 steps: 0.488 ms against
    // 0.499 ms, so the steps are kernel-bound (about 5 us each), not
    // launch-bound, and the graph was not kept.
    double *a = u, *b = u2;
    for (int s = 0; s < nsteps; ++s) {
        op.a = a;
        op.out = b;
        rc = launch<HeatConstOp, double>(c, op, geo, -1);
        if (rc) return rc;
        std::swap(a, b);
    }
    if (result_in_second) *result_in_second = nsteps % 2;
    return SCAFES_OK;
}

// Configuration index space: [0, kNumCfgs) one-step tiles, then the
// two-steps-per-sweep tiles of heatvar_tb2.hip.
int scafes_heat7var_num_cfgs(void) { return kNumCfgs + heatvar_tb2_num_cfgs(); }

int scafes_heat7var_cfg_desc(int cfg, int* tx, int* ty, int* zchunk, int* depth, int* steps_per_sweep)
{
    if (cfg < 0 || cfg >= scafes_heat7var_num_cfgs()) return fail(SCAFES_ERR_ARG, "bad cfg");
    int a, b, z, d, st;
    if (cfg < kNumCfgs) {
        a = kCfgs[cfg].tx;
        b = kCfgs[cfg].ty;
        z = kCfgs[cfg].zc;
        d = kCfgs[cfg].d;
        st = 1;
    } else {
        heatvar_tb2_cfg(cfg - kNumCfgs, &a, &b, &z, &d, &st);
    }
    if (tx) *tx = a;
    if (ty) *ty = b;
    if (zchunk) *zchunk = z;
    if (depth) *depth = d;
    if (steps_per_sweep) *steps_per_sweep = st;
    return SCAFES_OK;
}

// The autotuned configuration of this grid and mode, or -1.
static int tuned_cfg(const Ctx* c, const Geo& geo, int mode)
{
    for (int i = 0; i < c->ntuned; ++i) {
        const Ctx::Tuned& t = c->tuned[i];
        if (t.mode == mode && t.nx == geo.nx && t.ny == geo.ny && t.nz == geo.nz && t.e0 == geo.e0 &&
            t.e01 == geo.e01)
            return t.cfg;
    }
    return -1;
}

// cfg < 0: the autotuned choice for this grid and mode, else the default; a
// multi-step choice on a grid that kernel does not cover falls back to the
// mode's one-step tile.
static int resolve_heatvar_cfg(const Ctx* c, const Geo& geo, int mode, int cfg)
{
    if (cfg < 0) {
        cfg = tuned_cfg(c, geo, mode);
        if (cfg < 0) cfg = kDefaultCfgByMode[mode];
    }
    if (cfg >= kNumCfgs && !heatvar_tb2_supported(geo, mode)) cfg = kOneStepCfgByMode[mode];
    return cfg;
}

int scafes_heat7var_active_cfg(scafes_ctx* ctx, const scafes_geom* g, int hw, int mode, int* cfg)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !g || !cfg) return fail(SCAFES_ERR_ARG, "heat7var_active_cfg: null pointer");
    if (mode < 0 || mode > 2) return fail(SCAFES_ERR_ARG, "heat7var_active_cfg: bad mode");
    Geo geo;
    const int rc = make_geo(g, hw, &geo);
    if (rc) return rc;
    *cfg = resolve_heatvar_cfg(c, geo, mode, -1);
    return SCAFES_OK;
}

int scafes_heat7var_plan(scafes_ctx* ctx, const scafes_geom* g, int hw, int mode, int nsteps, int* sweeps, int max,
                         int* count)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !g || !count || (max > 0 && !sweeps)) return fail(SCAFES_ERR_ARG, "heat7var_plan: null pointer");
    if (mode < 0 || mode > 2) return fail(SCAFES_ERR_ARG, "heat7var_plan: bad mode");
    if (nsteps < 0) return fail(SCAFES_ERR_ARG, "heat7var_plan: negative step count");
    Geo geo;
    const int rc = make_geo(g, hw, &geo);
    if (rc) return rc;
    int n = 0;
    for (long s = 0; s < nsteps; ++n) {
        const int k = heat7var_next_sweep(c, geo, mode, nsteps - s);
        if (n < max) sweeps[n] = k;
        s += k;
    }
    *count = n;
    return SCAFES_OK;
}

int scafes_heat7var_run(scafes_ctx* ctx, const scafes_geom* g, int hw, int mode, void* u, void* u2,
                        void* cmp, void* cmp2, const void* kappa, double tau, int nsteps,
                        int border_mode, double border_value, int cfg, int* result_in_second)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !g || !u || !u2 || !kappa) return fail(SCAFES_ERR_ARG, "heat7var: null pointer");
    if (mode < 0 || mode > 2) return fail(SCAFES_ERR_ARG, "heat7var: bad mode");
    if (mode == SCAFES_HEATVAR_F32_KAHAN && (!cmp || !cmp2))
        return fail(SCAFES_ERR_ARG, "heat7var: Kahan mode needs both compensation buffers");
    if (border_mode != SCAFES_BORDER_KEEP && border_mode != SCAFES_BORDER_VALUE)
        return fail(SCAFES_ERR_UNSUPPORTED, "heat7var: border must be KEEP or VALUE");
    if (nsteps < 0) return fail(SCAFES_ERR_ARG, "heat7var: negative step count");
    int rc = check_stencil_geo(g, hw);
    if (rc) return rc;
    Geo geo;
    rc = make_geo(g, hw, &geo);
    if (rc) return rc;
    if (cfg >= scafes_heat7var_num_cfgs()) return fail(SCAFES_ERR_ARG, "heat7var: bad cfg");
    if (nsteps > 0) {
        double kmax = 0.0;
        rc = coef_abs_max(c, geo, hw, mode == SCAFES_HEATVAR_F64 ? SCAFES_F64 : SCAFES_F32, kappa, true, &kmax);
        if (rc) return rc;
        rc = check_explicit_tau(tau, diffusion_rate(g, kmax), "heat7var");
        if (rc) return rc;
    }
    const bool planned = cfg < 0;  // the default or autotuned tile
    cfg = resolve_heatvar_cfg(c, geo, mode, cfg);
    void *a = u, *b = u2, *ca = cmp, *cb = cmp2;
    int launches = 0;
    int s = 0;
    if (planned && cfg >= kNumCfgs) {
        while (s < nsteps) {
            const int k = heat7var_next_sweep(c, geo, mode, nsteps - s);
            rc = heat7var_sweep_k(c, geo, g, mode, k, a, b, ca, cb, kappa, tau, border_mode, border_value);
            if (rc) return rc;
            std::swap(a, b);
            std::swap(ca, cb);
            s += k;
            ++launches;
        }
    } else if (cfg >= kNumCfgs) {
        int tx_, ty_, zc_, d_, st = 1;
        heatvar_tb2_cfg(cfg - kNumCfgs, &tx_, &ty_, &zc_, &d_, &st);
        for (; s + st <= nsteps; s += st, ++launches) {
            rc = heatvar_tb2_launch(c, geo, g, mode, a, b, ca, cb, kappa, tau, border_mode, border_value,
                                    cfg - kNumCfgs);
            if (rc) return rc;
            std::swap(a, b);
            std::swap(ca, cb);
        }
        // two steps left behind a three-step tile: one two-step sweep
        if (nsteps - s == 2) {
            rc = heatvar_tb2_launch(c, geo, g, mode, a, b, ca, cb, kappa, tau, border_mode, border_value,
                                    kTwoStepCfgByMode[mode] - kNumCfgs);
            if (rc) return rc;
            std::swap(a, b);
            std::swap(ca, cb);
            s += 2;
            ++launches;
        }
        cfg = kOneStepCfgByMode[mode];  // a last step one per sweep
    }
    for (; s < nsteps; ++s, ++launches) {
        rc = heat7var_launch(c, geo, g, mode, a, b, ca, cb, kappa, tau, border_mode, border_value, cfg);
        if (rc) return rc;
        std::swap(a, b);
        std::swap(ca, cb);
    }
    if (result_in_second) *result_in_second = (launches % 2 == 1) ? 1 : 0;
    return SCAFES_OK;
}

int scafes_autotune_heat7var(scafes_ctx* ctx, const scafes_geom* g, int hw, int mode, void* u,
                             void* u2, void* cmp, void* cmp2, const void* kappa, double tau,
                             int nsteps, int* best_cfg, float* best_ms)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !g) return fail(SCAFES_ERR_ARG, "autotune: null pointer");
    if (mode < 0 || mode > 2) return fail(SCAFES_ERR_ARG, "autotune: bad mode");
    Geo geo;
    int rc = make_geo(g, hw, &geo);
    if (rc) return rc;
    // a multiple of 24 runs every configuration (1, 2, 3 or 4 steps per
    // sweep) without a remainder sweep, and an even sweep count keeps u in
    // place
    const int steps = 24 * std::max(1, (nsteps + 23) / 24);
    hipEvent_t e0, e1;
    SCAFES_HIP(hipEventCreate(&e0));
    SCAFES_HIP(hipEventCreate(&e1));
    int best = -1;
    float bms = 1e30f;
    const int ncfg = scafes_heat7var_num_cfgs();
    for (int cfg = 0; cfg < ncfg && rc == SCAFES_OK; ++cfg) {
        int st = 1, tx_, ty_, zc_, d_;
        scafes_heat7var_cfg_desc(cfg, &tx_, &ty_, &zc_, &d_, &st);
        if (st > 1 && !heatvar_tb2_supported(geo, mode)) continue;  // would run the one-step tile
        int which = 0;
        // warm-up: one launch of each sweep the timed run makes
        rc = scafes_heat7var_run(ctx, g, hw, mode, u, u2, cmp, cmp2, kappa, tau, 2 * st, SCAFES_BORDER_KEEP, 0.0,
                                 cfg, &which);
        if (rc) break;
        hipError_t e = hipEventRecord(e0, c->stream);
        if (e == hipSuccess)
            rc = scafes_heat7var_run(ctx, g, hw, mode, u, u2, cmp, cmp2, kappa, tau, steps, SCAFES_BORDER_KEEP, 0.0,
                                     cfg, &which);
        if (rc) break;
        if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
            rc = check_hip(e, "autotune timing");
            break;
        }
        ms /= (float)steps;
        if (ms < bms) {
            bms = ms;
            best = cfg;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best < 0) return fail(SCAFES_ERR_UNSUPPORTED, "autotune: no configuration runs on this grid");
    // keep the choice for this grid and mode (replacing an earlier one)
    int slot = -1;
    for (int i = 0; i < c->ntuned; ++i) {
        const Ctx::Tuned& t = c->tuned[i];
        if (t.mode == mode && t.nx == geo.nx && t.ny == geo.ny && t.nz == geo.nz && t.e0 == geo.e0 &&
            t.e01 == geo.e01)
            slot = i;
    }
    if (slot < 0) {
        if (c->ntuned < 8) {
            slot = c->ntuned++;
        } else {
            for (int i = 0; i < 7; ++i) c->tuned[i] = c->tuned[i + 1];  // drop the oldest
            slot = 7;
        }
    }
    c->tuned[slot] = Ctx::Tuned{mode, geo.nx, geo.ny, geo.nz, geo.e0, geo.e01, best};
    if (best_cfg) *best_cfg = best;
    if (best_ms) *best_ms = bms;
    return SCAFES_OK;
}

int scafes_heat7var_init(scafes_ctx* ctx, const scafes_geom* g, int hw, int dtype, void* u, void* kappa,
                         uint64_t seed, double amp)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !g || !u || !kappa) return fail(SCAFES_ERR_ARG, "init: null pointer");
    Geo geo;
    int rc = make_geo(g, hw, &geo);
    if (rc) return rc;
    coef_invalidate(c, nullptr, 0);  // kappa is rewritten
    const long total = (long)geo.nx * geo.ny * geo.nz;
    const int blocks = (int)std::min<long>((total + 255) / 256, 65536);
    if (dtype == SCAFES_F64)
        heatvar_init_kernel<double><<<blocks, 256, 0, c->stream>>>(geo, g->h[0], g->h[1], g->h[2],
                                                                   (double*)u, (double*)kappa, seed, amp);
    else
        heatvar_init_kernel<float><<<blocks, 256, 0, c->stream>>>(geo, g->h[0], g->h[1], g->h[2],
                                                                  (float*)u, (float*)kappa, seed, amp);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

int scafes_pennes_step(scafes_ctx* ctx, const scafes_geom* g, int hw, int dtype, const void* t_old,
                       void* t_new, const void* w, const void* q, const scafes_pennes_params* p)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !g || !t_old || !t_new || !w || !q || !p) return fail(SCAFES_ERR_ARG, "pennes: null pointer");
    if (t_old == t_new) return fail(SCAFES_ERR_ARG, "pennes: in-place update not allowed");
    int rc = check_stencil_geo(g, hw);
    if (rc) return rc;
    Geo geo;
    rc = make_geo(g, hw, &geo);
    if (rc) return rc;
    if (dtype != SCAFES_F32 && dtype != SCAFES_F64) return fail(SCAFES_ERR_ARG, "pennes: dtype F32 or F64");
    rc = pennes_check_tau(c, geo, g, hw, dtype, w, p, true, "pennes");
    if (rc) return rc;
    return pennes_step_launch(c, geo, g, dtype, t_old, t_new, w, q, p, nullptr, nullptr);
}

int scafes_pennes_step_kahan(scafes_ctx* ctx, const scafes_geom* g, int hw, const float* t_old, float* t_new,
                             const float* c_old, float* c_new, const float* w, const float* q,
                             const scafes_pennes_params* p)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !g || !t_old || !t_new || !c_old || !c_new || !w || !q || !p)
        return fail(SCAFES_ERR_ARG, "pennes kahan: null pointer");
    if (t_old == t_new || c_old == c_new) return fail(SCAFES_ERR_ARG, "pennes kahan: in-place update not allowed");
    int rc = check_stencil_geo(g, hw);
    if (rc) return rc;
    Geo geo;
    rc = make_geo(g, hw, &geo);
    if (rc) return rc;
    rc = pennes_check_tau(c, geo, g, hw, SCAFES_F32, w, p, true, "pennes kahan");
    if (rc) return rc;
    return pennes_step_launch(c, geo, g, SCAFES_F32_KAHAN, t_old, t_new, w, q, p, c_old, c_new);
}

int scafes_mri_coef_f64(scafes_ctx* ctx, const scafes_geom* g, int hw, const double* const* mat,
                        const scafes_mri_params* p, double* coef)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !g || !mat || !p || !coef) return fail(SCAFES_ERR_ARG, "mri coef: null pointer");
    for (int f = 0; f < SCAFES_MRI_SURFACE; ++f)
        if (!mat[f]) return fail(SCAFES_ERR_ARG, "mri coef: null material field");
    Geo geo, gk;
    int rc = make_geo(g, hw, &geo);
    if (rc) return rc;
    rc = make_geo(g, 0, &gk);
    if (rc) return rc;
    long ne = 1;
    for (int q = 0; q < g->dim; ++q) ne *= (long)g->n[q] + 2L * hw;
    const long nn = (long)geo.nx * geo.ny * geo.nz;
    const int blocks = (int)std::min<long>(ceil_div(nn, 256L), 8192L);
    mri_coef_kernel<<<blocks, 256, 0, c->stream>>>(geo, gk, mat[SCAFES_MRI_RHO], mat[SCAFES_MRI_C],
                                                   mat[SCAFES_MRI_LAMBDA], mat[SCAFES_MRI_RHO_BLOOD],
                                                   mat[SCAFES_MRI_C_BLOOD], mat[SCAFES_MRI_OMEGA],
                                                   mat[SCAFES_MRI_T_BLOOD], mat[SCAFES_MRI_Q], p->tau, coef, ne);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

int scafes_mri_step_f64(scafes_ctx* ctx, const scafes_geom* g, int hw, const double* t_old, double* t_new,
                        const double* const* mat, const double* coef, const scafes_mri_params* p)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !g || !t_old || !t_new || !mat || !coef || !p) return fail(SCAFES_ERR_ARG, "mri: null pointer");
    if (!mat[SCAFES_MRI_LAMBDA] || !mat[SCAFES_MRI_SURFACE])
        return fail(SCAFES_ERR_ARG, "mri: the border needs lambda and surface");
    if (t_old == t_new) return fail(SCAFES_ERR_ARG, "mri: in-place update not allowed");
    int rc = check_stencil_geo(g, hw);
    if (rc) return rc;
    Geo geo, gk;
    rc = make_geo(g, hw, &geo);
    if (rc) return rc;
    rc = make_geo(g, 0, &gk);
    if (rc) return rc;
    long ne = 1;
    for (int q = 0; q < g->dim; ++q) ne *= (long)g->n[q] + 2L * hw;
    double hh[3];
    hh_of(g, hh);
    MriOp op{};
    op.a = t_old;
    op.out = t_new;
    for (int f = 0; f < 4; ++f) op.x[f] = coef + f * ne;
    op.lam = mat[SCAFES_MRI_LAMBDA];
    op.surf = mat[SCAFES_MRI_SURFACE];
    op.gk = gk;
    op.hh0 = hh[0];
    op.hh1 = hh[1];
    op.hh2 = hh[2];
    for (int q = 0; q < 3; ++q) op.h[q] = q < g->dim ? g->h[q] : 1.0;
    op.hconv = p->h_conv;
    op.tamb = p->t_amb;
    op.qbc = p->q_bc;
    op.qskull = p->q_skull;
    op.epsilon = p->epsilon;
    op.sigma = p->sigma;
    const double ta = p->t_amb + 273.15;
    op.tamb4 = ((ta * ta) * ta) * ta;
    return launch<MriOp, double>(c, op, geo, -1);
}

int scafes_pennes_adjoint_step(scafes_ctx* ctx, const scafes_geom* g, int hw, int dtype, const void* lam,
                               void* lam_new, const void* w, const void* t_prev, double* grad,
                               const scafes_pennes_params* p)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !g || !lam || !lam_new || !w || !t_prev || !grad || !p)
        return fail(SCAFES_ERR_ARG, "pennes adjoint: null pointer");
    if (lam == lam_new) return fail(SCAFES_ERR_ARG, "pennes adjoint: in-place update not allowed");
    int rc = check_stencil_geo(g, hw);
    if (rc) return rc;
    Geo geo;
    rc = make_geo(g, hw, &geo);
    if (rc) return rc;
    if (dtype != SCAFES_F32 && dtype != SCAFES_F64) return fail(SCAFES_ERR_ARG, "pennes adjoint: dtype F32 or F64");
    return pennes_adj_launch(c, geo, g, dtype, lam, lam_new, w, t_prev, grad, p, nullptr);
}

}  // extern "C"

namespace scafes {

// Steps per sweep of a tile configuration index (one-step tiles: 1).
static int cfg_steps(int cfg)
{
    if (cfg < kNumCfgs) return 1;
    int tx, ty, zc, d, st = 1;
    heatvar_tb2_cfg(cfg - kNumCfgs, &tx, &ty, &zc, &d, &st);
    return st;
}

// Behind a three-step tile, a remainder of two steps takes one two-step
// sweep; where the mode has a four-step tile of about the same cost per
// step, step counts that are not a multiple of three are made of three-
// and four-step sweeps instead (4 = 4, 7 = 4 + 3, 8 = 4 + 4, 20 = 4 + 4 +
// 4 x 3; only 5 keeps 3 + 2), so no short sweep pays for a whole pass over
// HBM.  A tuned four-step tile takes a remainder of three with the mode's
// three-step default tile, of two with its two-step tile.
int heat7var_next_sweep(const Ctx* c, const Geo& geo, int mode, long left)
{
    const int st = cfg_steps(resolve_heatvar_cfg(c, geo, mode, -1));
    if (left <= 1 || st == 1) return 1;
    if (st == 2) return 2;
    if (st == 4) return left >= 4 ? 4 : (int)left;
    if (kFourStepCfgByMode[mode] >= 0 && left >= 4 && (left % 3 == 1 || (left % 3 == 2 && left >= 8))) return 4;
    return left >= 3 ? 3 : 2;
}

int heat7var_sweep_k(Ctx* c, const Geo& geo, const scafes_geom* g, int mode, int k, const void* u, void* un,
                     const void* cin, void* cout, const void* kap, double tau, int border_mode, double border_value)
{
    if (k == 1) return heat7var_step_one(c, geo, g, mode, u, un, cin, cout, kap, tau, border_mode, border_value);
    int cfg = resolve_heatvar_cfg(c, geo, mode, -1);
    if (cfg_steps(cfg) != k) {
        if (k == 2) cfg = kTwoStepCfgByMode[mode];
        else if (k == 3) cfg = kDefaultCfgByMode[mode];
        else if (k == 4) cfg = kFourStepCfgByMode[mode];
        else cfg = -1;
    }
    if (cfg < kNumCfgs || cfg_steps(cfg) != k || !heatvar_tb2_supported(geo, mode))
        return fail(SCAFES_ERR_UNSUPPORTED, "no sweep of this many steps for this grid and mode");
    return heatvar_tb2_launch(c, geo, g, mode, u, un, cin, cout, kap, tau, border_mode, border_value,
                              cfg - kNumCfgs);
}

bool heat7var_states_sweep_ok(const Ctx* c, const Geo& geo, int mode)
{
    const int cfg = resolve_heatvar_cfg(c, geo, mode, -1);
    return cfg >= kNumCfgs && cfg_steps(cfg) == 3 && heatvar_tb_states_supported(cfg - kNumCfgs);
}

bool heat7var_states4_ok(const Ctx* c, const Geo& geo, int mode)
{
    const int f = kFourStepCfgByMode[mode];
    return heat7var_states_sweep_ok(c, geo, mode) && f >= kNumCfgs && heatvar_tb_states_supported(f - kNumCfgs);
}

int heat7var_sweep_states(Ctx* c, const Geo& geo, const scafes_geom* g, int mode, int k, const void* u, void* un,
                          const void* cin, void* cout, void* const* mid, const void* kap, double tau,
                          int border_mode, double border_value)
{
    int cfg = -1;
    if (k == 3 && heat7var_states_sweep_ok(c, geo, mode)) cfg = resolve_heatvar_cfg(c, geo, mode, -1);
    if (k == 4 && heat7var_states4_ok(c, geo, mode)) cfg = kFourStepCfgByMode[mode];
    if (cfg < kNumCfgs) return fail(SCAFES_ERR_UNSUPPORTED, "no state-writing sweep of this many steps for this grid");
    return heatvar_tb2_launch(c, geo, g, mode, u, un, cin, cout, kap, tau, border_mode, border_value,
                              cfg - kNumCfgs, mid);
}

}  // namespace scafes
