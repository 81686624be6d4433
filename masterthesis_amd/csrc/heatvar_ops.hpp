// heatvar_ops.hpp -- per-node arithmetic of the variable-kappa heat update
// and of its kappa adjoint, shared by the multi-step kernels (heatvar_tb2.hip,
// heatvar_adjk.hip); the one-step operators in stencil3d.hip use the same
// operation order.
#ifndef SCAFES_HEATVAR_OPS_HPP_
#define SCAFES_HEATVAR_OPS_HPP_

#include <hip/hip_runtime.h>

namespace scafes {

struct VarCoef {
    double tau, hh0, hh1, hh2;  // FP64 path (CPU operation order)
    float c0, c1, c2;           // 0.5*tau/h_p^2, FP32 paths
    double border_value;
    int border_mode;
};

// Face coefficient between the centre and one neighbour: 0.5*(k_nb + k_c)
// in the FP64 path (the CPU path's grouping), k_nb + k_c in FP32 (the 0.5 is
// folded into c_p).  IEEE addition is commutative, so the face shared by two
// nodes has one value.
template <typename T, int MODE>
__device__ inline T face(T a, T b)
{
    if constexpr (MODE == 0)
        return 0.5 * (a + b);
    else
        return a + b;
}

template <typename T>
struct Faces {
    T e, w, n, s, u, d;
};

// One explicit step at one interior node from its six face coefficients.
// MODE 0: FP64 in the CPU path's order; 1: FP32; 2: FP32 + Kahan (cin -> *cout).
template <typename T, int MODE>
__device__ inline T var_update(const VarCoef& cf, T uc, T uw, T ue, T us, T un, T ud, T uu, const Faces<T>& f, T cin,
                               T* cout)
{
    if constexpr (MODE == 0) {
        double v = uc;
        v += (cf.tau * (f.e * (ue - uc) - f.w * (uc - uw))) / cf.hh0;
        v += (cf.tau * (f.n * (un - uc) - f.s * (uc - us))) / cf.hh1;
        v += (cf.tau * (f.u * (uu - uc) - f.d * (uc - ud))) / cf.hh2;
        *cout = T(0);
        return (T)v;
    } else {
        float d = cf.c0 * __builtin_fmaf(f.e, ue - uc, -(f.w * (uc - uw)));
        d = __builtin_fmaf(cf.c1, __builtin_fmaf(f.n, un - uc, -(f.s * (uc - us))), d);
        d = __builtin_fmaf(cf.c2, __builtin_fmaf(f.u, uu - uc, -(f.d * (uc - ud))), d);
        if constexpr (MODE == 2) {
            const float y = d + cin;
            const float t = uc + y;
            *cout = y - (t - uc);
            return t;
        } else {
            *cout = 0.0f;
            return uc + d;
        }
    }
}

struct AdjVarCoef {
    VarCoef v;                // lam update (the forward step's arithmetic)
    double th0, th1, th2;     // tau / h_p^2
    float f0, f1, f2;         // 0.5 * th_p rounded to FP32 (exactly half of th_p rounded)
};

__device__ inline double adj_face_sum(bool lo, bool hi, double lc, double lm, double lp, double uc, double um,
                                      double up)
{
    double sp = 0.0;
    if (lo) sp += (lc - lm) * (um - uc);
    if (hi) sp += (lc - lp) * (up - uc);
    return sp;
}

// The same face sum in FP32 (FP32 fields).
__device__ inline float adj_face_sum_f(bool lo, bool hi, float lc, float lm, float lp, float uc, float um, float up)
{
    const float a = lo ? (lc - lm) * (um - uc) : 0.0f;
    const float b = hi ? (lc - lp) * (up - uc) : 0.0f;
    return a + b;
}

// Kappa-gradient term of one adjoint step at a node of a 3-D grid,
//   t = sum_p th_p * sum_{nb = -+e_p in the grid} (lam_c - lam_nb)(u_nb - u_c);
// returns the increment 0.5 t the FP64 gradient adds.  FP64 fields: t in
// FP64 in the oracle's order, then 0.5 t.  FP32 fields: 0.5 t in FP32 from
// the FP32 values with the coefficients 0.5 th_p rounded to FP32 and the
// axes joined with fmaf, which is 0.5 times the term formed with th_p
// rounded to FP32 (scaling by 2^-1 is exact at every rounding step away
// from the subnormal range), so no FP64 multiply is left per term (the
// one-step operator in stencil3d.hip uses the same arithmetic, so every
// schedule gives the same gradient).
template <typename T>
__device__ inline double adj_term3(const AdjVarCoef& cf, bool xlo, bool xhi, bool ylo, bool yhi, bool zlo, bool zhi,
                                   T lc, T lw, T le, T ls, T ln, T ld, T lu, T uc, T uw, T ue, T us, T un, T ud,
                                   T uu)
{
    if constexpr (sizeof(T) == 8) {
        double t = 0.0;
        t += cf.th0 * adj_face_sum(xlo, xhi, lc, lw, le, uc, uw, ue);
        t += cf.th1 * adj_face_sum(ylo, yhi, lc, ls, ln, uc, us, un);
        t += cf.th2 * adj_face_sum(zlo, zhi, lc, ld, lu, uc, ud, uu);
        return 0.5 * t;
    } else {
        float t = cf.f0 * adj_face_sum_f(xlo, xhi, lc, lw, le, uc, uw, ue);
        t = __builtin_fmaf(cf.f1, adj_face_sum_f(ylo, yhi, lc, ls, ln, uc, us, un), t);
        t = __builtin_fmaf(cf.f2, adj_face_sum_f(zlo, zhi, lc, ld, lu, uc, ud, uu), t);
        return (double)t;  // 0.5 t: the halved coefficients scale every rounding step by 2^-1 exactly
    }
}

}  // namespace scafes
#endif
