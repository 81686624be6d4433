// pennes_adj2.hip -- two reverse-time Pennes adjoint steps per HBM sweep
// (lam^n -> lam^{n-2}) with the gradient contributions of both steps, for
// 3-D grids.  Per node the arithmetic is that of PennesAdjOp in
// stencil3d.hip (the one-step adjoint), in the same order, and the FP64
// gradient receives the step-n term before the step-(n-1) term, so the
// result is bit-identical to two one-step launches.
//
//   lam^{n-1}_i = lam^n_i + sum_p ctd*((lam_W+lam_E) - 2 lam_i)/h_p^2 - (ctp*w_i)*lam_i
//   grad_i     += lam^n_i * (ctp*(T_a - T^{n-1}_i))        (inner nodes)
//   lam^{n-1} = 0 on border nodes.
//
// Same skeleton as heatvar_tb2.hip: level 1 on the TX x TY thread tile
// (output tile plus a one-node ring) from an LDS plane of lam^n, level 2 on
// the (TX-2) x (TY-2) output tile one plane behind from the LDS plane of
// level 1; z neighbours ride in register queues.  Per two steps HBM moves
// lam r+w, w, T^{n-1}, T^{n-2} and the FP64 gradient r+w once.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "scafes_common.hpp"

namespace scafes {

namespace {

using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 family: 32-bit data format, raw (stride 0) buffer

__device__ inline Rsrc plane_rsrc(const void* plane, uint32_t range)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(plane), (short)0, (int)range, kRsrcWord3);
}

template <typename T>
__device__ inline T bload(Rsrc r, uint32_t off)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

template <typename T>
__device__ inline void bstore(T v, Rsrc r, uint32_t off)
{
    if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)off, 0, 0);
    } else {
        typedef int v2i __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2i, v), r, (int)off, 0, 0);
    }
}

struct AdjCoef {
    double ctd, ctp, ta, hh0, hh1, hh2;  // FP64 path
    float fd0, fd1, fd2, fp;             // FP32 path: ctd/h_p^2, ctp
};

// One adjoint step at an inner node (PennesAdjOp's order).
template <typename T>
__device__ inline T adj_update(const AdjCoef& cf, T c, T w_, T e, T s, T n, T d, T u, T wf)
{
    if constexpr (sizeof(T) == 8) {
        double v = c;
        v += (cf.ctd * ((w_ + e) - 2.0 * c)) / cf.hh0;
        v += (cf.ctd * ((s + n) - 2.0 * c)) / cf.hh1;
        v += (cf.ctd * ((d + u) - 2.0 * c)) / cf.hh2;
        v -= (cf.ctp * wf) * c;
        return v;
    } else {
        float dd = cf.fd0 * ((w_ + e) - 2.0f * c);
        dd = __builtin_fmaf(cf.fd1, (s + n) - 2.0f * c, dd);
        dd = __builtin_fmaf(cf.fd2, (d + u) - 2.0f * c, dd);
        dd = __builtin_fmaf(-cf.fp * wf, c, dd);
        return c + dd;
    }
}

template <typename T, bool LO, int TX, int TY, int D>
__global__ __launch_bounds__(TX* TY) void pennes_adj2_kernel(Geo g, const T* __restrict__ lam, T* __restrict__ lamo,
                                                              const T* __restrict__ wf, const T* __restrict__ t1,
                                                              const T* __restrict__ t1lo, const T* __restrict__ t2,
                                                              const T* __restrict__ t2lo, double* __restrict__ grad,
                                                              AdjCoef cf, int zc, int tiles_x, int tiles_y)
{
    static_assert(D >= 2 && D % 2 == 0, "prefetch depth must be even");
    constexpr int OX = TX - 2, OY = TY - 2;
    __shared__ T s0[2][TY + 2][TX + 2];  // lam^n plane with ring
    __shared__ T s1[2][TY][TX];          // lam^{n-1} plane
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per_plane = tiles_x * tiles_y;
    const int tz = lb / per_plane;
    const int r = lb - tz * per_plane;
    const int tyi = r / tiles_x;
    const int txi = r - tyi * tiles_x;
    const int i = txi * OX - 1 + tx, j = tyi * OY - 1 + ty;
    const int ic = min(max(i, 0), g.nx - 1), jc = min(max(j, 0), g.ny - 1);
    const bool inner_xy = i >= g.nl0 && i <= g.nx - 1 - g.nl0 && j >= g.nl1 && j <= g.ny - 1 - g.nl1;
    const bool out_thread = tx >= 1 && tx <= TX - 2 && ty >= 1 && ty <= TY - 2 && i < g.nx && j < g.ny;
    const int k0 = tz * zc, k1 = min(k0 + zc, g.nz);
    const int kstart = k0 - 1, kcen = k1 + 1;
    const int klo = g.nl2, khi = g.nz - 1 - g.nl2;
    const bool hx = (tx == 0) || (tx == TX - 1);
    const bool hy = (ty == 0) || (ty == TY - 1);
    const int hxi = (tx == 0) ? max(ic - 1, 0) : min(ic + 1, g.nx - 1);
    const int hyj = (ty == 0) ? max(jc - 1, 0) : min(jc + 1, g.ny - 1);
    const int sxh = (tx == 0) ? 0 : TX + 1;
    const int syh = (ty == 0) ? 0 : TY + 1;
    constexpr uint32_t B = sizeof(T);
    const uint32_t oc = (uint32_t)(g.base + ic + g.e0 * jc) * B;
    const uint32_t ohx = (uint32_t)(g.base + hxi + g.e0 * jc) * B;
    const uint32_t ohy = (uint32_t)(g.base + ic + g.e0 * hyj) * B;
    const uint32_t range = (uint32_t)((g.base + g.e01) * B);
    const uint32_t og = (uint32_t)(g.base + ic + g.e0 * jc) * 8u;  // FP64 gradient
    const uint32_t grange = (uint32_t)((g.base + g.e01) * 8);
    const size_t pb = (size_t)g.e01 * B, pg = (size_t)g.e01 * 8;
    auto pl = [&](int kk) { return (size_t)min(max(kk, 0), g.nz - 1); };

    T qa[D], qw[D], qax[D], qay[D], q1[D], q1l[D], q2[D], q2l[D];
    double qg[D];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        qa[s] = qw[s] = qax[s] = qay[s] = q1[s] = q1l[s] = q2[s] = q2l[s] = T(0);
        qg[s] = 0.0;
    }
    auto fetch = [&](const int slot, const int kk) {
        if (kk <= kcen) {
            const size_t p = pl(kk);
            const Rsrc rl = plane_rsrc((const char*)lam + p * pb, range);
            qa[slot] = bload<T>(rl, oc);
            if (kk <= k1) {
                qw[slot] = bload<T>(plane_rsrc((const char*)wf + p * pb, range), oc);
                if (hx) qax[slot] = bload<T>(rl, ohx);
                if (hy) qay[slot] = bload<T>(rl, ohy);
                if (out_thread && kk >= k0 && kk < k1) {
                    q1[slot] = bload<T>(plane_rsrc((const char*)t1 + p * pb, range), oc);
                    q2[slot] = bload<T>(plane_rsrc((const char*)t2 + p * pb, range), oc);
                    if constexpr (LO) {
                        q1l[slot] = bload<T>(plane_rsrc((const char*)t1lo + p * pb, range), oc);
                        q2l[slot] = bload<T>(plane_rsrc((const char*)t2lo + p * pb, range), oc);
                    }
                    qg[slot] = bload<double>(plane_rsrc((const char*)grad + p * pg, grange), og);
                }
            }
        }
    };
    T am = bload<T>(plane_rsrc((const char*)lam + pl(kstart - 1) * pb, range), oc);
#pragma unroll
    for (int s = 0; s < D; ++s) fetch(s, kstart + s);

    T l1m2 = T(0), l1m1 = T(0);  // level 1 (lam^{n-1}) at k-2, k-1
    T wm1 = T(0);                // w at k-1
    T t2m1 = T(0), t2lm1 = T(0); // T^{n-2} at k-1
    double gm1 = 0.0, g1m1 = 0.0;  // gradient at k-1 and its step-n term
    for (int kb = kstart; kb <= k1; kb += D) {
#pragma unroll
        for (int s = 0; s < D; ++s) {
            const int k = kb + s;
            if (k > k1) break;  // uniform over the workgroup
            const int buf = s & 1;
            const T ac = qa[s], ap = qa[(s + 1) % D];
            const T wc = qw[s];
            const T tp1 = q1[s], tp1l = q1l[s];
            const T tp2 = q2[s], tp2l = q2l[s];
            const double gc = qg[s];
            s0[buf][ty + 1][tx + 1] = ac;
            if (hx) s0[buf][ty + 1][sxh] = qax[s];
            if (hy) s0[buf][syh][tx + 1] = qay[s];
            fetch(s, k + D);
            __syncthreads();
            // ---- level 1: lam^{n-1} at (i, j, k) and the step-n gradient term
            const bool in1 = inner_xy && k >= klo && k <= khi;
            T l1 = adj_update<T>(cf, ac, s0[buf][ty + 1][tx], s0[buf][ty + 1][tx + 2], s0[buf][ty][tx + 1],
                                 s0[buf][ty + 2][tx + 1], am, ap, wc);
            if (!in1) l1 = T(0);
            s1[buf][ty][tx] = l1;
            double g1 = 0.0;
            if (out_thread && in1) {
                double tp = (double)tp1;
                if constexpr (LO) tp += (double)tp1l;
                g1 = (double)ac * (cf.ctp * (cf.ta - tp));
            }
            // ---- level 2: lam^{n-2} at (i, j, k-1), both gradient terms
            const int k2 = k - 1;
            if (k2 >= k0 && out_thread) {
                const int pbf = buf ^ 1;
                const bool in2 = inner_xy && k2 >= klo && k2 <= khi;
                T v = T(0);
                const uint32_t oo = oc;  // out threads are never clamped
                if (in2) {
                    v = adj_update<T>(cf, l1m1, s1[pbf][ty][tx - 1], s1[pbf][ty][tx + 1], s1[pbf][ty - 1][tx],
                                      s1[pbf][ty + 1][tx], l1m2, l1, wm1);
                    double tp = (double)t2m1;
                    if constexpr (LO) tp += (double)t2lm1;
                    double gv = gm1 + g1m1;
                    gv += (double)l1m1 * (cf.ctp * (cf.ta - tp));
                    bstore<double>(gv, plane_rsrc((const char*)grad + (size_t)k2 * pg, grange), og);
                }
                bstore<T>(v, plane_rsrc((const char*)lamo + (size_t)k2 * pb, range), oo);
            }
            l1m2 = l1m1;
            l1m1 = l1;
            wm1 = wc;
            t2m1 = tp2;
            t2lm1 = tp2l;
            gm1 = gc;
            g1m1 = g1;
            am = ac;
        }
    }
}

constexpr int kATX = 64, kATY = 8, kAZC = 64, kAD = 2;

template <typename T, bool LO>
int adj2_launch_t(Ctx* c, const Geo& g, const void* lam, void* lamo, const void* w, const void* t1, const void* t1lo,
                  const void* t2, const void* t2lo, double* grad, const AdjCoef& cf)
{
    const int tiles_x = ceil_div(g.nx, kATX - 2), tiles_y = ceil_div(g.ny, kATY - 2), tiles_z = ceil_div(g.nz, kAZC);
    const long nb = (long)tiles_x * tiles_y * tiles_z;
    if (nb > 0x7fffffffL) return fail(SCAFES_ERR_ARG, "grid too large");
    pennes_adj2_kernel<T, LO, kATX, kATY, kAD><<<dim3((unsigned)nb), dim3(kATX, kATY), 0, c->stream>>>(
        g, (const T*)lam, (T*)lamo, (const T*)w, (const T*)t1, (const T*)t1lo, (const T*)t2, (const T*)t2lo, grad, cf,
        kAZC, tiles_x, tiles_y);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

}  // namespace

bool pennes_adj2_supported(const Geo& geo)
{
    // 3-D only; in-plane byte offsets of the FP64 gradient must fit 32 bits
    return geo.dim == 3 && (geo.base + geo.e01) * 8L < (1L << 32);
}

int pennes_adj2_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, const void* lam, void* lam_out,
                       const void* w, const void* t1, const void* t1lo, const void* t2, const void* t2lo,
                       double* grad, const scafes_pennes_params* p)
{
    if (!pennes_adj2_supported(geo)) return fail(SCAFES_ERR_UNSUPPORTED, "two-step adjoint needs a 3-D grid");
    if (p->border_mode != SCAFES_BORDER_KEEP)
        return fail(SCAFES_ERR_UNSUPPORTED, "two-step adjoint: only border_mode KEEP has an adjoint");
    if ((t1lo == nullptr) != (t2lo == nullptr)) return fail(SCAFES_ERR_ARG, "two-step adjoint: lo parts mismatch");
    AdjCoef cf;
    double hh[3];
    for (int q = 0; q < 3; ++q) hh[q] = g->h[q] * g->h[q];
    cf.ctd = p->tau * p->cdiff;
    cf.ctp = p->tau * p->cperf;
    cf.ta = p->t_a;
    cf.hh0 = hh[0];
    cf.hh1 = hh[1];
    cf.hh2 = hh[2];
    cf.fd0 = (float)(cf.ctd / hh[0]);
    cf.fd1 = (float)(cf.ctd / hh[1]);
    cf.fd2 = (float)(cf.ctd / hh[2]);
    cf.fp = (float)cf.ctp;
    if (dtype == SCAFES_F64) {
        if (t1lo) return fail(SCAFES_ERR_ARG, "two-step adjoint: FP64 states have no lo part");
        return adj2_launch_t<double, false>(c, geo, lam, lam_out, w, t1, nullptr, t2, nullptr, grad, cf);
    }
    if (t1lo) return adj2_launch_t<float, true>(c, geo, lam, lam_out, w, t1, t1lo, t2, t2lo, grad, cf);
    return adj2_launch_t<float, false>(c, geo, lam, lam_out, w, t1, nullptr, t2, nullptr, grad, cf);
}

}  // namespace scafes
