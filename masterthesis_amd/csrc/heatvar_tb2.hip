// heatvar_tb2.hip -- two time steps per HBM sweep for the variable-kappa heat
// update (temporal blocking), same arithmetic per node as the one-step
// kernel in stencil3d.hip, so FP64 results stay bit-identical to the CPU path.
//
// A TX x TY workgroup streams z planes.  Level 1 (u after one step) is
// computed for every thread position, i.e. for the output tile plus a
// one-node ring; level 2 (after two steps) for the (TX-2) x (TY-2) interior,
// which is the output tile.  Per plane k:
//   - level-0 plane k (u, kappa, plus a one-node ring) goes to LDS from a
//     D-deep register prefetch ring; centre-only Kahan compensation too;
//   - level 1 at plane k from LDS neighbours and the z queue (k-1, k+1);
//   - level 2 at plane k-1 from the level-1 LDS plane written one plane
//     before (in-plane neighbours) and a level-1 z queue (k-2, k).
// One barrier per plane: u planes are double-buffered in LDS; kappa planes are
// triple-buffered because level 2 reads plane k-1's kappa while the next
// plane's kappa is being written.  HBM moves the one-step bytes once per two steps
// (plus the ring, mostly L2 hits from the neighbouring tiles on the same XCD).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "scafes_common.hpp"

namespace scafes {

struct VarCoef {
    double tau, hh0, hh1, hh2;  // FP64 path (CPU operation order)
    float c0, c1, c2;           // 0.5*tau/h_p^2, FP32 paths
    double border_value;
    int border_mode;
};

// One explicit step of the conservative form at one node.  MODE 0: FP64 in
// the CPU path's order; 1: FP32; 2: FP32 + Kahan (cin -> *cout).
template <typename T, int MODE>
__device__ inline T var_update(const Geo& g, const VarCoef& cf, T uc, T uw, T ue, T us, T un, T ud, T uu, T kc,
                               T kw, T ke, T ks, T kn, T kd, T ku, T cin, T* cout)
{
    if constexpr (MODE == 0) {
        double v = uc;
        v += (cf.tau * (0.5 * (ke + kc) * (ue - uc) - 0.5 * (kc + kw) * (uc - uw))) / cf.hh0;
        if (g.dim > 1) v += (cf.tau * (0.5 * (kn + kc) * (un - uc) - 0.5 * (kc + ks) * (uc - us))) / cf.hh1;
        if (g.dim > 2) v += (cf.tau * (0.5 * (ku + kc) * (uu - uc) - 0.5 * (kc + kd) * (uc - ud))) / cf.hh2;
        *cout = T(0);
        return (T)v;
    } else {
        float d = cf.c0 * __builtin_fmaf(ke + kc, ue - uc, -((kc + kw) * (uc - uw)));
        if (g.dim > 1) d = __builtin_fmaf(cf.c1, __builtin_fmaf(kn + kc, un - uc, -((kc + ks) * (uc - us))), d);
        if (g.dim > 2) d = __builtin_fmaf(cf.c2, __builtin_fmaf(ku + kc, uu - uc, -((kc + kd) * (uc - ud))), d);
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

template <typename T, int MODE, int TX, int TY, int D>
__global__ __launch_bounds__(TX* TY) void heatvar_tb2_kernel(Geo g, const T* __restrict__ u, const T* __restrict__ cin,
                                                              const T* __restrict__ kap, T* __restrict__ uout,
                                                              T* __restrict__ cout, VarCoef cf, int zc, int tiles_x,
                                                              int tiles_y)
{
    static_assert(D >= 2 && D % 2 == 0, "prefetch depth must be even");
    constexpr int OX = TX - 2, OY = TY - 2;
    constexpr bool KH = (MODE == 2);
    __shared__ T s0[2][TY + 2][TX + 2];  // level-0 plane with ring
    __shared__ T sk[3][TY + 2][TX + 2];  // kappa plane with ring (read again one plane later)
    __shared__ T s1[2][TY][TX];          // level-1 plane
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per_plane = tiles_x * tiles_y;
    const int tz = lb / per_plane;
    const int r = lb - tz * per_plane;
    const int tyi = r / tiles_x;
    const int txi = r - tyi * tiles_x;
    const int i = txi * OX - 1 + tx, j = tyi * OY - 1 + ty;
    const int ic = min(max(i, 0), g.nx - 1), jc = min(max(j, 0), g.ny - 1);
    const bool in_xy = i >= 0 && i < g.nx && j >= 0 && j < g.ny;
    const bool out_thread = tx >= 1 && tx <= TX - 2 && ty >= 1 && ty <= TY - 2 && i < g.nx && j < g.ny;
    const int k0 = tz * zc, k1 = min(k0 + zc, g.nz);
    const int kstart = k0 - 1;         // first plane with a level-1 value
    const int kcen = k1 + 1;           // last plane whose centre is needed
    // halo duty for the level-0 ring
    const bool hx = (tx == 0) || (tx == TX - 1);
    const bool hy = (ty == 0) || (ty == TY - 1);
    const int hxi = (tx == 0) ? max(ic - 1, 0) : min(ic + 1, g.nx - 1);
    const int hyj = (ty == 0) ? max(jc - 1, 0) : min(jc + 1, g.ny - 1);
    const int sxh = (tx == 0) ? 0 : TX + 1;
    const int syh = (ty == 0) ? 0 : TY + 1;
    const long oc = at(g, ic, jc, 0), ohx = at(g, hxi, jc, 0), ohy = at(g, ic, hyj, 0);
    auto kcl = [&](int kk) { return min(max(kk, 0), g.nz - 1); };

    T qa[D], qb[D], qc[D], qax[D], qbx[D], qay[D], qby[D];
#pragma unroll
    for (int s = 0; s < D; ++s) qa[s] = qb[s] = qc[s] = qax[s] = qbx[s] = qay[s] = qby[s] = T(0);
    auto fetch = [&](const int slot, const int kk) {
        if (kk <= kcen) {
            const long o = g.e01 * kcl(kk);
            qa[slot] = u[oc + o];
            qb[slot] = kap[oc + o];
            if (kk <= k1) {
                if constexpr (KH) qc[slot] = cin[oc + o];
                if (hx) {
                    qax[slot] = u[ohx + o];
                    qbx[slot] = kap[ohx + o];
                }
                if (hy) {
                    qay[slot] = u[ohy + o];
                    qby[slot] = kap[ohy + o];
                }
            }
        }
    };
    T am = u[oc + g.e01 * kcl(kstart - 1)];
    T bm = kap[oc + g.e01 * kcl(kstart - 1)];
#pragma unroll
    for (int s = 0; s < D; ++s) fetch(s, kstart + s);

    T l1m2 = T(0), l1m1 = T(0), c1m1 = T(0);  // level-1 at k-2, k-1 and compensation at k-1
    T km2 = T(0), km1 = T(0);                  // kappa centre at k-2, k-1
    int kcur = 0, kprev = 2;                   // kappa LDS slots of planes k and k-1
    const T bv = (T)cf.border_value;
    for (int kb = kstart; kb <= k1; kb += D) {
#pragma unroll
        for (int s = 0; s < D; ++s) {
            const int k = kb + s;
            if (k > k1) break;  // uniform over the workgroup
            const int buf = s & 1;
            const T ac = qa[s], ap = qa[(s + 1) % D];
            const T bc = qb[s], bp = qb[(s + 1) % D];
            const T c0 = qc[s];
            s0[buf][ty + 1][tx + 1] = ac;
            sk[kcur][ty + 1][tx + 1] = bc;
            if (hx) {
                s0[buf][ty + 1][sxh] = qax[s];
                sk[kcur][ty + 1][sxh] = qbx[s];
            }
            if (hy) {
                s0[buf][syh][tx + 1] = qay[s];
                sk[kcur][syh][tx + 1] = qby[s];
            }
            fetch(s, k + D);
            __syncthreads();
            // ---- level 1 at (i, j, k)
            T l1, c1 = T(0);
            const bool in1 = in_xy && k >= 0 && k < g.nz;
            if (in1 && !is_border(g, i, j, k)) {
                l1 = var_update<T, MODE>(g, cf, ac, s0[buf][ty + 1][tx], s0[buf][ty + 1][tx + 2], s0[buf][ty][tx + 1],
                                         s0[buf][ty + 2][tx + 1], am, ap, bc, sk[kcur][ty + 1][tx],
                                         sk[kcur][ty + 1][tx + 2], sk[kcur][ty][tx + 1], sk[kcur][ty + 2][tx + 1], bm,
                                         bp, c0, &c1);
            } else if (cf.border_mode == SCAFES_BORDER_VALUE) {
                l1 = bv;
            } else {
                l1 = ac;
                c1 = c0;
            }
            s1[buf][ty][tx] = l1;
            // ---- level 2 at (i, j, k-1) from the level-1 plane of the previous iteration
            const int k2 = k - 1;
            if (k2 >= k0 && out_thread) {
                const long m = at(g, i, j, k2);
                T v, cv = T(0);
                if (!is_border(g, i, j, k2)) {
                    const int pb = buf ^ 1;
                    v = var_update<T, MODE>(g, cf, l1m1, s1[pb][ty][tx - 1], s1[pb][ty][tx + 1], s1[pb][ty - 1][tx],
                                            s1[pb][ty + 1][tx], l1m2, l1, km1, sk[kprev][ty + 1][tx],
                                            sk[kprev][ty + 1][tx + 2], sk[kprev][ty][tx + 1], sk[kprev][ty + 2][tx + 1],
                                            km2, bc, c1m1, &cv);
                } else if (cf.border_mode == SCAFES_BORDER_VALUE) {
                    v = bv;
                } else {
                    v = l1m1;
                    cv = c1m1;
                }
                uout[m] = v;
                if constexpr (KH) cout[m] = cv;
            }
            l1m2 = l1m1;
            l1m1 = l1;
            c1m1 = c1;
            km2 = km1;
            km1 = bc;
            am = ac;
            bm = bc;
            kprev = kcur;
            kcur = kcur == 2 ? 0 : kcur + 1;
        }
    }
}

struct Tb2Cfg {
    int tx, ty, zc, d;
};
static const Tb2Cfg kTb2Cfgs[] = {
    {64, 16, 64, 2}, {64, 16, 128, 2}, {64, 8, 64, 2}, {128, 8, 64, 2}, {64, 16, 64, 4}, {128, 4, 64, 2},
};
static constexpr int kNumTb2 = sizeof(kTb2Cfgs) / sizeof(kTb2Cfgs[0]);

template <typename T, int MODE, int TX, int TY, int D>
static int tb2_launch_t(Ctx* c, const Geo& g, const void* u, const void* cin, const void* kap, void* uo, void* co,
                        const VarCoef& cf, int zc)
{
    const int tiles_x = ceil_div(g.nx, TX - 2), tiles_y = ceil_div(g.ny, TY - 2), tiles_z = ceil_div(g.nz, zc);
    const long nb = (long)tiles_x * tiles_y * tiles_z;
    if (nb > 0x7fffffffL) return fail(SCAFES_ERR_ARG, "grid too large");
    heatvar_tb2_kernel<T, MODE, TX, TY, D><<<dim3((unsigned)nb), dim3(TX, TY), 0, c->stream>>>(
        g, (const T*)u, (const T*)cin, (const T*)kap, (T*)uo, (T*)co, cf, zc, tiles_x, tiles_y);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

template <typename T, int MODE>
static int tb2_dispatch(Ctx* c, const Geo& g, const void* u, const void* cin, const void* kap, void* uo, void* co,
                        const VarCoef& cf, int cfg)
{
    const Tb2Cfg t = kTb2Cfgs[cfg];
    switch (t.tx * 100000 + t.ty * 100 + t.d) {
    case 6401602: return tb2_launch_t<T, MODE, 64, 16, 2>(c, g, u, cin, kap, uo, co, cf, t.zc);
    case 6400802: return tb2_launch_t<T, MODE, 64, 8, 2>(c, g, u, cin, kap, uo, co, cf, t.zc);
    case 12800802: return tb2_launch_t<T, MODE, 128, 8, 2>(c, g, u, cin, kap, uo, co, cf, t.zc);
    case 6401604: return tb2_launch_t<T, MODE, 64, 16, 4>(c, g, u, cin, kap, uo, co, cf, t.zc);
    case 12800402: return tb2_launch_t<T, MODE, 128, 4, 2>(c, g, u, cin, kap, uo, co, cf, t.zc);
    default: return fail(SCAFES_ERR_ARG, "bad two-step tile configuration");
    }
}

int heatvar_tb2_num_cfgs() { return kNumTb2; }

void heatvar_tb2_cfg(int cfg, int* tx, int* ty, int* zc, int* d)
{
    const Tb2Cfg t = kTb2Cfgs[cfg];
    *tx = t.tx;
    *ty = t.ty;
    *zc = t.zc;
    *d = t.d;
}

// Two steps u -> uo (and Kahan compensation cin -> co).
int heatvar_tb2_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int mode, const void* u, void* uo,
                       const void* cin, void* co, const void* kap, double tau, int border_mode, double border_value,
                       int cfg)
{
    if (cfg < 0 || cfg >= kNumTb2) return fail(SCAFES_ERR_ARG, "bad two-step tile configuration");
    VarCoef cf;
    double hh[3];
    for (int p = 0; p < 3; ++p) hh[p] = p < g->dim ? g->h[p] * g->h[p] : 1.0;
    cf.tau = tau;
    cf.hh0 = hh[0];
    cf.hh1 = hh[1];
    cf.hh2 = hh[2];
    cf.c0 = (float)(0.5 * tau / hh[0]);
    cf.c1 = (float)(0.5 * tau / hh[1]);
    cf.c2 = (float)(0.5 * tau / hh[2]);
    cf.border_value = border_value;
    cf.border_mode = border_mode;
    if (mode == SCAFES_HEATVAR_F64) return tb2_dispatch<double, 0>(c, geo, u, nullptr, kap, uo, nullptr, cf, cfg);
    if (mode == SCAFES_HEATVAR_F32) return tb2_dispatch<float, 1>(c, geo, u, nullptr, kap, uo, nullptr, cf, cfg);
    return tb2_dispatch<float, 2>(c, geo, u, cin, kap, uo, co, cf, cfg);
}

}  // namespace scafes
