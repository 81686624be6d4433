// heatvar_adjk.hip -- K reverse steps of the variable-kappa adjoint per HBM
// sweep (lam^n -> lam^{n-K}, K = 2..kAdjMaxK) with the kappa-gradient terms
// of all K steps, for 3-D grids.  Per node the arithmetic is that of
// HeatVarAdjOp in stencil3d.hip, in the same order:
//   lam^{n-1} = the forward step's operator applied to lam^n (var_update
//               with the node's six face coefficients), 0 on border nodes;
//   grad     += 0.5 * sum_p th_p * sum_{nb = -+e_p in the grid}
//               (lam^n_c - lam^n_nb)(u^{n-1}_nb - u^{n-1}_c)   (every node),
// and the FP64 gradient of a node receives the step-n term, then the
// step-(n-1) term, ..., so the result is bit-identical to K one-step launches.
//
// A TX x TY thread tile walks a chunk of z planes.  In iteration k level 0
// (lam^n) and kappa are staged in LDS at plane k with a one-node ring;
// level L >= 1 (lam^{n-L}) is computed at plane k-(L-1) on the tile shrunk
// by L-1 rings, from level L-1's LDS plane of the previous iteration (x, y
// neighbours) and its register history (z neighbours).  The six face
// coefficients of a plane are formed once, by level 1, and kept in a
// register ring for levels 2..K.  The gradient term of step n-l at plane p
// needs lam^{n-l} and u^{n-1-l} with their six neighbours: it is formed in
// the iteration in which level l+1 is computed at plane p (level l's plane p
// is then in LDS, its z neighbours in registers), with u^{n-1-l} staged in
// LDS for that plane, and added to the node's running FP64 sum.  The sum
// starts from the stored gradient with step n's term and is stored with
// lam^{n-K}, on the (TX-2(K-1)) x (TY-2(K-1)) output tile.  One barrier per
// plane; LDS planes are double-buffered.
//
// Per K steps HBM moves lam r+w, kappa, u^{n-1}..u^{n-K} and the FP64
// gradient r+w once: (28 + 4K)/K bytes per step for FP32 fields (13.3 at
// K = 3, 11 at K = 4), against 32 for one-step launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "buffer_chunk.hpp"
#include "heatvar_ops.hpp"
#include "scafes_common.hpp"

namespace scafes {

namespace {

using namespace chunk;

struct AdjkStates {
    const void* u[kAdjMaxK];  // u[l] = u^{n-1-l}
};

template <typename T, int TX, int TY, int K, int D>
__global__ __launch_bounds__(TX* TY) void heatvar_adjk_kernel(Geo g, const T* __restrict__ lam,
                                                               T* __restrict__ lamo, const T* __restrict__ kap,
                                                               AdjkStates st, double* __restrict__ grad,
                                                               AdjVarCoef cf, int zc, int tiles_x, int tiles_y)
{
    static_assert(D >= 2 && D % 2 == 0, "prefetch depth must be even");
    static_assert(K >= 2 && K <= kAdjMaxK, "2 <= K <= kAdjMaxK");
    static_assert(TX % 64 == 0, "whole rows per wavefront: ty is wave-uniform");
    constexpr int MODE = sizeof(T) == 8 ? 0 : 1;
    constexpr int R = K - 1;  // rings the tile loses over the sweep
    constexpr int OX = TX - 2 * R, OY = TY - 2 * R;
    static_assert(OX >= 1 && OY >= 1, "tile too small for K levels");
    // Register rings: a value formed in iteration it sits at ring(it); ages
    // up to max(2, K-1) are read.  The plane loop is unrolled by U, a
    // multiple of the ring size and of D, so every index is a constant.
    constexpr int H = K > 3 ? K : 3;
    constexpr int U = H / gcd_c(H, D) * D;
    auto ring = [](int c) { return ((c % H) + H) % H; };
    // lam^n and kappa with ring; row TY + 2 takes the y-ring writes of the
    // rows off the ring (a sink that is never read)
    __shared__ T s0[2][TY + 3][TX + 2];
    __shared__ T sk[2][TY + 3][TX + 2];
    __shared__ T sl[K - 1][2][TY][TX];  // levels 1..K-1
    __shared__ T su[K][2][TY][TX];      // u^{n-1-l} for the step-(n-l) term
    const int tx = threadIdx.x;
    const int ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per_plane = tiles_x * tiles_y;
    const int tz = lb / per_plane;
    const int r = lb - tz * per_plane;
    const int tyi = r / tiles_x;
    const int txi = r - tyi * tiles_x;
    const int i = txi * OX - R + tx, j = tyi * OY - R + ty;
    const int ic = min(max(i, 0), g.nx - 1), jc = min(max(j, 0), g.ny - 1);
    // lane-dependent (x) and wave-uniform (y) parts of the node predicates
    const bool inner_i = i >= g.nl0 && i <= g.nx - 1 - g.nl0;
    const bool inner_j = j >= g.nl1 && j <= g.ny - 1 - g.nl1;
    const bool col_out = tx >= R && tx <= TX - 1 - R && i < g.nx;
    const bool row_out = ty >= R && ty <= TY - 1 - R && j < g.ny;
    const bool xlo = i > 0, xhi = i < g.nx - 1, ylo = j > 0, yhi = j < g.ny - 1;
    const int k0 = tz * zc, k1 = min(k0 + zc, g.nz);
    const int kstart = k0 - R, kend = k1 + R - 1;
    const int klo = g.nl2, khi = g.nz - 1 - g.nl2;
    const bool hx = (tx == 0) || (tx == TX - 1);
    const bool hy = (ty == 0) || (ty == TY - 1);
    const int hxi = (tx == 0) ? max(ic - 1, 0) : min(ic + 1, g.nx - 1);
    const int hyj = (ty == 0) ? max(jc - 1, 0) : min(jc + 1, g.ny - 1);
    // x ring: a lane off the ring reloads and rewrites its own centre value
    const int sxc = hx ? ((tx == 0) ? 0 : TX + 1) : tx + 1;
    const int syc = hy ? ((ty == 0) ? 0 : TY + 1) : TY + 2;
    // LDS neighbours of levels >= 2 and of the staged states, clamped into
    // the tile: only lanes whose result is discarded read a clamped position
    const int txm = max(tx - 1, 0), txp = min(tx + 1, TX - 1);
    const int tym = max(ty - 1, 0), typ = min(ty + 1, TY - 1);
    constexpr uint32_t B = sizeof(T);
    const uint32_t oc = (uint32_t)(g.base + ic + g.e0 * jc) * B;
    const uint32_t ohx = hx ? (uint32_t)(g.base + hxi + g.e0 * jc) * B : oc;
    const uint32_t ohy = hy ? (uint32_t)(g.base + ic + g.e0 * hyj) * B : kSkip;
    const bool out = col_out && row_out;
    const uint32_t oo = out ? oc : kSkip;                                         // output nodes
    const uint32_t og = out ? (uint32_t)(g.base + ic + g.e0 * jc) * 8u : kSkip;  // FP64 gradient
    // the states are read by the output rows and their y neighbours
    const uint32_t ous = (ty >= R - 1 && ty <= TY - R) ? oc : kSkip;
    const uint32_t pb = (uint32_t)(g.e01 * B), pg = (uint32_t)(g.e01 * 8);
    // every plane this workgroup reads or writes lies in [kbase, nz-1]
    const int kbase = max(kstart - 1, 0);
    auto pl = [&](int kk) { return (uint32_t)(min(max(kk, kbase), g.nz - 1) - kbase); };
    const Rsrc rl = chunk_rsrc((const char*)lam + (size_t)kbase * pb);
    const Rsrc rk = chunk_rsrc((const char*)kap + (size_t)kbase * pb);
    const Rsrc rlo = chunk_rsrc((char*)lamo + (size_t)kbase * pb);
    const Rsrc rg = chunk_rsrc((char*)grad + (size_t)kbase * pg);
    Rsrc ru[K];
#pragma unroll
    for (int l = 0; l < K; ++l) ru[l] = chunk_rsrc((const char*)st.u[l] + (size_t)kbase * pb);

    T qa[D], qax[D], qay[D], qk[D], qkx[D], qky[D], qu[K][D];
    double qg[D];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        qa[s] = qax[s] = qay[s] = qk[s] = qkx[s] = qky[s] = T(0);
#pragma unroll
        for (int l = 0; l < K; ++l) qu[l][s] = T(0);
        qg[s] = 0.0;
    }
    // slot <- plane kk's lam^n and kappa (centre and ring), its stored
    // gradient, and u^{n-1-l} at plane kk - l.  Unconditional: planes past
    // the chunk are clamped into the grid and their values never used;
    // values no output node needs are skipped through the offset.
    auto fetch = [&](const int slot, const int kk) {
        const uint32_t p = pl(kk) * pb;
        qa[slot] = cload<T>(rl, oc, p);
        qax[slot] = cload<T>(rl, ohx, p);
        qay[slot] = cload<T>(rl, ohy, p);
        qk[slot] = cload<T>(rk, oc, p);
        qkx[slot] = cload<T>(rk, ohx, p);
        qky[slot] = cload<T>(rk, ohy, p);
#pragma unroll
        for (int l = 0; l < K; ++l) {
            const int pp = kk - l;
            qu[l][slot] = cload<T>(ru[l], (pp >= k0 - 1 && pp <= k1) ? ous : kSkip, pl(pp) * pb);
        }
        qg[slot] = cload<double>(rg, (kk >= k0 && kk < k1) ? og : kSkip, pl(kk) * pg);
    };
#pragma unroll
    for (int s = 0; s < D; ++s) fetch(s, kstart + s);

    T hist[K][H];         // hist[l][ring(it)]: level l at the plane it computed in iteration it
    Faces<T> fr[H];       // fr[ring(it)]: faces of plane kstart + it
    double sum[H];        // sum[ring(it)]: running gradient of plane kstart + it
    T um[K];              // u^{n-1-l} one plane below the plane its term is formed at
#pragma unroll
    for (int a = 0; a < H; ++a) {
        fr[a] = Faces<T>{};
        sum[a] = 0.0;
#pragma unroll
        for (int l = 0; l < K; ++l) hist[l][a] = T(0);
    }
    hist[0][ring(-1)] = cload<T>(rl, oc, pl(kstart - 1) * pb);
    fr[ring(-1)].u = face<T, MODE>(qk[0], cload<T>(rk, oc, pl(kstart - 1) * pb));
#pragma unroll
    for (int l = 0; l < K; ++l) {
        const int pp = kstart - l - 1;
        um[l] = cload<T>(ru[l], (pp >= k0 - 1 && pp <= k1) ? ous : kSkip, pl(pp) * pb);
    }
    for (int kb = kstart; kb <= kend; kb += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = kb + u;
            // uniform over the workgroup; nothing follows the plane loop, so
            // the exit leaves the register rings dead (no copies on this path)
            if (k > kend) return;
            const int s = u % D, buf = u & 1, pbf = buf ^ 1;
            const T ac = qa[s], ap = qa[(s + 1) % D], am = hist[0][ring(u - 1)];
            const T bc = qk[s], bp = qk[(s + 1) % D];
            T uc[K], up[K];
#pragma unroll
            for (int l = 0; l < K; ++l) {
                uc[l] = qu[l][s];
                up[l] = qu[l][(s + 1) % D];
            }
            const double gl = qg[s];
            hist[0][ring(u)] = ac;
            s0[buf][ty + 1][tx + 1] = ac;
            s0[buf][ty + 1][sxc] = qax[s];
            s0[buf][syc][tx + 1] = qay[s];
            sk[buf][ty + 1][tx + 1] = bc;
            sk[buf][ty + 1][sxc] = qkx[s];
            sk[buf][syc][tx + 1] = qky[s];
#pragma unroll
            for (int l = 0; l < K; ++l) su[l][buf][ty][tx] = uc[l];
            fetch(s, k + D);
            __syncthreads();
            // ---- level 1: lam^{n-1} at plane k, every thread; the faces of plane k
            const T lw = s0[buf][ty + 1][tx], le = s0[buf][ty + 1][tx + 2];
            const T ls = s0[buf][ty][tx + 1], ln = s0[buf][ty + 2][tx + 1];
            Faces<T> f;
            f.e = face<T, MODE>(sk[buf][ty + 1][tx + 2], bc);
            f.w = face<T, MODE>(bc, sk[buf][ty + 1][tx]);
            f.n = face<T, MODE>(sk[buf][ty + 2][tx + 1], bc);
            f.s = face<T, MODE>(bc, sk[buf][ty][tx + 1]);
            f.u = face<T, MODE>(bp, bc);
            f.d = fr[ring(u - 1)].u;
            fr[ring(u)] = f;
            {
                T dummy;
                T l1 = var_update<T, MODE>(cf.v, ac, lw, le, ls, ln, am, ap, f, T(0), &dummy);
                const bool in1 = inner_i && inner_j && k >= klo && k <= khi;
                l1 = in1 ? l1 : T(0);
                hist[1][ring(u)] = l1;
                sl[0][buf][ty][tx] = l1;
            }
            // ---- step-n term at plane k (output rows and planes)
            if (row_out && k >= k0 && k < k1) {  // wave-uniform
                const double t = adj_term3<T>(cf, xlo, xhi, ylo, yhi, k > 0, k < g.nz - 1, ac, lw, le, ls, ln, am,
                                              ap, uc[0], su[0][buf][ty][txm], su[0][buf][ty][txp],
                                              su[0][buf][tym][tx], su[0][buf][typ][tx], um[0], up[0]);
                sum[ring(u)] = gl + t;
            }
            // ---- levels 2..K: lam^{n-L} at plane k-(L-1), and the step-(n-L+1) term there
#pragma unroll
            for (int L = 2; L <= K; ++L) {
                const int p = k - (L - 1);
                const bool inz = p >= klo && p <= khi;
                const T lc = hist[L - 1][ring(u - 1)], lm = hist[L - 1][ring(u - 2)], lp = hist[L - 1][ring(u)];
                T v = T(0);
                if (ty >= L - 1 && ty <= TY - L && inner_j && inz) {  // wave-uniform
                    T dummy;
                    v = var_update<T, MODE>(cf.v, lc, sl[L - 2][pbf][ty][txm], sl[L - 2][pbf][ty][txp],
                                            sl[L - 2][pbf][tym][tx], sl[L - 2][pbf][typ][tx], lm, lp,
                                            fr[ring(u - (L - 1))], T(0), &dummy);
                    v = (tx >= L - 1 && tx <= TX - L && inner_i) ? v : T(0);
                }
                if (row_out && p >= k0 && p < k1) {  // wave-uniform
                    // level L-1 (lam^{n-L+1}) at plane p with u^{n-L} = state L-1
                    const int l = L - 1;
                    const double t =
                        adj_term3<T>(cf, xlo, xhi, ylo, yhi, p > 0, p < g.nz - 1, lc, sl[l - 1][pbf][ty][txm],
                                     sl[l - 1][pbf][ty][txp], sl[l - 1][pbf][tym][tx], sl[l - 1][pbf][typ][tx], lm,
                                     lp, uc[l], su[l][buf][ty][txm], su[l][buf][ty][txp], su[l][buf][tym][tx],
                                     su[l][buf][typ][tx], um[l], up[l]);
                    sum[ring(u - l)] = sum[ring(u - l)] + t;
                }
                if (L < K) {
                    hist[L][ring(u)] = v;
                    sl[L - 1][buf][ty][tx] = v;
                } else if (row_out && p >= k0) {  // wave-uniform; p < k1 always
                    cstore<T>(v, rlo, oo, pl(p) * pb);
                    cstore<double>(sum[ring(u - R)], rg, og, pl(p) * pg);
                }
            }
#pragma unroll
            for (int l = 0; l < K; ++l) um[l] = uc[l];
        }
    }
}

// Tile: 64 x 16 threads, prefetch depth 2, 64-plane z chunks for FP32 (the
// 20-step gradient at 512^3 with every state stored: 18.9-19.4 ms with 64-
// and 128-plane chunks, 19.5-19.9 with 48 and 96, 19.6-19.9 with 32, 20.9-21.4
// with 16), 32 for FP64.  Measured on
// an MI355X (512^3 FP32 + Kahan, 20 steps store-all, whole gradient in ms):
// K <= 4 64x16 24.0, 64x12 27.9; K <= 3 64x12 24.6, 64x16 25.2; K = 2
// 64x8 24.5, 64x16 24.1; the two-step kernel of heatvar_tb2.hip 25.0.
// Those figures were taken with FP64 gradient terms.  With FP32 terms
// (adj_term3) the K = 4 sweep takes 2.14 ms at 512^3 (2.01 with 64-plane
// chunks) and still needs 123 VGPRs: one workgroup per CU, bound by vector
// issue (0.60 of its rate), not by HBM (0.41 of 8 TB/s).  Two 12-wave workgroups per CU would need
// at most 80 VGPRs.
// FP64 fields: 64 x 8 (a 1024-thread workgroup caps a thread at 128 VGPRs,
// which the FP64 rings exceed).
// FP32, K = 3: 64 x 12 (64 x 16 caps a thread at 128 VGPRs, which the
// three-step rings exceed by a 16-byte spill).
template <typename T, int K>
struct AdjkTile {
    static constexpr int TX = 64, TY = sizeof(T) == 8 ? 8 : (K == 3 ? 12 : 16), D = 2, ZC = sizeof(T) == 8 ? 32 : 64;
};

template <typename T, int K>
int adjk_launch_t(Ctx* c, const Geo& g, const void* lam, void* lamo, const void* kap, const AdjkStates& st,
                  double* grad, const AdjVarCoef& cf)
{
    typedef AdjkTile<T, K> A;
    constexpr int R = K - 1;
    const int tiles_x = ceil_div(g.nx, A::TX - 2 * R), tiles_y = ceil_div(g.ny, A::TY - 2 * R);
    const long nb = (long)tiles_x * tiles_y * ceil_div(g.nz, A::ZC);
    if (nb > 0x7fffffffL) return fail(SCAFES_ERR_ARG, "grid too large");
    heatvar_adjk_kernel<T, A::TX, A::TY, K, A::D><<<dim3((unsigned)nb), dim3(A::TX, A::TY), 0, c->stream>>>(
        g, (const T*)lam, (T*)lamo, (const T*)kap, st, grad, cf, A::ZC, tiles_x, tiles_y);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

template <typename T>
int adjk_launch_k(Ctx* c, const Geo& g, int K, const void* lam, void* lamo, const void* kap, const AdjkStates& st,
                  double* grad, const AdjVarCoef& cf)
{
    switch (K) {
    case 2: return adjk_launch_t<T, 2>(c, g, lam, lamo, kap, st, grad, cf);
    case 3: return adjk_launch_t<T, 3>(c, g, lam, lamo, kap, st, grad, cf);
    case 4: return adjk_launch_t<T, 4>(c, g, lam, lamo, kap, st, grad, cf);
    default: return fail(SCAFES_ERR_ARG, "multi-step var-kappa adjoint: K must be 2..4");
    }
}

}  // namespace

bool heatvar_adjk_supported(const Geo& geo)
{
    // 3-D only; a workgroup's chunk of the FP64 gradient (at most the z chunk
    // + 2K + the prefetch depth + 2 planes) plus the in-plane offsets must
    // stay below kRecords bytes
    constexpr long kPlanes = 64 + 2 * kAdjMaxK + 2 + 2;
    return geo.dim == 3 && (kPlanes * geo.e01 + geo.base + geo.e01) * 8L < (long)kRecords;
}

int heatvar_adjk_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, int K, const void* lam,
                        void* lamo, const void* kap, const void* const* u, double* grad, double tau)
{
    if (!heatvar_adjk_supported(geo))
        return fail(SCAFES_ERR_UNSUPPORTED, "multi-step var-kappa adjoint needs a 3-D grid");
    if (K < 2 || K > kAdjMaxK) return fail(SCAFES_ERR_ARG, "multi-step var-kappa adjoint: K must be 2..4");
    AdjkStates st{};
    for (int l = 0; l < K; ++l) {
        if (!u[l]) return fail(SCAFES_ERR_ARG, "multi-step var-kappa adjoint: missing state");
        st.u[l] = u[l];
    }
    AdjVarCoef cf{};
    double hh[3];
    for (int p = 0; p < 3; ++p) hh[p] = g->h[p] * g->h[p];
    cf.v.tau = tau;
    cf.v.hh0 = hh[0];
    cf.v.hh1 = hh[1];
    cf.v.hh2 = hh[2];
    cf.v.c0 = (float)(0.5 * tau / hh[0]);
    cf.v.c1 = (float)(0.5 * tau / hh[1]);
    cf.v.c2 = (float)(0.5 * tau / hh[2]);
    cf.v.border_value = 0.0;
    cf.v.border_mode = SCAFES_BORDER_VALUE;
    cf.th0 = tau / hh[0];
    cf.th1 = tau / hh[1];
    cf.th2 = tau / hh[2];
    cf.f0 = (float)(0.5 * cf.th0);
    cf.f1 = (float)(0.5 * cf.th1);
    cf.f2 = (float)(0.5 * cf.th2);
    if (dtype == SCAFES_F64) return adjk_launch_k<double>(c, geo, K, lam, lamo, kap, st, grad, cf);
    return adjk_launch_k<float>(c, geo, K, lam, lamo, kap, st, grad, cf);
}

}  // namespace scafes
