// pennes_adjk.hip -- K reverse-time Pennes adjoint steps per HBM sweep
// (lam^n -> lam^{n-K}, K = 2..kAdjMaxK) with the gradient contributions of
// all K steps, for 3-D grids.  Per node the arithmetic is that of
// PennesAdjOp in stencil3d.hip (the one-step adjoint), in the same order,
// and the FP64 gradient receives the step-n term, then the step-(n-1) term,
// ..., so the result is bit-identical to K one-step launches.
//
//   lam^{n-1}_i = lam^n_i + sum_p ctd*((lam_W+lam_E) - 2 lam_i)/h_p^2 - (ctp*w_i)*lam_i
//   grad_i     += lam^n_i * (ctp*(T_a - T^{n-1}_i))        (inner nodes)
//   lam^{n-1} = 0 on border nodes.
// Under the Tumor border every node is updated with the transposed face
// weights of the one-step operator (adj_update_tumor below) and the
// gradient covers every node.
//
// A TX x TY thread tile walks a chunk of z planes.  In iteration k, level 0
// (lam^n) is staged in LDS at plane k with a one-node ring; level L >= 1
// (lam^{n-L}) is computed at plane k-(L-1) from the LDS plane level L-1 wrote
// one iteration earlier (x, y neighbours) and from level L-1's register
// history (z neighbours).  Each level loses one ring of the tile, so level K
// is valid on the (TX-2(K-1)) x (TY-2(K-1)) output tile.  One barrier per
// plane (double-buffered LDS planes).  The gradient terms of an output node
// need lam^{n-L} at that node: level L's value, held in its register history
// until the output plane is reached.  The states T^{n-1..n-K} and the
// gradient are fetched directly for the output plane, D planes ahead.
//
// Per K steps HBM moves lam r+w, w, T^{n-1}..T^{n-K} and the FP64 gradient
// r+w once: (28 + 4K)/K bytes per step for FP32 fields (18 at K = 2, 11 at
// K = 4), against 32 for one-step launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "buffer_chunk.hpp"
#include "scafes_common.hpp"

namespace scafes {

namespace {

using namespace chunk;

struct AdjCoef {
    double ctd, ctp, ta, hh0, hh1, hh2;  // FP64 path
    float fd0, fd1, fd2, fp;             // FP32 path: ctd/h_p^2, ctp
    double a2h2, hc;                     // Tumor border: (2 h_z)/K and H (Cauchy face z = nz-1)
    float fa2h2, fhc;                    // the same rounded to FP32
};

// T^{n-1}, ..., T^{n-K} (and their lo parts for compensated states)
struct AdjStates {
    const void* t[kAdjMaxK];
    const void* lo[kAdjMaxK];
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

// One axis of the transposed Tumor step at a node with index idx of n
// (PennesAdjOp::tumor_adj's order): the neighbour row at idx-1 weighs this
// node twice when it lies on the face, the one at idx+1 twice when it has
// index n-1, neighbours beyond the grid are selected away (their values are
// never used); on the Cauchy face the diagonal gets -cau = -(2 h/K) H.
template <typename T>
__device__ inline T tumor_t(int idx, int n, bool cauchy, T cau, T lw, T le, T l)
{
    lw = idx >= 1 ? lw : T(0);
    le = idx <= n - 2 ? le : T(0);
    if (idx == 1) lw = lw + lw;
    if (idx == n - 2) le = le + le;
    if (cauchy && idx == n - 1) return ((lw + le) - cau * l) - (T)2 * l;
    return (lw + le) - (T)2 * l;
}

// One adjoint step of the Tumor-border Pennes update at node (i, j, k) of a
// 3-D grid (every node is updated; the Cauchy face is z = nz-1).
template <typename T>
__device__ inline T adj_update_tumor(const AdjCoef& cf, const Geo& g, int i, int j, int k, T c, T w_, T e, T s,
                                     T n, T d, T u, T wf)
{
    const T t0 = tumor_t<T>(i, g.nx, false, T(0), w_, e, c);
    const T t1 = tumor_t<T>(j, g.ny, false, T(0), s, n, c);
    if constexpr (sizeof(T) == 8) {
        const T t2 = tumor_t<T>(k, g.nz, true, cf.a2h2 * cf.hc, d, u, c);
        double v = c;
        v += (cf.ctd * t0) / cf.hh0;
        v += (cf.ctd * t1) / cf.hh1;
        v += (cf.ctd * t2) / cf.hh2;
        v -= (cf.ctp * wf) * c;
        return v;
    } else {
        const T t2 = tumor_t<T>(k, g.nz, true, cf.fa2h2 * cf.fhc, d, u, c);
        float dd = cf.fd0 * t0;
        dd = __builtin_fmaf(cf.fd1, t1, dd);
        dd = __builtin_fmaf(cf.fd2, t2, dd);
        dd = __builtin_fmaf(-cf.fp * wf, c, dd);
        return c + dd;
    }
}

// W: waves per SIMD the register budget must allow.  Two workgroups of 12
// waves per CU need at most 80 VGPRs (W = 6); the FP64 and compensated
// variants would spill under that cap and run one workgroup per CU (W = 1).
// TU: the Tumor border (every node of the grid is updated, with the
// transposed face weights of adj_update_tumor; the gradient covers every
// node) instead of border nodes kept.
template <typename T, bool LO, int TX, int TY, int K, int D, bool TU = false,
          int W = (sizeof(T) == 4 && !LO) ? 6 : 1>
__global__ __launch_bounds__(TX* TY) __attribute__((amdgpu_waves_per_eu(W))) void pennes_adjk_kernel(Geo g, const T* __restrict__ lam, T* __restrict__ lamo,
                                                              const T* __restrict__ wf, AdjStates st,
                                                              double* __restrict__ grad, AdjCoef cf, int zc,
                                                              int tiles_x, int tiles_y)
{
    static_assert(D >= 2 && D % 2 == 0, "prefetch depth must be even");
    static_assert(K >= 2 && K <= kAdjMaxK, "2 <= K <= kAdjMaxK");
    static_assert(TX == 64, "one tile row per wavefront: ty is wave-uniform");
    constexpr int R = K - 1;  // rings the tile loses over the sweep
    constexpr int OX = TX - 2 * R, OY = TY - 2 * R;
    static_assert(OX >= 1 && OY >= 1, "tile too small for K levels");
    // Register rings: the value level l computed a iterations ago sits at
    // ring(it - a), ring size H (ages up to max(2, K-1) are read).  The plane
    // loop is unrolled by U, a multiple of H and D, so every ring and
    // prefetch-slot index is a compile-time constant: no register moves.
    constexpr int H = K > 3 ? K : 3;
    constexpr int U = H / gcd_c(H, D) * D;
    auto ring = [](int c) { return ((c % H) + H) % H; };
    __shared__ T s0[2][TY + 2][TX + 2];  // lam^n plane with ring
    __shared__ T sl[K - 1][2][TY][TX];   // levels 1..K-1
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
    // inner: updated by the step (every node of the grid under TU)
    const bool inner_i = TU ? (i >= 0 && i <= g.nx - 1) : (i >= g.nl0 && i <= g.nx - 1 - g.nl0);
    const bool inner_j = TU ? (j >= 0 && j <= g.ny - 1) : (j >= g.nl1 && j <= g.ny - 1 - g.nl1);
    const bool col_out = tx >= R && tx <= TX - 1 - R && i < g.nx;
    const bool row_out = ty >= R && ty <= TY - 1 - R && j < g.ny;
    const int k0 = tz * zc, k1 = min(k0 + zc, g.nz);
    const int kstart = k0 - R, kend = k1 + R - 1;
    const int klo = TU ? 0 : g.nl2, khi = TU ? g.nz - 1 : g.nz - 1 - g.nl2;
    const bool hx = (tx == 0) || (tx == TX - 1);
    const bool hy = (ty == 0) || (ty == TY - 1);
    const int hxi = (tx == 0) ? max(ic - 1, 0) : min(ic + 1, g.nx - 1);
    const int hyj = (ty == 0) ? max(jc - 1, 0) : min(jc + 1, g.ny - 1);
    // x ring: a lane off the ring reloads and rewrites its own centre value
    const int sxc = hx ? ((tx == 0) ? 0 : TX + 1) : tx + 1;
    const int syh = (ty == 0) ? 0 : TY + 1;
    // LDS neighbours of levels >= 2, clamped into the tile: only lanes whose
    // result is discarded read a clamped position
    const int txm = max(tx - 1, 0), txp = min(tx + 1, TX - 1);
    const int tym = max(ty - 1, 0), typ = min(ty + 1, TY - 1);
    constexpr uint32_t B = sizeof(T);
    const uint32_t oc = (uint32_t)(g.base + ic + g.e0 * jc) * B;
    const uint32_t ohx = hx ? (uint32_t)(g.base + hxi + g.e0 * jc) * B : oc;
    const uint32_t ohy = hy ? (uint32_t)(g.base + ic + g.e0 * hyj) * B : kSkip;
    const bool out = col_out && row_out;
    const uint32_t oo = out ? oc : kSkip;                                     // output nodes
    const uint32_t og = out ? (uint32_t)(g.base + ic + g.e0 * jc) * 8u : kSkip;  // FP64 gradient
    const uint32_t ogs = inner_i ? og : kSkip;  // the gradient is written at inner nodes
    const uint32_t pb = (uint32_t)(g.e01 * B), pg = (uint32_t)(g.e01 * 8);
    // every plane this workgroup reads or writes lies in [kbase, nz-1]
    const int kbase = max(kstart - 1, 0);
    auto pl = [&](int kk) { return (uint32_t)(min(max(kk, kbase), g.nz - 1) - kbase); };
    const Rsrc rl = chunk_rsrc((const char*)lam + (size_t)kbase * pb);
    const Rsrc rw = chunk_rsrc((const char*)wf + (size_t)kbase * pb);
    const Rsrc rlo = chunk_rsrc((char*)lamo + (size_t)kbase * pb);
    const Rsrc rg = chunk_rsrc((char*)grad + (size_t)kbase * pg);
    Rsrc rt[K], rtl[K];
#pragma unroll
    for (int l = 0; l < K; ++l) {
        rt[l] = chunk_rsrc((const char*)st.t[l] + (size_t)kbase * pb);
        rtl[l] = chunk_rsrc(LO ? (const char*)st.lo[l] + (size_t)kbase * pb : nullptr);
    }

    T qa[D], qax[D], qay[D], qw[D], qt[K][D], qtl[K][D];
    double qg[D];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        qa[s] = qax[s] = qay[s] = qw[s] = T(0);
#pragma unroll
        for (int l = 0; l < K; ++l) qt[l][s] = qtl[l][s] = T(0);
        qg[s] = 0.0;
    }
    // slot <- level-0 input at plane kk (centre, halo, w) and the output
    // plane kk - R's states and gradient.  Unconditional: planes past the
    // chunk are clamped into the grid and their values never used; output
    // planes outside [k0, k1) are skipped through the offset.
    auto fetch = [&](const int slot, const int kk) {
        const uint32_t p = pl(kk) * pb;
        qa[slot] = cload<T>(rl, oc, p);
        qw[slot] = cload<T>(rw, oc, p);
        qax[slot] = cload<T>(rl, ohx, p);
        qay[slot] = cload<T>(rl, ohy, p);
        const int po = kk - R;
        const bool need = po >= k0 && po < k1;
        const uint32_t os = need ? oo : kSkip, osg = need ? og : kSkip;
        const uint32_t q = pl(po);
#pragma unroll
        for (int l = 0; l < K; ++l) {
            qt[l][slot] = cload<T>(rt[l], os, q * pb);
            if constexpr (LO) qtl[l][slot] = cload<T>(rtl[l], os, q * pb);
        }
        qg[slot] = cload<double>(rg, osg, q * pg);
    };
#pragma unroll
    for (int s = 0; s < D; ++s) fetch(s, kstart + s);

    T hist[K][H];  // hist[l][ring(it)]: level l at the plane it computed in iteration it
    T wh[H];       // wh[ring(it)]: w at plane kstart + it
#pragma unroll
    for (int a = 0; a < H; ++a) {
        wh[a] = T(0);
#pragma unroll
        for (int l = 0; l < K; ++l) hist[l][a] = T(0);
    }
    hist[0][ring(-1)] = cload<T>(rl, oc, pl(kstart - 1) * pb);
    for (int kb = kstart; kb <= kend; kb += U) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = kb + u;
            // uniform over the workgroup; nothing follows the plane loop, so
            // the exit leaves the register rings dead (no copies on this path)
            if (k > kend) return;
            const int s = u % D, buf = u & 1;
            const T ac = qa[s], ap = qa[(s + 1) % D], am = hist[0][ring(u - 1)];
            hist[0][ring(u)] = ac;
            wh[ring(u)] = qw[s];
            T tl[K], tll[K];
#pragma unroll
            for (int l = 0; l < K; ++l) {
                tl[l] = qt[l][s];
                tll[l] = qtl[l][s];
            }
            const double gl = qg[s];
            s0[buf][ty + 1][tx + 1] = ac;
            s0[buf][ty + 1][sxc] = qax[s];
            if (hy) s0[buf][syh][tx + 1] = qay[s];
            fetch(s, k + D);
            __syncthreads();
            // ---- level 1: lam^{n-1} at plane k, every thread
            {
                const bool in1 = inner_i && inner_j && k >= klo && k <= khi;
                T l1;
                if constexpr (TU)
                    l1 = adj_update_tumor<T>(cf, g, i, j, k, ac, s0[buf][ty + 1][tx], s0[buf][ty + 1][tx + 2],
                                             s0[buf][ty][tx + 1], s0[buf][ty + 2][tx + 1], am, ap, wh[ring(u)]);
                else
                    l1 = adj_update<T>(cf, ac, s0[buf][ty + 1][tx], s0[buf][ty + 1][tx + 2], s0[buf][ty][tx + 1],
                                       s0[buf][ty + 2][tx + 1], am, ap, wh[ring(u)]);
                l1 = in1 ? l1 : T(0);
                hist[1][ring(u)] = l1;
                sl[0][buf][ty][tx] = l1;
            }
            // ---- levels 2..K: lam^{n-L} at plane k-(L-1) on the tile shrunk by L-1 rings
#pragma unroll
            for (int L = 2; L <= K; ++L) {
                const int p = k - (L - 1);
                const bool inz = p >= klo && p <= khi;
                T v = T(0);
                if (ty >= L - 1 && ty <= TY - L && inner_j && inz) {  // wave-uniform
                    const int pbf = buf ^ 1;
                    if constexpr (TU)
                        v = adj_update_tumor<T>(cf, g, i, j, p, hist[L - 1][ring(u - 1)], sl[L - 2][pbf][ty][txm],
                                                sl[L - 2][pbf][ty][txp], sl[L - 2][pbf][tym][tx],
                                                sl[L - 2][pbf][typ][tx], hist[L - 1][ring(u - 2)],
                                                hist[L - 1][ring(u)], wh[ring(u - (L - 1))]);
                    else
                        v = adj_update<T>(cf, hist[L - 1][ring(u - 1)], sl[L - 2][pbf][ty][txm],
                                          sl[L - 2][pbf][ty][txp], sl[L - 2][pbf][tym][tx], sl[L - 2][pbf][typ][tx],
                                          hist[L - 1][ring(u - 2)], hist[L - 1][ring(u)], wh[ring(u - (L - 1))]);
                    v = (tx >= L - 1 && tx <= TX - L && inner_i) ? v : T(0);
                }
                if (L < K) {
                    hist[L][ring(u)] = v;
                    sl[L - 1][buf][ty][tx] = v;
                } else if (row_out && p >= k0) {  // wave-uniform
                    // output plane: lam^{n-K} and the K gradient terms in step order
                    if (inner_j && inz) {
                        double gv = gl;
#pragma unroll
                        for (int l = 0; l < K; ++l) {
                            // lam^{n-l} at plane p: level l computed it K-1 (l = 0) or K-l iterations ago
                            const T lv = hist[l][ring(u - (l == 0 ? K - 1 : K - l))];
                            double tp = (double)tl[l];
                            if constexpr (LO) tp += (double)tll[l];
                            gv += (double)lv * (cf.ctp * (cf.ta - tp));
                        }
                        cstore<double>(gv, rg, ogs, pl(p) * pg);
                    }
                    cstore<T>(v, rlo, oo, pl(p) * pb);
                }
            }
        }
    }
}

// Tile: 64 x 12 threads, 32 z planes per workgroup, prefetch depth 2.
// Measured on an MI355X (Pennes 256^3 FP32, us per adjoint step, best z
// chunk of 8..64 planes): K = 4 64x12 64.3, 64x16 72.3; K = 3 64x8 70.4,
// 64x16 74.1; K = 2 64x16 72.6, 64x8 74.9.  The sweeps are not HBM-bound
// (2.9 TB/s algorithmic at K = 4): the SQ counters put the scalar unit at
// 0.68 of its issue rate against 0.47 for the vector units.  One resource
// per array and workgroup (the plane in the scalar offset) took K = 4 to
// 241 us per launch; wave-uniform rows, skipped lanes through out-of-range
// offsets instead of branches and register rings instead of shifts take it
// to 207 us (K = 3 157, K = 2 139), scalar instructions per wave 4493 ->
// 2624.  With the branch-free kernel 64x12 with prefetch depth 2 is still
// the best tile: 64x16 is 8 % slower, 64x16 limited to 64 VGPRs (8 waves
// per SIMD, spills) 44 %, prefetch depth 4 16 %, 64x8 24 %.  More steps per
// sweep do not pay: K = 5 on 64x16 (9.6 B per step, 118 VGPRs, one
// workgroup per CU) gave the 2000-step gradient with 62 snapshots in 302.6
// ms against 300.4 for K = 4 on 64x12, and K = 6 (8.7 B per step) 335 ms.
constexpr int kKTX = 64, kKTY = 12, kKZC = 32, kKD = 2;

template <typename T, bool LO, int K, bool TU>
int adjk_launch_t(Ctx* c, const Geo& g, const void* lam, void* lamo, const void* w, const AdjStates& st,
                  double* grad, const AdjCoef& cf)
{
    constexpr int R = K - 1;
    const int tiles_x = ceil_div(g.nx, kKTX - 2 * R), tiles_y = ceil_div(g.ny, kKTY - 2 * R);
    const int tiles_z = ceil_div(g.nz, kKZC);
    const long nb = (long)tiles_x * tiles_y * tiles_z;
    if (nb > 0x7fffffffL) return fail(SCAFES_ERR_ARG, "grid too large");
    pennes_adjk_kernel<T, LO, kKTX, kKTY, K, kKD, TU><<<dim3((unsigned)nb), dim3(kKTX, kKTY), 0, c->stream>>>(
        g, (const T*)lam, (T*)lamo, (const T*)w, st, grad, cf, kKZC, tiles_x, tiles_y);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

template <typename T, bool LO, bool TU>
int adjk_launch_k(Ctx* c, const Geo& g, int K, const void* lam, void* lamo, const void* w, const AdjStates& st,
                  double* grad, const AdjCoef& cf)
{
    switch (K) {
    case 2: return adjk_launch_t<T, LO, 2, TU>(c, g, lam, lamo, w, st, grad, cf);
    case 3: return adjk_launch_t<T, LO, 3, TU>(c, g, lam, lamo, w, st, grad, cf);
    case 4: return adjk_launch_t<T, LO, 4, TU>(c, g, lam, lamo, w, st, grad, cf);
    default: return fail(SCAFES_ERR_ARG, "multi-step adjoint: K must be 2..4");
    }
}

}  // namespace

bool pennes_adjk_supported(const Geo& geo)
{
    // 3-D only; a workgroup's chunk of the FP64 gradient (at most kKZC + 2K
    // + kKD planes) plus the in-plane offsets must stay below 2^31 - 1 bytes
    return geo.dim == 3 && ((long)(kKZC + 2 * kAdjMaxK + kKD + 2) * geo.e01 + geo.base + geo.e01) * 8L <
                               (long)kRecords;
}

int pennes_adjk_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, int K, const void* lam,
                       void* lam_out, const void* w, const void* const* t, const void* const* tlo, double* grad,
                       const scafes_pennes_params* p)
{
    if (!pennes_adjk_supported(geo)) return fail(SCAFES_ERR_UNSUPPORTED, "multi-step adjoint needs a 3-D grid");
    if (K < 2 || K > kAdjMaxK) return fail(SCAFES_ERR_ARG, "multi-step adjoint: K must be 2..4");
    if (p->border_mode != SCAFES_BORDER_KEEP && p->border_mode != SCAFES_BORDER_TUMOR)
        return fail(SCAFES_ERR_ARG, "multi-step adjoint: border_mode must be KEEP or TUMOR");
    const bool tumor = p->border_mode == SCAFES_BORDER_TUMOR;
    if (tumor && p->k_cond == 0.0) return fail(SCAFES_ERR_ARG, "multi-step adjoint: the TUMOR border needs k_cond != 0");
    AdjStates st{};
    for (int l = 0; l < K; ++l) {
        if (!t[l]) return fail(SCAFES_ERR_ARG, "multi-step adjoint: missing state");
        st.t[l] = t[l];
        st.lo[l] = tlo ? tlo[l] : nullptr;
        if (tlo && !tlo[l]) return fail(SCAFES_ERR_ARG, "multi-step adjoint: lo parts mismatch");
    }
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
    // the one-step Tumor adjoint's coefficients: (2 h_p)/K per axis, H
    cf.a2h2 = tumor ? (2.0 * g->h[2]) / p->k_cond : 0.0;
    cf.hc = tumor ? p->h_conv : 0.0;
    cf.fa2h2 = (float)cf.a2h2;
    cf.fhc = (float)cf.hc;
    if (dtype == SCAFES_F64) {
        if (tlo) return fail(SCAFES_ERR_ARG, "multi-step adjoint: FP64 states have no lo part");
        return tumor ? adjk_launch_k<double, false, true>(c, geo, K, lam, lam_out, w, st, grad, cf)
                     : adjk_launch_k<double, false, false>(c, geo, K, lam, lam_out, w, st, grad, cf);
    }
    if (tlo)
        return tumor ? adjk_launch_k<float, true, true>(c, geo, K, lam, lam_out, w, st, grad, cf)
                     : adjk_launch_k<float, true, false>(c, geo, K, lam, lam_out, w, st, grad, cf);
    return tumor ? adjk_launch_k<float, false, true>(c, geo, K, lam, lam_out, w, st, grad, cf)
                 : adjk_launch_k<float, false, false>(c, geo, K, lam, lam_out, w, st, grad, cf);
}

}  // namespace scafes
