// heatvar_tb2.hip -- two (or three) time steps per HBM sweep for the
// variable-kappa heat update (temporal blocking) on 3-D grids, same arithmetic per node as the
// one-step kernel in stencil3d.hip, so results are bit-identical to it in
// every precision (and, in FP64, to the CPU path).
//
// A TX x TY workgroup streams z planes.  Level 1 (u after one step) is
// computed for every thread position, i.e. for the output tile plus a
// one-node ring; level 2 (after two steps) for the (TX-2) x (TY-2) interior,
// which is the output tile.  Per plane k:
//   - level-0 plane k (u, kappa, plus a one-node ring) goes to LDS from a
//     D-deep register prefetch ring; the centre-only Kahan compensation too;
//   - level 1 at plane k from LDS neighbours and the z queue (k-1, k+1); its
//     six face coefficients (0.5*(kE+kC) ... in FP64, kE+kC ... in FP32) stay
//     in registers, since level 2 at the same node one plane later needs
//     exactly these values, and the upper z face of plane k is the lower one
//     of plane k+1;
//   - level 2 at plane k-1 from the level-1 LDS plane written one plane
//     before (in-plane neighbours) and a level-1 z queue (k-2, k).
// One barrier per plane; u, kappa and level-1 planes are double-buffered.
// Loads and stores are buffer accesses: a wave-uniform per-plane resource
// plus a 32-bit in-plane byte offset per thread, which keeps address
// arithmetic off the vector ALUs.
// HBM moves the one-step bytes once per two steps, plus the ring, mostly L2
// hits from the neighbouring tiles on the same XCD.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "heatvar_ops.hpp"
#include "scafes_common.hpp"

namespace scafes {

// Buffer access to one z plane: a wave-uniform resource (base = the plane's
// first memAll element, range = the in-plane extent incl. ghost planes below
// it) plus a 32-bit per-thread byte offset; the hardware range check returns
// 0 for, and drops, accesses beyond the range.
using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 family: 32-bit data format, raw (stride 0) buffer

__device__ inline Rsrc plane_rsrc(const char* plane, uint32_t range)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(plane), (short)0, (int)range, kRsrcWord3);
}

template <typename T>
__device__ inline T bload(Rsrc r, uint32_t off)
{
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
    }
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

template <typename T, int MODE, int TX, int TY, int D>
__global__ __launch_bounds__(TX* TY) void heatvar_tb2_kernel(Geo g, const T* __restrict__ u, const T* __restrict__ cin,
                                                              const T* __restrict__ kap, T* __restrict__ uout,
                                                              T* __restrict__ cout, VarCoef cf, int zc, int tiles_x,
                                                              int tiles_y)
{
    static_assert(D >= 2 && D % 2 == 0, "prefetch depth must be even");
    static_assert(TX % 64 == 0, "whole rows per wavefront: ty is wave-uniform");
    constexpr int OX = TX - 2, OY = TY - 2;
    constexpr bool KH = (MODE == 2);
    // level-0 and kappa planes with ring; row TY + 2 takes the y-ring writes
    // of the rows off the ring (a sink that is never read)
    __shared__ T s0[2][TY + 3][TX + 2];
    __shared__ T sk[2][TY + 3][TX + 2];
    __shared__ T s1[2][TY][TX];          // level-1 plane
    const int tx = threadIdx.x;
    const int ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per_plane = tiles_x * tiles_y;
    const int tz = lb / per_plane;
    const int r = lb - tz * per_plane;
    const int tyi = r / tiles_x;
    const int txi = r - tyi * tiles_x;
    const int i = txi * OX - 1 + tx, j = tyi * OY - 1 + ty;
    const int ic = min(max(i, 0), g.nx - 1), jc = min(max(j, 0), g.ny - 1);
    // in-plane interior (also false outside the grid): border and outside
    // positions keep (or take the border value) at both levels
    const bool inner_xy = i >= g.nl0 && i <= g.nx - 1 - g.nl0 && j >= g.nl1 && j <= g.ny - 1 - g.nl1;
    const bool col_out = tx >= 1 && tx <= TX - 2 && i < g.nx;  // lane-dependent
    const bool row_out = ty >= 1 && ty <= TY - 2 && j < g.ny;  // wave-uniform
    const int txm = max(tx - 1, 0), txp = min(tx + 1, TX - 1);
    const int k0 = tz * zc, k1 = min(k0 + zc, g.nz);
    const int kstart = k0 - 1;  // first plane with a level-1 value
    const int kcen = k1 + 1;    // last plane whose centre is needed
    const int klo = g.nl2, khi = g.nz - 1 - g.nl2;  // interior planes
    // halo duty for the level-0 ring
    const bool hx = (tx == 0) || (tx == TX - 1);
    const bool hy = (ty == 0) || (ty == TY - 1);
    const int hxi = (tx == 0) ? max(ic - 1, 0) : min(ic + 1, g.nx - 1);
    const int hyj = (ty == 0) ? max(jc - 1, 0) : min(jc + 1, g.ny - 1);
    // x ring: a lane off the ring reloads and rewrites its own centre value
    const int sxc = hx ? ((tx == 0) ? 0 : TX + 1) : tx + 1;
    const int syc = hy ? ((ty == 0) ? 0 : TY + 1) : TY + 2;
    constexpr uint32_t B = sizeof(T);
    const uint32_t oc = (uint32_t)(g.base + ic + g.e0 * jc) * B;
    const uint32_t ohx = hx ? (uint32_t)(g.base + hxi + g.e0 * jc) * B : oc;
    const uint32_t ohy = (uint32_t)(g.base + ic + g.e0 * hyj) * B;
    const size_t pb = (size_t)g.e01 * B;                      // plane stride in bytes
    const uint32_t range = (uint32_t)((g.base + g.e01) * B);  // in-plane offsets stay below this
    const uint32_t oo = col_out ? (uint32_t)(g.base + i + g.e0 * j) * B : range;  // output node
    const char* ub = (const char*)u;
    const char* kb8 = (const char*)kap;
    const char* cb = (const char*)cin;
    auto plane = [&](int kk) { return (size_t)min(max(kk, 0), g.nz - 1) * pb; };

    T qa[D], qb[D], qc[D], qax[D], qbx[D], qay[D], qby[D];
#pragma unroll
    for (int s = 0; s < D; ++s) qa[s] = qb[s] = qc[s] = qax[s] = qbx[s] = qay[s] = qby[s] = T(0);
    // Every iteration issues the same loads and stores; unneeded ones use
    // offset `range` (out of range: loads read 0 without a memory access,
    // stores are dropped), so the compiler can wait for exactly the slot a
    // plane consumes instead of for every outstanding access.
    auto fetch = [&](const int slot, const int kk) {
        const size_t o = plane(kk);
        const Rsrc ru = plane_rsrc(ub + o, range), rk = plane_rsrc(kb8 + o, range);
        const bool cen = kk <= kcen, mid = kk <= k1;
        const uint32_t occ = cen ? oc : range, ocm = mid ? oc : range;
        const uint32_t ox = mid ? ohx : range, oy = (mid && hy) ? ohy : range;
        qa[slot] = bload<T>(ru, occ);
        qb[slot] = bload<T>(rk, occ);
        if constexpr (KH) qc[slot] = bload<T>(plane_rsrc(cb + o, range), ocm);
        qax[slot] = bload<T>(ru, ox);
        qbx[slot] = bload<T>(rk, ox);
        qay[slot] = bload<T>(ru, oy);
        qby[slot] = bload<T>(rk, oy);
    };
    T am = bload<T>(plane_rsrc(ub + plane(kstart - 1), range), oc);
    const T bm = bload<T>(plane_rsrc(kb8 + plane(kstart - 1), range), oc);
#pragma unroll
    for (int s = 0; s < D; ++s) fetch(s, kstart + s);

    T l1m2 = T(0), l1m1 = T(0), c1m1 = T(0);  // level 1 at k-2, k-1 and its compensation at k-1
    Faces<T> fp = {};                          // faces of plane k-1 (level 2 reuses them)
    T fup = face<T, MODE>(qb[0], bm);          // upper z face of plane k-1 = lower face of plane k
    const T bv = (T)cf.border_value;
    const bool keep = cf.border_mode != SCAFES_BORDER_VALUE;
    for (int kb = kstart; kb <= k1; kb += D) {
#pragma unroll
        for (int s = 0; s < D; ++s) {
            const int k = kb + s;
            // uniform over the workgroup; nothing follows the plane loop, so
            // the exit leaves the register queues dead (no copies on this path)
            if (k > k1) return;
            const int buf = s & 1;
            const T ac = qa[s], ap = qa[(s + 1) % D];
            const T bc = qb[s], bp = qb[(s + 1) % D];
            const T c0 = qc[s];
            s0[buf][ty + 1][tx + 1] = ac;
            sk[buf][ty + 1][tx + 1] = bc;
            s0[buf][ty + 1][sxc] = qax[s];
            sk[buf][ty + 1][sxc] = qbx[s];
            s0[buf][syc][tx + 1] = qay[s];
            sk[buf][syc][tx + 1] = qby[s];
            fetch(s, k + D);
            __syncthreads();
            // ---- level 1 at (i, j, k), computed for every position, kept where not interior
            Faces<T> f;
            f.e = face<T, MODE>(sk[buf][ty + 1][tx + 2], bc);
            f.w = face<T, MODE>(bc, sk[buf][ty + 1][tx]);
            f.n = face<T, MODE>(sk[buf][ty + 2][tx + 1], bc);
            f.s = face<T, MODE>(bc, sk[buf][ty][tx + 1]);
            f.u = face<T, MODE>(bp, bc);
            f.d = fup;
            T c1;
            T l1 = var_update<T, MODE>(cf, ac, s0[buf][ty + 1][tx], s0[buf][ty + 1][tx + 2], s0[buf][ty][tx + 1],
                                       s0[buf][ty + 2][tx + 1], am, ap, f, c0, &c1);
            if (!(inner_xy && k >= klo && k <= khi)) {
                l1 = keep ? ac : bv;
                c1 = keep ? c0 : T(0);
            }
            s1[buf][ty][tx] = l1;
            // ---- level 2 at (i, j, k-1) from the level-1 plane of the previous iteration
            const int k2 = k - 1;
            {
                // computed by the output rows on output planes (a wave-uniform
                // branch); the store is issued by every thread, out of range
                // off the output columns and planes, so dropped there
                const bool st = k2 >= k0 && row_out;
                T v = T(0), cv = T(0);
                if (st) {
                    const int qb1 = buf ^ 1;
                    v = var_update<T, MODE>(cf, l1m1, s1[qb1][ty][txm], s1[qb1][ty][txp], s1[qb1][ty - 1][tx],
                                            s1[qb1][ty + 1][tx], l1m2, l1, fp, c1m1, &cv);
                    if (!(inner_xy && k2 >= klo && k2 <= khi)) {
                        v = keep ? l1m1 : bv;
                        cv = keep ? c1m1 : T(0);
                    }
                }
                const size_t o = plane(k2);
                const uint32_t os = st ? oo : range;
                bstore<T>(v, plane_rsrc((const char*)uout + o, range), os);
                if constexpr (KH) bstore<T>(cv, plane_rsrc((const char*)cout + o, range), os);
            }
            l1m2 = l1m1;
            l1m1 = l1;
            c1m1 = c1;
            fp = f;
            fup = f.u;
            am = ac;
        }
    }
}

// Three steps per sweep.  Threads cover the output tile plus a two-node
// ring: level 1 (u after one step) on every thread position, level 2 on
// positions one node inside, level 3 on the (TX-4) x (TY-4) output tile.
// Iteration kk computes level 1 at plane kk, level 2 at kk-1 and level 3 at
// kk-2; levels 2 and 3 read the in-plane neighbours of the level below from
// an LDS plane written one iteration earlier (one barrier per plane, all
// planes double-buffered).  The face coefficients of a node are the same
// at every level, so levels 2 and 3 take them from a register queue.
// A tile row is one wavefront (TX a multiple of 64), so row predicates are
// wave-uniform branches; lane predicates never branch: off-ring lanes load
// and rewrite their own centre, off-ring rows write the y ring into a sink
// row, and lanes off the output columns store out of range.  Measured on an
// MI355X at 512^3 (64x12, us per three-step sweep): F32 + Kahan 835 -> 813,
// F32 701 -> 631, F64 (64x16) 1882 -> 1767 against the branching kernel;
// SQ counters for F32 + Kahan: vector ALU 0.64 of its issue rate, scalar
// 0.51 (0.66 / 0.65 before).
// MO: the states after one and two steps are written too (u1o, u2o; the
// value part only in the compensated mode), at every node of the output
// tile, as the one-step kernel would write them.
template <typename T, int MODE, int TX, int TY, int D, bool MO = false>
__global__ __launch_bounds__(TX* TY) void heatvar_tb3_kernel(Geo g, const T* __restrict__ u, const T* __restrict__ cin,
                                                              const T* __restrict__ kap, T* __restrict__ uout,
                                                              T* __restrict__ cout, T* __restrict__ u1o,
                                                              T* __restrict__ u2o, VarCoef cf, int zc, int tiles_x,
                                                              int tiles_y)
{
    static_assert(D >= 2 && D % 2 == 0, "prefetch depth must be even");
    static_assert(TX % 64 == 0, "whole rows per wavefront: ty is wave-uniform");
    constexpr int OX = TX - 4, OY = TY - 4;
    constexpr bool KH = (MODE == 2);
    // level-0 and kappa planes with ring; row TY + 2 takes the y-ring writes
    // of the rows off the ring (a sink that is never read)
    __shared__ T s0[2][TY + 3][TX + 2];
    __shared__ T sk[2][TY + 3][TX + 2];
    __shared__ T s1[2][TY][TX];          // level-1 plane
    __shared__ T s2[2][TY][TX];          // level-2 plane
    const int tx = threadIdx.x;
    const int ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per_plane = tiles_x * tiles_y;
    const int tz = lb / per_plane;
    const int r = lb - tz * per_plane;
    const int tyi = r / tiles_x;
    const int txi = r - tyi * tiles_x;
    const int i = txi * OX - 2 + tx, j = tyi * OY - 2 + ty;
    const int ic = min(max(i, 0), g.nx - 1), jc = min(max(j, 0), g.ny - 1);
    const bool inner_xy = i >= g.nl0 && i <= g.nx - 1 - g.nl0 && j >= g.nl1 && j <= g.ny - 1 - g.nl1;
    const bool col_out = tx >= 2 && tx <= TX - 3 && i < g.nx;  // lane-dependent
    const bool row_out = ty >= 2 && ty <= TY - 3 && j < g.ny;  // wave-uniform
    // clamped neighbour slots of levels 2 and 3 (clamped reads only feed
    // positions that are never stored)
    const int txm = max(tx - 1, 0), txp = min(tx + 1, TX - 1), tym = max(ty - 1, 0), typ = min(ty + 1, TY - 1);
    const int k0 = tz * zc, k1 = min(k0 + zc, g.nz);
    const int kfirst = k0 - 2;  // first plane with a level-1 value
    const int klast = k1 + 1;   // last iteration (level 1 at k1 + 1)
    const int kcen = k1 + 2;    // last plane whose centre is read
    const int klo = g.nl2, khi = g.nz - 1 - g.nl2;
    const bool hx = (tx == 0) || (tx == TX - 1);
    const bool hy = (ty == 0) || (ty == TY - 1);
    const int hxi = (tx == 0) ? max(ic - 1, 0) : min(ic + 1, g.nx - 1);
    const int hyj = (ty == 0) ? max(jc - 1, 0) : min(jc + 1, g.ny - 1);
    // x ring: a lane off the ring reloads and rewrites its own centre value
    const int sxc = hx ? ((tx == 0) ? 0 : TX + 1) : tx + 1;
    const int syc = hy ? ((ty == 0) ? 0 : TY + 1) : TY + 2;
    constexpr uint32_t B = sizeof(T);
    const uint32_t oc = (uint32_t)(g.base + ic + g.e0 * jc) * B;
    const uint32_t ohx = hx ? (uint32_t)(g.base + hxi + g.e0 * jc) * B : oc;
    const uint32_t ohy = (uint32_t)(g.base + ic + g.e0 * hyj) * B;
    const size_t pb = (size_t)g.e01 * B;
    const uint32_t range = (uint32_t)((g.base + g.e01) * B);
    const uint32_t oo = col_out ? (uint32_t)(g.base + i + g.e0 * j) * B : range;  // output node
    const char* ub = (const char*)u;
    const char* kb8 = (const char*)kap;
    const char* cb = (const char*)cin;
    auto plane = [&](int kk) { return (size_t)min(max(kk, 0), g.nz - 1) * pb; };

    T qa[D], qb[D], qc[D], qax[D], qbx[D], qay[D], qby[D];
#pragma unroll
    for (int s = 0; s < D; ++s) qa[s] = qb[s] = qc[s] = qax[s] = qbx[s] = qay[s] = qby[s] = T(0);
    // Every iteration issues the same loads (and stores, below): a load that
    // is not needed reads at offset `range`, which the buffer unit answers
    // with 0 without a memory access.  With a fixed count of memory
    // instructions per plane the compiler waits only for the slot a plane
    // consumes (vmcnt(n) instead of vmcnt(0)), so the prefetch ring hides the
    // latency of the loads issued D planes ahead.
    auto fetch = [&](const int slot, const int kk) {
        const size_t o = plane(kk);
        const Rsrc ru = plane_rsrc(ub + o, range), rk = plane_rsrc(kb8 + o, range);
        const bool cen = kk <= kcen, mid = kk <= klast;
        const uint32_t occ = cen ? oc : range, ocm = mid ? oc : range;
        const uint32_t ox = mid ? ohx : range, oy = (mid && hy) ? ohy : range;
        qa[slot] = bload<T>(ru, occ);
        qb[slot] = bload<T>(rk, occ);
        if constexpr (KH) qc[slot] = bload<T>(plane_rsrc(cb + o, range), ocm);
        qax[slot] = bload<T>(ru, ox);
        qbx[slot] = bload<T>(rk, ox);
        qay[slot] = bload<T>(ru, oy);
        qby[slot] = bload<T>(rk, oy);
    };
    T am = bload<T>(plane_rsrc(ub + plane(kfirst - 1), range), oc);
    const T bm = bload<T>(plane_rsrc(kb8 + plane(kfirst - 1), range), oc);
#pragma unroll
    for (int s = 0; s < D; ++s) fetch(s, kfirst + s);

    T l1m1 = T(0), l1m2 = T(0), c1m1 = T(0);  // level 1 at kk-1, kk-2; compensation at kk-1
    T l2m1 = T(0), l2m2 = T(0), c2m1 = T(0);  // level 2 at kk-2, kk-3; compensation at kk-2
    Faces<T> fq1 = {}, fq2 = {};                // faces of planes kk-1, kk-2
    T fup = face<T, MODE>(qb[0], bm);
    const T bv = (T)cf.border_value;
    const bool keep = cf.border_mode != SCAFES_BORDER_VALUE;
    // The z queues below (faces, level values) shift every plane with period
    // 3; the body covers lcm(D, 3) planes (6 at D = 2).  The shifts still cost
    // about 30 register moves per plane (the wave-uniform branches merge
    // values), but this form needs 78 VGPRs (FP32 + Kahan): two 12-wave
    // workgroups per CU.  Register rings indexed by the unrolled plane (no
    // moves, 28 % fewer vector instructions) needed 102 VGPRs and took a
    // three-step sweep at 512^3 from 815 to 880-960 us (one workgroup per CU;
    // capped at 80 VGPRs they spill: 2100 us), the FP64 leg from 222 to 148
    // GLUPS.  s is the prefetch slot, buf the LDS parity.
    constexpr int U = (D == 2) ? 6 : D;
    for (int kb = kfirst; kb <= klast; kb += U) {
#pragma unroll
        for (int uu = 0; uu < U; ++uu) {
            const int s = uu % D;
            const int k = kb + uu;
            // uniform over the workgroup; nothing follows the plane loop, so
            // the exit leaves the register queues dead (no copies on this path)
            if (k > klast) return;
            const int buf = uu & 1, pbf = buf ^ 1;
            const T ac = qa[s], ap = qa[(s + 1) % D];
            const T bc = qb[s], bp = qb[(s + 1) % D];
            const T c0 = qc[s];
            s0[buf][ty + 1][tx + 1] = ac;
            sk[buf][ty + 1][tx + 1] = bc;
            s0[buf][ty + 1][sxc] = qax[s];
            sk[buf][ty + 1][sxc] = qbx[s];
            s0[buf][syc][tx + 1] = qay[s];
            sk[buf][syc][tx + 1] = qby[s];
            fetch(s, k + D);
            __syncthreads();
            // ---- level 1 at plane k
            Faces<T> f;
            f.e = face<T, MODE>(sk[buf][ty + 1][tx + 2], bc);
            f.w = face<T, MODE>(bc, sk[buf][ty + 1][tx]);
            f.n = face<T, MODE>(sk[buf][ty + 2][tx + 1], bc);
            f.s = face<T, MODE>(bc, sk[buf][ty][tx + 1]);
            f.u = face<T, MODE>(bp, bc);
            f.d = fup;
            T c1;
            T l1 = var_update<T, MODE>(cf, ac, s0[buf][ty + 1][tx], s0[buf][ty + 1][tx + 2], s0[buf][ty][tx + 1],
                                       s0[buf][ty + 2][tx + 1], am, ap, f, c0, &c1);
            if (!(inner_xy && k >= klo && k <= khi)) {
                l1 = keep ? ac : bv;
                c1 = keep ? c0 : T(0);
            }
            s1[buf][ty][tx] = l1;
            if constexpr (MO) {
                const uint32_t os1 = (k >= k0 && k < k1 && row_out) ? oo : range;
                bstore<T>(l1, plane_rsrc((const char*)u1o + plane(k), range), os1);
            }
            // ---- level 2 at plane k-1 from the level-1 plane written one
            // iteration ago; level 3 reads it on rows 1..TY-2 only (a
            // wave-uniform branch skips the outer rows)
            const int k2 = k - 1;
            T c2 = T(0), l2 = T(0);
            if (ty >= 1 && ty <= TY - 2) {
                l2 = var_update<T, MODE>(cf, l1m1, s1[pbf][ty][txm], s1[pbf][ty][txp], s1[pbf][tym][tx],
                                         s1[pbf][typ][tx], l1m2, l1, fq1, c1m1, &c2);
                if (!(inner_xy && k2 >= klo && k2 <= khi)) {
                    l2 = keep ? l1m1 : bv;
                    c2 = keep ? c1m1 : T(0);
                }
            }
            s2[buf][ty][tx] = l2;
            if constexpr (MO) {
                const uint32_t os2 = (k2 >= k0 && k2 < k1 && row_out) ? oo : range;
                bstore<T>(l2, plane_rsrc((const char*)u2o + plane(k2), range), os2);
            }
            // ---- level 3 at plane k-2 from the level-2 plane written one iteration ago
            const int k3 = k - 2;
            {
                // computed by the output rows on output planes (a wave-uniform
                // branch); the store is issued by every thread, out of range
                // off the output columns and planes, so dropped there
                const bool st = k3 >= k0 && k3 < k1 && row_out;
                T l3 = T(0), c3 = T(0);
                if (st) {
                    l3 = var_update<T, MODE>(cf, l2m1, s2[pbf][ty][txm], s2[pbf][ty][txp], s2[pbf][ty - 1][tx],
                                             s2[pbf][ty + 1][tx], l2m2, l2, fq2, c2m1, &c3);
                    if (!(inner_xy && k3 >= klo && k3 <= khi)) {
                        l3 = keep ? l2m1 : bv;
                        c3 = keep ? c2m1 : T(0);
                    }
                }
                const size_t o = plane(k3);
                const uint32_t os = st ? oo : range;
                bstore<T>(l3, plane_rsrc((const char*)uout + o, range), os);
                if constexpr (KH) bstore<T>(c3, plane_rsrc((const char*)cout + o, range), os);
            }
            l2m2 = l2m1;
            l2m1 = l2;
            c2m1 = c2;
            l1m2 = l1m1;
            l1m1 = l1;
            c1m1 = c1;
            fq2 = fq1;
            fq1 = f;
            fup = f.u;
            am = ac;
        }
    }
}

// Four steps per sweep: the three-step scheme with one more level.  Threads
// cover the output tile plus a three-node ring; level L (u after L steps) is
// valid L-1 nodes inside the thread tile, level 4 on the (TX-6) x (TY-6)
// output tile.  Iteration kk computes level L at plane kk-(L-1); levels 2..4
// read the in-plane neighbours of the level below from its LDS plane written
// one iteration earlier and the z neighbours from register queues (one
// barrier per plane, planes double-buffered).  The faces of a node ride in a
// three-deep register queue for levels 2..4.  Same branch-free form as the
// three-step kernel: wave-uniform row branches, off-ring rows writing the y
// ring into a sink row, lanes off the output columns storing out of range.
// Per four steps HBM moves one sweep of u, kappa (and c) plus the ring.
// MO: the states after one, two and three steps are written too (u1o, u2o,
// u3o; the value part only in the compensated mode) at every node of the
// output tile, as the one-step kernel would write them.
template <typename T, int MODE, int TX, int TY, int D, bool MO = false>
__global__ __launch_bounds__(TX* TY) void heatvar_tb4_kernel(Geo g, const T* __restrict__ u, const T* __restrict__ cin,
                                                              const T* __restrict__ kap, T* __restrict__ uout,
                                                              T* __restrict__ cout, T* __restrict__ u1o,
                                                              T* __restrict__ u2o, T* __restrict__ u3o, VarCoef cf,
                                                              int zc, int tiles_x, int tiles_y)
{
    static_assert(D >= 2 && D % 2 == 0, "prefetch depth must be even");
    static_assert(TX % 64 == 0, "whole rows per wavefront: ty is wave-uniform");
    static_assert(TY >= 8, "the output tile needs rows");
    constexpr int OX = TX - 6, OY = TY - 6;
    constexpr bool KH = (MODE == 2);
    __shared__ T s0[2][TY + 3][TX + 2];
    __shared__ T sk[2][TY + 3][TX + 2];
    __shared__ T s1[2][TY][TX];  // level-1 plane
    __shared__ T s2[2][TY][TX];  // level-2 plane
    __shared__ T s3[2][TY][TX];  // level-3 plane
    const int tx = threadIdx.x;
    const int ty = __builtin_amdgcn_readfirstlane((int)threadIdx.y);
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per_plane = tiles_x * tiles_y;
    const int tz = lb / per_plane;
    const int r = lb - tz * per_plane;
    const int tyi = r / tiles_x;
    const int txi = r - tyi * tiles_x;
    const int i = txi * OX - 3 + tx, j = tyi * OY - 3 + ty;
    const int ic = min(max(i, 0), g.nx - 1), jc = min(max(j, 0), g.ny - 1);
    const bool inner_xy = i >= g.nl0 && i <= g.nx - 1 - g.nl0 && j >= g.nl1 && j <= g.ny - 1 - g.nl1;
    const bool col_out = tx >= 3 && tx <= TX - 4 && i < g.nx;  // lane-dependent
    const bool row_out = ty >= 3 && ty <= TY - 4 && j < g.ny;  // wave-uniform
    // clamped neighbour slots of level 2 (clamped reads only feed positions
    // that are never used); levels 3 and 4 run on rows whose neighbours exist
    const int txm = max(tx - 1, 0), txp = min(tx + 1, TX - 1), tym = max(ty - 1, 0), typ = min(ty + 1, TY - 1);
    const int k0 = tz * zc, k1 = min(k0 + zc, g.nz);
    const int kfirst = k0 - 3;  // first plane with a level-1 value
    const int klast = k1 + 2;   // last iteration (level 1 at k1 + 2)
    const int kcen = k1 + 3;    // last plane whose centre is read
    const int klo = g.nl2, khi = g.nz - 1 - g.nl2;
    const bool hx = (tx == 0) || (tx == TX - 1);
    const bool hy = (ty == 0) || (ty == TY - 1);
    const int hxi = (tx == 0) ? max(ic - 1, 0) : min(ic + 1, g.nx - 1);
    const int hyj = (ty == 0) ? max(jc - 1, 0) : min(jc + 1, g.ny - 1);
    const int sxc = hx ? ((tx == 0) ? 0 : TX + 1) : tx + 1;
    const int syc = hy ? ((ty == 0) ? 0 : TY + 1) : TY + 2;
    constexpr uint32_t B = sizeof(T);
    const uint32_t oc = (uint32_t)(g.base + ic + g.e0 * jc) * B;
    const uint32_t ohx = hx ? (uint32_t)(g.base + hxi + g.e0 * jc) * B : oc;
    const uint32_t ohy = (uint32_t)(g.base + ic + g.e0 * hyj) * B;
    const size_t pb = (size_t)g.e01 * B;
    const uint32_t range = (uint32_t)((g.base + g.e01) * B);
    const uint32_t oo = col_out ? (uint32_t)(g.base + i + g.e0 * j) * B : range;  // output node
    const char* ub = (const char*)u;
    const char* kb8 = (const char*)kap;
    const char* cb = (const char*)cin;
    auto plane = [&](int kk) { return (size_t)min(max(kk, 0), g.nz - 1) * pb; };

    T qa[D], qb[D], qc[D], qax[D], qbx[D], qay[D], qby[D];
#pragma unroll
    for (int s = 0; s < D; ++s) qa[s] = qb[s] = qc[s] = qax[s] = qbx[s] = qay[s] = qby[s] = T(0);
    // the same fixed set of loads every plane (unneeded ones read at offset
    // `range`, answered with 0 without a memory access)
    auto fetch = [&](const int slot, const int kk) {
        const size_t o = plane(kk);
        const Rsrc ru = plane_rsrc(ub + o, range), rk = plane_rsrc(kb8 + o, range);
        const bool cen = kk <= kcen, mid = kk <= klast;
        const uint32_t occ = cen ? oc : range, ocm = mid ? oc : range;
        const uint32_t ox = mid ? ohx : range, oy = (mid && hy) ? ohy : range;
        qa[slot] = bload<T>(ru, occ);
        qb[slot] = bload<T>(rk, occ);
        if constexpr (KH) qc[slot] = bload<T>(plane_rsrc(cb + o, range), ocm);
        qax[slot] = bload<T>(ru, ox);
        qbx[slot] = bload<T>(rk, ox);
        qay[slot] = bload<T>(ru, oy);
        qby[slot] = bload<T>(rk, oy);
    };
    T am = bload<T>(plane_rsrc(ub + plane(kfirst - 1), range), oc);
    const T bm = bload<T>(plane_rsrc(kb8 + plane(kfirst - 1), range), oc);
#pragma unroll
    for (int s = 0; s < D; ++s) fetch(s, kfirst + s);

    T l1m1 = T(0), l1m2 = T(0), c1m1 = T(0);  // level 1 at kk-1, kk-2; compensation at kk-1
    T l2m1 = T(0), l2m2 = T(0), c2m1 = T(0);  // level 2 at kk-2, kk-3; compensation at kk-2
    T l3m1 = T(0), l3m2 = T(0), c3m1 = T(0);  // level 3 at kk-3, kk-4; compensation at kk-3
    Faces<T> fq1 = {}, fq2 = {}, fq3 = {};      // faces of planes kk-1, kk-2, kk-3
    T fup = face<T, MODE>(qb[0], bm);
    const T bv = (T)cf.border_value;
    const bool keep = cf.border_mode != SCAFES_BORDER_VALUE;
    constexpr int U = (D == 2) ? 6 : D;
    for (int kb = kfirst; kb <= klast; kb += U) {
#pragma unroll
        for (int uu = 0; uu < U; ++uu) {
            const int s = uu % D;
            const int k = kb + uu;
            if (k > klast) return;  // uniform over the workgroup
            const int buf = uu & 1, pbf = buf ^ 1;
            const T ac = qa[s], ap = qa[(s + 1) % D];
            const T bc = qb[s], bp = qb[(s + 1) % D];
            const T c0 = qc[s];
            s0[buf][ty + 1][tx + 1] = ac;
            sk[buf][ty + 1][tx + 1] = bc;
            s0[buf][ty + 1][sxc] = qax[s];
            sk[buf][ty + 1][sxc] = qbx[s];
            s0[buf][syc][tx + 1] = qay[s];
            sk[buf][syc][tx + 1] = qby[s];
            fetch(s, k + D);
            __syncthreads();
            // ---- level 1 at plane k, every thread position
            Faces<T> f;
            f.e = face<T, MODE>(sk[buf][ty + 1][tx + 2], bc);
            f.w = face<T, MODE>(bc, sk[buf][ty + 1][tx]);
            f.n = face<T, MODE>(sk[buf][ty + 2][tx + 1], bc);
            f.s = face<T, MODE>(bc, sk[buf][ty][tx + 1]);
            f.u = face<T, MODE>(bp, bc);
            f.d = fup;
            T c1;
            T l1 = var_update<T, MODE>(cf, ac, s0[buf][ty + 1][tx], s0[buf][ty + 1][tx + 2], s0[buf][ty][tx + 1],
                                       s0[buf][ty + 2][tx + 1], am, ap, f, c0, &c1);
            if (!(inner_xy && k >= klo && k <= khi)) {
                l1 = keep ? ac : bv;
                c1 = keep ? c0 : T(0);
            }
            s1[buf][ty][tx] = l1;
            if constexpr (MO) {
                const uint32_t os1 = (k >= k0 && k < k1 && row_out) ? oo : range;
                bstore<T>(l1, plane_rsrc((const char*)u1o + plane(k), range), os1);
            }
            // ---- level 2 at plane k-1 on rows 1..TY-2
            const int k2 = k - 1;
            T c2 = T(0), l2 = T(0);
            if (ty >= 1 && ty <= TY - 2) {
                l2 = var_update<T, MODE>(cf, l1m1, s1[pbf][ty][txm], s1[pbf][ty][txp], s1[pbf][tym][tx],
                                         s1[pbf][typ][tx], l1m2, l1, fq1, c1m1, &c2);
                if (!(inner_xy && k2 >= klo && k2 <= khi)) {
                    l2 = keep ? l1m1 : bv;
                    c2 = keep ? c1m1 : T(0);
                }
            }
            s2[buf][ty][tx] = l2;
            if constexpr (MO) {
                const uint32_t os2 = (k2 >= k0 && k2 < k1 && row_out) ? oo : range;
                bstore<T>(l2, plane_rsrc((const char*)u2o + plane(k2), range), os2);
            }
            // ---- level 3 at plane k-2 on rows 2..TY-3
            const int k3 = k - 2;
            T c3 = T(0), l3 = T(0);
            if (ty >= 2 && ty <= TY - 3) {
                l3 = var_update<T, MODE>(cf, l2m1, s2[pbf][ty][txm], s2[pbf][ty][txp], s2[pbf][ty - 1][tx],
                                         s2[pbf][ty + 1][tx], l2m2, l2, fq2, c2m1, &c3);
                if (!(inner_xy && k3 >= klo && k3 <= khi)) {
                    l3 = keep ? l2m1 : bv;
                    c3 = keep ? c2m1 : T(0);
                }
            }
            s3[buf][ty][tx] = l3;
            if constexpr (MO) {
                const uint32_t os3 = (k3 >= k0 && k3 < k1 && row_out) ? oo : range;
                bstore<T>(l3, plane_rsrc((const char*)u3o + plane(k3), range), os3);
            }
            // ---- level 4 at plane k-3 on the output rows and planes
            const int k4 = k - 3;
            {
                const bool st = k4 >= k0 && k4 < k1 && row_out;
                T l4 = T(0), c4 = T(0);
                if (st) {
                    l4 = var_update<T, MODE>(cf, l3m1, s3[pbf][ty][txm], s3[pbf][ty][txp], s3[pbf][ty - 1][tx],
                                             s3[pbf][ty + 1][tx], l3m2, l3, fq3, c3m1, &c4);
                    if (!(inner_xy && k4 >= klo && k4 <= khi)) {
                        l4 = keep ? l3m1 : bv;
                        c4 = keep ? c3m1 : T(0);
                    }
                }
                const size_t o = plane(k4);
                const uint32_t os = st ? oo : range;
                bstore<T>(l4, plane_rsrc((const char*)uout + o, range), os);
                if constexpr (KH) bstore<T>(c4, plane_rsrc((const char*)cout + o, range), os);
            }
            l3m2 = l3m1;
            l3m1 = l3;
            c3m1 = c3;
            l2m2 = l2m1;
            l2m1 = l2;
            c2m1 = c2;
            l1m2 = l1m1;
            l1m1 = l1;
            c1m1 = c1;
            fq3 = fq2;
            fq2 = fq1;
            fq1 = f;
            fup = f.u;
            am = ac;
        }
    }
}

// ---------------------------------------------------------------------------
// Two reverse steps of the variable-kappa adjoint per sweep (lam^n ->
// lam^{n-2}) with both steps' kappa-gradient terms, 3-D grids.  Per node the
// arithmetic is that of HeatVarAdjOp in stencil3d.hip, in the same order:
//   lam^{n-1} = the forward step's operator applied to lam^n (var_update
//               with the node's six face coefficients), 0 on border nodes;
//   grad     += 0.5 * sum_p th_p * sum_{nb = -+e_p in the grid}
//               (lam^n_c - lam^n_nb)(u^{n-1}_nb - u^{n-1}_c)   (every node),
// and the FP64 gradient receives the step-n term, then the step-(n-1) term,
// so the result is bit-identical to two one-step launches.  The layout is
// the two-step forward's: level 1 (lam^{n-1}) at plane k for every thread
// position (the output tile plus a one-node ring; lam^n and kappa come with
// the ring), level 2 (lam^{n-2}) at plane k-1 on the output tile from the
// level-1 plane written one iteration earlier.  The gradient terms are formed
// on the output tile only, so u^{n-1} and u^{n-2} need no ring; the step-n
// term of plane k rides in a register to the level-2 pass one plane later.
// Per two reverse steps HBM moves lam r+w, kappa, u^{n-1}, u^{n-2} and the
// FP64 gradient r+w once: 36 B per node for FP32 fields (18 per step,
// against 32 for one-step launches).
template <typename T, int TX, int TY, int D>
__global__ __launch_bounds__(TX* TY) void heatvar_adj2_kernel(Geo g, const T* __restrict__ lam,
                                                               const T* __restrict__ kap, const T* __restrict__ u1,
                                                               const T* __restrict__ u2, T* __restrict__ lamo,
                                                               double* __restrict__ grad, AdjVarCoef cf, int zc,
                                                               int tiles_x, int tiles_y)
{
    static_assert(D >= 2 && D % 2 == 0, "prefetch depth must be even");
    constexpr int OX = TX - 2, OY = TY - 2;
    constexpr int MODE = sizeof(T) == 8 ? 0 : 1;
    __shared__ T s0[2][TY + 2][TX + 2];  // lam^n plane with ring
    __shared__ T sk[2][TY + 2][TX + 2];  // kappa plane with ring
    __shared__ T sa[2][TY][TX];          // u^{n-1} plane
    __shared__ T s1[2][TY][TX];          // lam^{n-1} plane
    __shared__ T sb[2][TY][TX];          // u^{n-2} plane
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
    const int kstart = k0 - 1;
    const int kcen = k1 + 1;
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
    const uint32_t oo = (uint32_t)(g.base + i + g.e0 * j) * B;
    const uint32_t og = (uint32_t)(g.base + (out_thread ? i : 0) + g.e0 * (out_thread ? j : 0)) * 8u;
    const size_t pb = (size_t)g.e01 * B, pg = (size_t)g.e01 * 8;
    const uint32_t range = (uint32_t)((g.base + g.e01) * B);
    const uint32_t rangeg = (uint32_t)((g.base + g.e01) * 8);
    const char* lb8 = (const char*)lam;
    const char* kb8 = (const char*)kap;
    const char* ab8 = (const char*)u1;
    const char* bb8 = (const char*)u2;
    const char* gb8 = (const char*)grad;
    auto plane = [&](int kk) { return (size_t)min(max(kk, 0), g.nz - 1); };

    T qa[D], qk[D], qu1[D], qu2[D], qax[D], qkx[D], qay[D], qky[D];
    double qg[D];
#pragma unroll
    for (int s = 0; s < D; ++s) {
        qa[s] = qk[s] = qu1[s] = qu2[s] = qax[s] = qkx[s] = qay[s] = qky[s] = T(0);
        qg[s] = 0.0;
    }
    auto fetch = [&](const int slot, const int kk) {
        const size_t pk = plane(kk);
        const size_t o = pk * pb;
        const Rsrc rl = plane_rsrc(lb8 + o, range), rk = plane_rsrc(kb8 + o, range);
        const bool cen = kk <= kcen, mid = kk <= k1;
        const uint32_t occ = cen ? oc : range, ocm = mid ? oc : range;
        const uint32_t ox = (mid && hx) ? ohx : range, oy = (mid && hy) ? ohy : range;
        qa[slot] = bload<T>(rl, occ);
        qk[slot] = bload<T>(rk, occ);
        qu1[slot] = bload<T>(plane_rsrc(ab8 + o, range), occ);
        qu2[slot] = bload<T>(plane_rsrc(bb8 + o, range), ocm);
        qax[slot] = bload<T>(rl, ox);
        qkx[slot] = bload<T>(rk, ox);
        qay[slot] = bload<T>(rl, oy);
        qky[slot] = bload<T>(rk, oy);
        qg[slot] = bload<double>(plane_rsrc(gb8 + pk * pg, rangeg), (mid && out_thread) ? og : rangeg);
    };
    T am = bload<T>(plane_rsrc(lb8 + plane(kstart - 1) * pb, range), oc);
    T u1m = bload<T>(plane_rsrc(ab8 + plane(kstart - 1) * pb, range), oc);
    const T km = bload<T>(plane_rsrc(kb8 + plane(kstart - 1) * pb, range), oc);
#pragma unroll
    for (int s = 0; s < D; ++s) fetch(s, kstart + s);

    T l1m2 = T(0), l1m1 = T(0);     // level 1 at planes k-2, k-1
    T u2m2 = T(0), u2m1 = T(0);     // u^{n-2} at planes k-2, k-1
    double t1m1 = 0.0, gm1 = 0.0;   // step-n term and gradient of plane k-1
    Faces<T> fp = {};
    T fup = face<T, MODE>(qk[0], km);
    for (int kb = kstart; kb <= k1; kb += D) {
#pragma unroll
        for (int s = 0; s < D; ++s) {
            const int k = kb + s;
            if (k > k1) break;  // uniform over the workgroup
            const int buf = s & 1, pbf = buf ^ 1;
            const T ac = qa[s], ap = qa[(s + 1) % D];
            const T bc = qk[s], bp = qk[(s + 1) % D];
            const T uc = qu1[s], up = qu1[(s + 1) % D];
            const T vc = qu2[s];
            const double gk = qg[s];
            s0[buf][ty + 1][tx + 1] = ac;
            sk[buf][ty + 1][tx + 1] = bc;
            sa[buf][ty][tx] = uc;
            if (hx) {
                s0[buf][ty + 1][sxh] = qax[s];
                sk[buf][ty + 1][sxh] = qkx[s];
            }
            if (hy) {
                s0[buf][syh][tx + 1] = qay[s];
                sk[buf][syh][tx + 1] = qky[s];
            }
            fetch(s, k + D);
            __syncthreads();
            // ---- level 1 at (i, j, k): lam^{n-1}, every position
            const T lw = s0[buf][ty + 1][tx], le = s0[buf][ty + 1][tx + 2];
            const T ls = s0[buf][ty][tx + 1], ln = s0[buf][ty + 2][tx + 1];
            Faces<T> f;
            f.e = face<T, MODE>(sk[buf][ty + 1][tx + 2], bc);
            f.w = face<T, MODE>(bc, sk[buf][ty + 1][tx]);
            f.n = face<T, MODE>(sk[buf][ty + 2][tx + 1], bc);
            f.s = face<T, MODE>(bc, sk[buf][ty][tx + 1]);
            f.u = face<T, MODE>(bp, bc);
            f.d = fup;
            T dummy;
            T l1 = var_update<T, MODE>(cf.v, ac, lw, le, ls, ln, am, ap, f, T(0), &dummy);
            if (!(inner_xy && k >= klo && k <= khi)) l1 = T(0);
            // planes read one iteration later are written after the barrier
            // (a write before it could overtake a slower thread still reading
            // the same buffer two planes back)
            s1[buf][ty][tx] = l1;
            sb[buf][ty][tx] = vc;
            // step-n gradient term at (i, j, k) (output positions: the in-plane
            // neighbours are inside the thread tile)
            double t1 = 0.0;
            if (out_thread) {
                const int txm = tx - 1, txp = tx + 1, tym = ty - 1, typ = ty + 1;
                t1 = adj_term3<T>(cf, i > 0, i < g.nx - 1, j > 0, j < g.ny - 1, k > 0, k < g.nz - 1, ac, lw, le, ls,
                                  ln, am, ap, uc, sa[buf][ty][txm], sa[buf][ty][txp], sa[buf][tym][tx],
                                  sa[buf][typ][tx], u1m, up);
            }
            // ---- level 2 at (i, j, k-1): lam^{n-2} and the step-(n-1) term
            const int k2 = k - 1;
            {
                const bool st = k2 >= k0 && out_thread;
                T v = T(0);
                double gv = 0.0;
                if (st) {
                    const T pw = s1[pbf][ty][tx - 1], pe = s1[pbf][ty][tx + 1];
                    const T ps = s1[pbf][ty - 1][tx], pn = s1[pbf][ty + 1][tx];
                    T dummy2;
                    v = var_update<T, MODE>(cf.v, l1m1, pw, pe, ps, pn, l1m2, l1, fp, T(0), &dummy2);
                    if (!(inner_xy && k2 >= klo && k2 <= khi)) v = T(0);
                    const double t2 = adj_term3<T>(cf, i > 0, i < g.nx - 1, j > 0, j < g.ny - 1, k2 > 0,
                                                   k2 < g.nz - 1, l1m1, pw, pe, ps, pn, l1m2, l1, u2m1,
                                                   sb[pbf][ty][tx - 1], sb[pbf][ty][tx + 1], sb[pbf][ty - 1][tx],
                                                   sb[pbf][ty + 1][tx], u2m2, vc);
                    gv = gm1 + t1m1;
                    gv = gv + t2;
                }
                const uint32_t os = st ? oo : range;
                bstore<T>(v, plane_rsrc((const char*)lamo + plane(k2) * pb, range), os);
                bstore<double>(gv, plane_rsrc((const char*)grad + plane(k2) * pg, rangeg), st ? og : rangeg);
            }
            l1m2 = l1m1;
            l1m1 = l1;
            u2m2 = u2m1;
            u2m1 = vc;
            t1m1 = t1;
            gm1 = gk;
            fp = f;
            fup = f.u;
            am = ac;
            u1m = uc;
        }
    }
}

// 64 x 8 threads, two planes in flight; FP32 64-plane z chunks, FP64 32.
// Measured at 512^3 FP32 on an MI355X (us per two-step launch): 64x8 with
// 64-plane chunks 1372, with 32 1395; 64x16 with 64 1408, with 32 1461;
// 64x12 with 32 1724 (the results are bit-identical across tiles).
template <typename T>
struct Adj2Tile {
    static constexpr int TX = 64, TY = 8, D = 2, ZC = sizeof(T) == 4 ? 64 : 32;
};

// 3-D grids whose in-plane byte offsets of the FP64 gradient fit 32 bits.
bool heatvar_adj2_supported(const Geo& geo)
{
    return geo.dim == 3 && (geo.base + geo.e01) * 8 < kOobLimit;
}

template <typename T>
static int adj2_launch_t(Ctx* c, const Geo& g, const void* lam, void* lamo, const void* kap, const void* u1,
                         const void* u2, double* grad, const AdjVarCoef& cf)
{
    typedef Adj2Tile<T> A;
    const int tiles_x = ceil_div(g.nx, A::TX - 2), tiles_y = ceil_div(g.ny, A::TY - 2);
    const long nb = (long)tiles_x * tiles_y * ceil_div(g.nz, A::ZC);
    if (nb > 0x7fffffffL) return fail(SCAFES_ERR_ARG, "grid too large");
    heatvar_adj2_kernel<T, A::TX, A::TY, A::D><<<dim3((unsigned)nb), dim3(A::TX, A::TY), 0, c->stream>>>(
        g, (const T*)lam, (const T*)kap, (const T*)u1, (const T*)u2, (T*)lamo, grad, cf, A::ZC, tiles_x, tiles_y);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

int heatvar_adj2_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int dtype, const void* lam, void* lamo,
                        const void* kap, const void* u1, const void* u2, double* grad, double tau)
{
    if (!heatvar_adj2_supported(geo)) return fail(SCAFES_ERR_UNSUPPORTED, "two-step adjoint needs a 3-D grid");
    AdjVarCoef cf;
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
    if (dtype == SCAFES_F64) return adj2_launch_t<double>(c, geo, lam, lamo, kap, u1, u2, grad, cf);
    return adj2_launch_t<float>(c, geo, lam, lamo, kap, u1, u2, grad, cf);
}

// Multi-step tiles: steps per sweep (2, 3 or 4), workgroup tile, z planes
// per workgroup, prefetch depth.  Earlier entries keep their indices.
struct TbCfg {
    int steps, tx, ty, zc, d;
};
static const TbCfg kTbCfgs[] = {
    {2, 64, 16, 64, 2}, {2, 64, 16, 128, 2}, {2, 64, 8, 64, 2},  {2, 128, 8, 64, 2}, {2, 64, 16, 64, 4},
    {2, 128, 4, 64, 2}, {2, 64, 8, 64, 4},   {2, 64, 8, 32, 2},  {2, 64, 8, 128, 2}, {2, 64, 4, 64, 2},
    {3, 64, 16, 64, 2}, {3, 64, 16, 32, 2},  {3, 64, 16, 128, 2}, {3, 64, 12, 64, 2}, {3, 64, 16, 64, 4},
    {3, 128, 8, 64, 2}, {3, 64, 8, 64, 2},   {3, 64, 12, 32, 2}, {3, 64, 12, 64, 4}, {3, 64, 12, 96, 2},
    {3, 64, 10, 64, 2}, {3, 64, 14, 64, 2},
    {4, 64, 16, 64, 2}, {4, 64, 16, 32, 2},  {4, 64, 16, 128, 2}, {4, 64, 16, 96, 2}, {4, 64, 16, 64, 4},
};
static constexpr int kNumTb = sizeof(kTbCfgs) / sizeof(kTbCfgs[0]);

template <int STEPS, typename T, int MODE, int TX, int TY, int D, bool MO = false>
static int tb_launch_t(Ctx* c, const Geo& g, const void* u, const void* cin, const void* kap, void* uo, void* co,
                       const VarCoef& cf, int zc, void* const* mid = nullptr)
{
    constexpr int RING = STEPS - 1;  // ring width of the thread tile on each side
    const int tiles_x = ceil_div(g.nx, TX - 2 * RING), tiles_y = ceil_div(g.ny, TY - 2 * RING);
    const int tiles_z = ceil_div(g.nz, zc);
    const long nb = (long)tiles_x * tiles_y * tiles_z;
    if (nb > 0x7fffffffL) return fail(SCAFES_ERR_ARG, "grid too large");
    if constexpr (STEPS == 2)
        heatvar_tb2_kernel<T, MODE, TX, TY, D><<<dim3((unsigned)nb), dim3(TX, TY), 0, c->stream>>>(
            g, (const T*)u, (const T*)cin, (const T*)kap, (T*)uo, (T*)co, cf, zc, tiles_x, tiles_y);
    else if constexpr (STEPS == 4)
        heatvar_tb4_kernel<T, MODE, TX, TY, D, MO><<<dim3((unsigned)nb), dim3(TX, TY), 0, c->stream>>>(
            g, (const T*)u, (const T*)cin, (const T*)kap, (T*)uo, (T*)co, MO ? (T*)mid[0] : nullptr,
            MO ? (T*)mid[1] : nullptr, MO ? (T*)mid[2] : nullptr, cf, zc, tiles_x, tiles_y);
    else
        heatvar_tb3_kernel<T, MODE, TX, TY, D, MO><<<dim3((unsigned)nb), dim3(TX, TY), 0, c->stream>>>(
            g, (const T*)u, (const T*)cin, (const T*)kap, (T*)uo, (T*)co, MO ? (T*)mid[0] : nullptr,
            MO ? (T*)mid[1] : nullptr, cf, zc, tiles_x, tiles_y);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

template <typename T, int MODE>
static int tb_dispatch(Ctx* c, const Geo& g, const void* u, const void* cin, const void* kap, void* uo, void* co,
                       const VarCoef& cf, int cfg)
{
    const TbCfg t = kTbCfgs[cfg];
#define SCAFES_TB(S, X, Y, DD)                                                                    \
    case S * 10000000 + X * 100000 + Y * 100 + DD:                                                \
        return tb_launch_t<S, T, MODE, X, Y, DD>(c, g, u, cin, kap, uo, co, cf, t.zc);
    switch (t.steps * 10000000 + t.tx * 100000 + t.ty * 100 + t.d) {
        SCAFES_TB(2, 64, 16, 2)
        SCAFES_TB(2, 64, 8, 2)
        SCAFES_TB(2, 64, 8, 4)
        SCAFES_TB(2, 64, 4, 2)
        SCAFES_TB(2, 128, 8, 2)
        SCAFES_TB(2, 64, 16, 4)
        SCAFES_TB(2, 128, 4, 2)
        SCAFES_TB(3, 64, 16, 2)
        SCAFES_TB(3, 64, 12, 2)
        SCAFES_TB(3, 64, 16, 4)
        SCAFES_TB(3, 128, 8, 2)
        SCAFES_TB(3, 64, 8, 2)
        SCAFES_TB(3, 64, 12, 4)
        SCAFES_TB(3, 64, 10, 2)
        SCAFES_TB(3, 64, 14, 2)
        SCAFES_TB(4, 64, 16, 2)
        SCAFES_TB(4, 64, 16, 4)
    default: return fail(SCAFES_ERR_ARG, "bad multi-step tile configuration");
    }
#undef SCAFES_TB
}

// Multi-step sweeps that also write the intermediate states: the default
// three-step tiles (F64 64x16, F32 64x12, prefetch depth 2) and the 64x16
// four-step tile with prefetch depth 2.
template <typename T, int MODE>
static int tb_dispatch_mo(Ctx* c, const Geo& g, const void* u, const void* cin, const void* kap, void* uo, void* co,
                          const VarCoef& cf, int cfg, void* const* mid)
{
    const TbCfg t = kTbCfgs[cfg];
    if (t.steps == 3 && t.tx == 64 && t.ty == 16 && t.d == 2)
        return tb_launch_t<3, T, MODE, 64, 16, 2, true>(c, g, u, cin, kap, uo, co, cf, t.zc, mid);
    if (t.steps == 3 && t.tx == 64 && t.ty == 12 && t.d == 2)
        return tb_launch_t<3, T, MODE, 64, 12, 2, true>(c, g, u, cin, kap, uo, co, cf, t.zc, mid);
    if (t.steps == 4 && t.tx == 64 && t.ty == 16 && t.d == 2)
        return tb_launch_t<4, T, MODE, 64, 16, 2, true>(c, g, u, cin, kap, uo, co, cf, t.zc, mid);
    return fail(SCAFES_ERR_UNSUPPORTED, "no state-writing variant of this multi-step tile");
}

bool heatvar_tb_states_supported(int cfg)
{
    if (cfg < 0 || cfg >= kNumTb) return false;
    const TbCfg t = kTbCfgs[cfg];
    return (t.steps == 3 && t.tx == 64 && (t.ty == 16 || t.ty == 12) && t.d == 2) ||
           (t.steps == 4 && t.tx == 64 && t.ty == 16 && t.d == 2);
}

int heatvar_tb2_num_cfgs() { return kNumTb; }

void heatvar_tb2_cfg(int cfg, int* tx, int* ty, int* zc, int* d, int* steps)
{
    const TbCfg t = kTbCfgs[cfg];
    *tx = t.tx;
    *ty = t.ty;
    *zc = t.zc;
    *d = t.d;
    *steps = t.steps;
}

bool heatvar_tb2_supported(const Geo& geo, int mode)
{
    // 3-D only; in-plane byte offsets (ghost planes included) must fit 32 bits
    const long bytes = mode == SCAFES_HEATVAR_F64 ? 8 : 4;
    return geo.dim == 3 && (geo.base + geo.e01) * bytes < kOobLimit;
}

// The configuration's number of steps u -> uo (and Kahan compensation cin -> co).
int heatvar_tb2_launch(Ctx* c, const Geo& geo, const scafes_geom* g, int mode, const void* u, void* uo,
                       const void* cin, void* co, const void* kap, double tau, int border_mode, double border_value,
                       int cfg, void* const* mid)
{
    if (cfg < 0 || cfg >= kNumTb) return fail(SCAFES_ERR_ARG, "bad multi-step tile configuration");
    if (!heatvar_tb2_supported(geo, mode)) return fail(SCAFES_ERR_UNSUPPORTED, "multi-step sweep needs a 3-D grid");
    VarCoef cf;
    double hh[3];
    for (int p = 0; p < 3; ++p) hh[p] = g->h[p] * g->h[p];
    cf.tau = tau;
    cf.hh0 = hh[0];
    cf.hh1 = hh[1];
    cf.hh2 = hh[2];
    cf.c0 = (float)(0.5 * tau / hh[0]);
    cf.c1 = (float)(0.5 * tau / hh[1]);
    cf.c2 = (float)(0.5 * tau / hh[2]);
    cf.border_value = border_value;
    cf.border_mode = border_mode;
    if (mid) {
        if (mode == SCAFES_HEATVAR_F64)
            return tb_dispatch_mo<double, 0>(c, geo, u, nullptr, kap, uo, nullptr, cf, cfg, mid);
        if (mode == SCAFES_HEATVAR_F32)
            return tb_dispatch_mo<float, 1>(c, geo, u, nullptr, kap, uo, nullptr, cf, cfg, mid);
        return tb_dispatch_mo<float, 2>(c, geo, u, cin, kap, uo, co, cf, cfg, mid);
    }
    if (mode == SCAFES_HEATVAR_F64) return tb_dispatch<double, 0>(c, geo, u, nullptr, kap, uo, nullptr, cf, cfg);
    if (mode == SCAFES_HEATVAR_F32) return tb_dispatch<float, 1>(c, geo, u, nullptr, kap, uo, nullptr, cf, cfg);
    return tb_dispatch<float, 2>(c, geo, u, cin, kap, uo, co, cf, cfg);
}

}  // namespace scafes
