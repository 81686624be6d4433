// yee_tb2.hip -- two Yee leapfrog steps per HBM sweep (temporal blocking),
// out of place (fa -> fb), with the per-value arithmetic of yee_h_kernel /
// yee_e_kernel in yee.hip, so the result is bit-identical to two fused or
// four half-step sweeps.
//
// Levels: A  H1 = H^{n+1/2} from E0 = E^n   (forward differences)
//         B  E1 = E^{n+1}   from H1         (backward differences)
//         C  H2 = H^{n+3/2} from E1         (forward differences)
//         D  E2 = E^{n+2}   from H2         (backward differences)
// A TX x TY workgroup covers nodes [a-2, a+TX-2) in x (likewise y); level A
// runs on all of them (E0 comes with one more node on the +x/+y side), B on
// positions with tx, ty >= 1, C on 1 <= tx, ty <= TX-2, D (and the stored
// E2, H2) on 2 <= tx, ty <= TX-2: a (TX-3) x (TY-3) output tile.
// Along z one iteration kk computes A at plane kk, B at kk-1, C at kk-2 and
// D at kk-3; each level reads the in-plane neighbours of the level below
// from an LDS plane written one iteration earlier, so one barrier per plane
// suffices with double-buffered planes.  Per two steps HBM moves the six
// fields once in and once out (plus the rings, mostly L2 hits).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "scafes_common.hpp"

namespace scafes {

namespace {

using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr int kRsrcWord3 = 0x00020000;  // gfx9 family: 32-bit data format, raw (stride 0) buffer

__device__ inline Rsrc rsrc(const char* p, uint32_t range)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(p), (short)0, (int)range, kRsrcWord3);
}

template <typename T>
__device__ inline T bld(Rsrc r, uint32_t off)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}

template <typename T>
__device__ inline void bst(T v, Rsrc r, uint32_t off)
{
    if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)off, 0, 0);
    } else {
        typedef int v2i __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2i, v), r, (int)off, 0, 0);
    }
}

template <typename T>
struct V3 {
    T x, y, z;
};

// H half step at one node (yee_h_kernel's expressions and guards).
template <typename T>
__device__ inline V3<T> h_update(int N, int i, int j, int k, T s, V3<T> h, V3<T> e, T exu, T eyu, T ey_xp, T ez_xp,
                                 T ex_yp, T ez_yp)
{
    V3<T> r = h;
    if (j < N && k < N) r.x = h.x + s * ((eyu - e.y) - (ez_yp - e.z));
    if (i < N && k < N) r.y = h.y + s * ((ez_xp - e.z) - (exu - e.x));
    if (i < N && j < N) r.z = h.z + s * ((ex_yp - e.x) - (ey_xp - e.y));
    return r;
}

// E half step at one node (yee_e_kernel's expressions and guards).
template <typename T>
__device__ inline V3<T> e_update(int N, int i, int j, int k, T s, V3<T> e, V3<T> h, T hxm, T hym, T hz_ym, T hx_ym,
                                 T hz_xm, T hy_xm)
{
    V3<T> r = e;
    const bool in_k = k >= 1 && k < N, in_j = j >= 1 && j < N, in_i = i >= 1 && i < N;
    if (i < N && in_j && in_k) r.x = e.x + s * ((h.z - hz_ym) - (h.y - hym));
    if (j < N && in_i && in_k) r.y = e.y + s * ((h.x - hxm) - (h.z - hz_xm));
    if (k < N && in_i && in_j) r.z = e.z + s * ((h.y - hy_xm) - (h.x - hx_ym));
    return r;
}

template <typename T, int TX, int TY>
struct Tb2Lds {
    T e0[3][2][TY + 1][TX + 1];  // E^n plane with the +x / +y ring
    T h1[3][2][TY][TX];
    T e1[3][2][TY][TX];
    T h2[3][2][TY][TX];
};

// EPS: the E levels use the per-node coefficient ce = s / eps (one value
// per node index, queued along z with the fields) in place of s.
// MO: the state after one step (E1, H1) is written to fm too, at the nodes
// whose E2, H2 the sweep writes (every node of the grid over all tiles), as a
// one-step sweep would write it.
template <typename T, int TX, int TY, int D, bool EPS, bool MO = false>
__global__ __launch_bounds__(TX* TY) void yee_tb2_kernel(YeeDims d, T s, const T* __restrict__ fa, T* __restrict__ fb,
                                                          const T* __restrict__ ce, int zc, int tiles_x, int tiles_y,
                                                          double* __restrict__ ep0, double* __restrict__ ep1,
                                                          T* __restrict__ fm)
{
    static_assert(D >= 2 && D % 2 == 0, "prefetch depth must be even");
    static_assert((TX * TY) % 64 == 0, "whole waves");
    constexpr int OX = TX - 3, OY = TY - 3;
    __shared__ Tb2Lds<T, TX, TY> L;
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per = tiles_x * tiles_y;
    const int tz = lb / per, r = lb - tz * per, tyi = r / tiles_x, txi = r - tyi * tiles_x;
    const int N = d.N;
    const int i = txi * OX - 2 + tx, j = tyi * OY - 2 + ty;
    const int ic = min(max(i, 0), N), jc = min(max(j, 0), N);
    const int ip = min(ic + 1, N), jp = min(jc + 1, N);
    const bool rx = tx == TX - 1, ry = ty == TY - 1;  // +x / +y ring duty for E^n
    // neighbour slots for the levels that read one node to the -x/-y or
    // +x/+y side: clamped so every LDS read stays in the tile (values read
    // through a clamped slot only feed positions that are never stored)
    const int txm = max(tx - 1, 0), tym = max(ty - 1, 0), txp = min(tx + 1, TX - 1), typ = min(ty + 1, TY - 1);
    const bool out = tx >= 2 && tx <= TX - 2 && ty >= 2 && ty <= TY - 2 && i <= N && j <= N;
    const int k0 = tz * zc, k1 = min(k0 + zc, N + 1);
    constexpr uint32_t B = sizeof(T);
    const uint32_t oc = (uint32_t)(ic + d.M * jc) * B;
    const uint32_t ox = (uint32_t)(ip + d.M * jc) * B;
    const uint32_t oy = (uint32_t)(ic + d.M * jp) * B;
    const uint32_t range = (uint32_t)(d.M2 * B);
    const size_t fs = (size_t)d.nf * B, pb = (size_t)d.M2 * B;
    const char* a8 = (const char*)fa;
    char* b8 = (char*)fb;
    const char* c8 = (const char*)ce;
    auto plane = [&](int kk) { return (size_t)min(max(kk, 0), N) * pb; };

    T qex[D], qey[D], qez[D], qhx[D], qhy[D], qhz[D], rxy[D], rxz[D], ryx[D], ryz[D], qce[D];
#pragma unroll
    for (int q = 0; q < D; ++q) qex[q] = qey[q] = qez[q] = qhx[q] = qhy[q] = qhz[q] = rxy[q] = rxz[q] = ryx[q] =
        ryz[q] = qce[q] = T(0);
    const int klast = k1 + 3;  // last plane of E^n read
    // Every iteration issues the same loads and stores; unneeded ones use
    // offset `range` (out of range: loads read 0 without a memory access,
    // stores are dropped), so the compiler can wait for exactly the slot a
    // plane consumes instead of for every outstanding access.
    auto fetch = [&](const int slot, const int kk) {
        const char* p = a8 + plane(kk);
        const Rsrc rEx = rsrc(p, range), rEy = rsrc(p + fs, range), rEz = rsrc(p + 2 * fs, range);
        const bool in = kk <= klast;
        const uint32_t o0 = in ? oc : range, o1 = (in && rx) ? ox : range, o2 = (in && ry) ? oy : range;
        qex[slot] = bld<T>(rEx, o0);
        qey[slot] = bld<T>(rEy, o0);
        qez[slot] = bld<T>(rEz, o0);
        qhx[slot] = bld<T>(rsrc(p + 3 * fs, range), o0);
        qhy[slot] = bld<T>(rsrc(p + 4 * fs, range), o0);
        qhz[slot] = bld<T>(rsrc(p + 5 * fs, range), o0);
        rxy[slot] = bld<T>(rEy, o1);
        rxz[slot] = bld<T>(rEz, o1);
        ryx[slot] = bld<T>(rEx, o2);
        ryz[slot] = bld<T>(rEz, o2);
        if constexpr (EPS) qce[slot] = bld<T>(rsrc(c8 + plane(kk), range), o0);
    };
    const int kfirst = k0 - 2;
#pragma unroll
    for (int q = 0; q < D; ++q) fetch(q, kfirst + q);

    V3<T> e0m1 = {}, h1m1 = {}, h1m2 = {}, e1m1 = {}, e1m2 = {}, h2m1 = {};
    T h2m2x = T(0), h2m2y = T(0);
    T cm1 = s, cm2 = s, cm3 = s;  // E coefficient at planes kk-1, kk-2, kk-3
    double es0 = 0.0, es1 = 0.0;
    const int kend = k1 + 2;
    for (int kb = kfirst; kb <= kend; kb += D) {
#pragma unroll
        for (int q = 0; q < D; ++q) {
            const int kk = kb + q;
            if (kk > kend) break;  // uniform over the workgroup
            const int buf = q & 1, pbf = buf ^ 1;
            const int q1 = (q + 1) % D;
            const V3<T> e0 = {qex[q], qey[q], qez[q]};
            const V3<T> h0 = {qhx[q], qhy[q], qhz[q]};
            const T exu = qex[q1], eyu = qey[q1];
            const T c0 = qce[q];  // E coefficient at plane kk (slot q is refilled below)
            L.e0[0][buf][ty][tx] = e0.x;
            L.e0[1][buf][ty][tx] = e0.y;
            L.e0[2][buf][ty][tx] = e0.z;
            if (rx) {
                L.e0[1][buf][ty][TX] = rxy[q];
                L.e0[2][buf][ty][TX] = rxz[q];
            }
            if (ry) {
                L.e0[0][buf][TY][tx] = ryx[q];
                L.e0[2][buf][TY][tx] = ryz[q];
            }
            fetch(q, kk + D);
            __syncthreads();
            // ---- A: H1 at plane kk
            const V3<T> h1 = h_update<T>(N, i, j, kk, s, h0, e0, exu, eyu, L.e0[1][buf][ty][tx + 1],
                                         L.e0[2][buf][ty][tx + 1], L.e0[0][buf][ty + 1][tx], L.e0[2][buf][ty + 1][tx]);
            L.h1[0][buf][ty][tx] = h1.x;
            L.h1[1][buf][ty][tx] = h1.y;
            L.h1[2][buf][ty][tx] = h1.z;
            if constexpr (MO) {
                const size_t o = plane(kk);
                const uint32_t os = (out && kk >= k0 && kk < k1) ? oc : range;
                char* m8 = (char*)fm;
                bst<T>(h1.x, rsrc(m8 + o + 3 * fs, range), os);
                bst<T>(h1.y, rsrc(m8 + o + 4 * fs, range), os);
                bst<T>(h1.z, rsrc(m8 + o + 5 * fs, range), os);
            }
            if (out && kk >= k0 && kk < k1 && ep0) {
                es0 += 0.5 * ((double)e0.x * e0.x + (double)e0.y * e0.y + (double)e0.z * e0.z) +
                       0.5 * ((double)h0.x * h1.x + (double)h0.y * h1.y + (double)h0.z * h1.z);
            }
            // ---- B: E1 at plane kk-1 (H1 plane kk-1 from the previous iteration)
            const int kb1 = kk - 1;
            const V3<T> e1 = e_update<T>(N, i, j, kb1, cm1, e0m1, h1m1, h1m2.x, h1m2.y, L.h1[2][pbf][tym][tx],
                                         L.h1[0][pbf][tym][tx], L.h1[2][pbf][ty][txm], L.h1[1][pbf][ty][txm]);
            L.e1[0][buf][ty][tx] = e1.x;
            L.e1[1][buf][ty][tx] = e1.y;
            L.e1[2][buf][ty][tx] = e1.z;
            if constexpr (MO) {
                const size_t o = plane(kb1);
                const uint32_t os = (out && kb1 >= k0 && kb1 < k1) ? oc : range;
                char* m8 = (char*)fm;
                bst<T>(e1.x, rsrc(m8 + o, range), os);
                bst<T>(e1.y, rsrc(m8 + o + fs, range), os);
                bst<T>(e1.z, rsrc(m8 + o + 2 * fs, range), os);
            }
            // ---- C: H2 at plane kk-2 (E1 plane kk-2 from the previous iteration, E1(kk-1) in registers)
            const int kc = kk - 2;
            const V3<T> h2 = h_update<T>(N, i, j, kc, s, h1m2, e1m1, e1.x, e1.y, L.e1[1][pbf][ty][txp],
                                         L.e1[2][pbf][ty][txp], L.e1[0][pbf][typ][tx], L.e1[2][pbf][typ][tx]);
            L.h2[0][buf][ty][tx] = h2.x;
            L.h2[1][buf][ty][tx] = h2.y;
            L.h2[2][buf][ty][tx] = h2.z;
            {
                const bool st = out && kc >= k0 && kc < k1;
                const size_t o = plane(kc);
                const uint32_t os = st ? oc : range;
                bst<T>(h2.x, rsrc(b8 + o + 3 * fs, range), os);
                bst<T>(h2.y, rsrc(b8 + o + 4 * fs, range), os);
                bst<T>(h2.z, rsrc(b8 + o + 5 * fs, range), os);
                if (st && ep1)
                    es1 += 0.5 * ((double)e1m1.x * e1m1.x + (double)e1m1.y * e1m1.y + (double)e1m1.z * e1m1.z) +
                           0.5 * ((double)h1m2.x * h2.x + (double)h1m2.y * h2.y + (double)h1m2.z * h2.z);
            }
            // ---- D: E2 at plane kk-3 (H2 plane kk-3 from the previous iteration)
            const int kd = kk - 3;
            {
                const bool st = out && kd >= k0 && kd < k1;
                V3<T> e2 = {};
                if (st)
                    e2 = e_update<T>(N, i, j, kd, cm3, e1m2, h2m1, h2m2x, h2m2y, L.h2[2][pbf][tym][tx],
                                     L.h2[0][pbf][tym][tx], L.h2[2][pbf][ty][txm], L.h2[1][pbf][ty][txm]);
                const size_t o = plane(kd);
                const uint32_t os = st ? oc : range;
                bst<T>(e2.x, rsrc(b8 + o, range), os);
                bst<T>(e2.y, rsrc(b8 + o + fs, range), os);
                bst<T>(e2.z, rsrc(b8 + o + 2 * fs, range), os);
            }
            // ---- shift the z queues
            h2m2x = h2m1.x;
            h2m2y = h2m1.y;
            h2m1 = h2;
            e1m2 = e1m1;
            e1m1 = e1;
            h1m2 = h1m1;
            h1m1 = h1;
            e0m1 = e0;
            if constexpr (EPS) {
                cm3 = cm2;
                cm2 = cm1;
                cm1 = c0;
            }
        }
    }
    if (ep0 || ep1) {
        // wave sums by lane shuffles, then the waves' sums in lane order
        constexpr int kWaves = (TX * TY) / 64;
        __shared__ double sh[2][kWaves];
        for (int w = 32; w > 0; w >>= 1) {
            es0 += __shfl_down(es0, w, 64);
            es1 += __shfl_down(es1, w, 64);
        }
        const int t = tx + TX * ty;
        if ((t & 63) == 0) {
            sh[0][t >> 6] = es0;
            sh[1][t >> 6] = es1;
        }
        __syncthreads();
        if (t == 0) {
            double a0 = 0.0, a1 = 0.0;
            for (int w = 0; w < kWaves; ++w) {
                a0 += sh[0][w];
                a1 += sh[1][w];
            }
            if (ep0) ep0[blockIdx.x] = a0;
            if (ep1) ep1[blockIdx.x] = a1;
        }
    }
}

// Tile per precision: FP32 64 x 16 with 4 planes in flight; FP64 64 x 12
// with 2 (its LDS planes take twice the space).  Measured on an MI355X at
// 384^3 FP32: 64x16 d4 and a whole z column per workgroup 144 Gcell-steps/s,
// 64x12 d2 with 64-plane chunks 106.
template <typename T>
struct TileOf {
    static constexpr int TX = 64, TY = sizeof(T) == 4 ? 16 : 12, D = sizeof(T) == 4 ? 4 : 2;
};

}  // namespace

bool yee_tb2_supported(const YeeDims& d, int elem_bytes) { return d.M2 * (long)elem_bytes < kOobLimit; }

// Work split: one workgroup per xy tile and z chunk; the chunk count is the
// smallest that keeps every CU busy for about the same time (a chunk costs
// its planes plus 5 pipeline fill planes; one workgroup per CU at a time).
Tb2Plan yee_tb2_plan(const YeeDims& d, int elem_bytes, int cu_count)
{
    Tb2Plan pl;
    const int tx = 64, ty = elem_bytes == 4 ? 16 : 12;
    pl.tiles_x = ceil_div(d.N + 1, tx - 3);
    pl.tiles_y = ceil_div(d.N + 1, ty - 3);
    const long nxy = (long)pl.tiles_x * pl.tiles_y;
    const long cus = std::max(cu_count, 1);
    long best = -1;
    pl.zc = d.N + 1;
    for (int nz = 1; nz <= 16; ++nz) {
        const int zc = ceil_div(d.N + 1, nz);
        const long rounds = (nxy * ceil_div(d.N + 1, zc) + cus - 1) / cus;
        const long cost = rounds * (zc + 5);
        if (best < 0 || cost < best) {
            best = cost;
            pl.zc = zc;
        }
    }
    pl.nb = nxy * ceil_div(d.N + 1, pl.zc);
    return pl;
}

template <typename T>
int yee_tb2_launch(Ctx* c, const YeeDims& d, const Tb2Plan& pl, T s, const T* a, T* b, double* ep0, double* ep1,
                   const T* ce, T* mid)
{
    typedef TileOf<T> Tl;
    if (pl.nb > 0x7fffffffL) return fail(SCAFES_ERR_ARG, "yee: grid too large");
    if (mid && (!ce || ep0 || ep1)) return fail(SCAFES_ERR_ARG, "yee: the mid state needs the eps sweep, no energy");
    const dim3 grid((unsigned)pl.nb), blk(Tl::TX, Tl::TY);
    if (mid)
        yee_tb2_kernel<T, Tl::TX, Tl::TY, Tl::D, true, true><<<grid, blk, 0, c->stream>>>(
            d, s, a, b, ce, pl.zc, pl.tiles_x, pl.tiles_y, nullptr, nullptr, mid);
    else if (ce)
        yee_tb2_kernel<T, Tl::TX, Tl::TY, Tl::D, true><<<grid, blk, 0, c->stream>>>(d, s, a, b, ce, pl.zc, pl.tiles_x,
                                                                                     pl.tiles_y, ep0, ep1, nullptr);
    else
        yee_tb2_kernel<T, Tl::TX, Tl::TY, Tl::D, false><<<grid, blk, 0, c->stream>>>(
            d, s, a, b, nullptr, pl.zc, pl.tiles_x, pl.tiles_y, ep0, ep1, nullptr);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

template int yee_tb2_launch<float>(Ctx*, const YeeDims&, const Tb2Plan&, float, const float*, float*, double*,
                                   double*, const float*, float*);
template int yee_tb2_launch<double>(Ctx*, const YeeDims&, const Tb2Plan&, double, const double*, double*, double*,
                                    double*, const double*, double*);

}  // namespace scafes
