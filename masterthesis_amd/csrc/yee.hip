// yee.hip -- Yee FDTD leapfrog in a PEC box (vacuum, normalised units).
//
// The reference's leapfrog mode runs updateInner (H from E), swaps, then
// updateInner2 (E from H) (src/ScaFES_Problem.hpp:2603-2619), with the
// curl stencil dest_old + constant*(dfdl - dfdr) of
// examples/EMFDTD/stencils/CompleteCPMLCurlStencil.hpp:38-80 (+cfln for H,
// -cfln for E, examples/EMFDTD/EMFDTD.hpp:207-212).  Without CPML and
// conductivity that is the plain Yee curl:
//   H += s * (forward differences of E),   E += s * (backward differences of H),
// tangential E zero on the walls.  Fields: Ex,Ey,Ez,Hx,Hy,Hz, each (N+1)^3
// nodes, x fastest; Ex(i,j,k) sits at (i+1/2, j, k), Hx at (i, j+1/2, k+1/2)...
//
// Two schedules.  With a second state buffer, one out-of-place sweep per
// step does both half steps (yee_fused_kernel, 48 B per cell-step).
// Without one, each half step is a z-streaming in-place sweep (a half step
// reads only the other field triple, so in place is race free): the centre column
// rides in a register queue, x+1 / y+1 (or x-1 / y-1) neighbours come
// from L1/L2 lines the neighbouring lanes and rows just fetched.  The H
// sweep optionally accumulates the discrete energy
//   W^n = 1/2 sum E^n.E^n + 1/2 sum H^{n-1/2}.H^{n+1/2}
// (conserved by the scheme) in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include <cstdio>
#include <mutex>

#include "scafes_common.hpp"

namespace scafes {

__device__ inline long yat(const YeeDims& d, int i, int j, int k) { return i + d.M * j + d.M2 * k; }

template <typename T, int TX, int TY>
__global__ __launch_bounds__(TX* TY) void yee_h_kernel(YeeDims d, T s, T* __restrict__ f, int zc, int tiles_x,
                                                        int tiles_y, double* __restrict__ epart)
{
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per = tiles_x * tiles_y;
    const int tz = lb / per, r = lb - tz * per, tyi = r / tiles_x, txi = r - tyi * tiles_x;
    const int N = d.N;
    const int i = txi * TX + tx, j = tyi * TY + ty;
    const bool active = i <= N && j <= N;
    const int ic = min(i, N), jc = min(j, N), ip = min(ic + 1, N), jp = min(jc + 1, N);
    const int k0 = tz * zc, k1 = min(k0 + zc, N + 1);
    const T* __restrict__ Ex = f;
    const T* __restrict__ Ey = f + d.nf;
    const T* __restrict__ Ez = f + 2 * d.nf;
    T* __restrict__ Hx = f + 3 * d.nf;
    T* __restrict__ Hy = f + 4 * d.nf;
    T* __restrict__ Hz = f + 5 * d.nf;
    double esum = 0.0;
    long m = yat(d, ic, jc, k0);
    T ex = Ex[m], ey = Ey[m], ez = Ez[m];
    for (int k = k0; k < k1; ++k) {
        const long mp = yat(d, ic, jc, min(k + 1, N));
        const T exn = Ex[mp], eyn = Ey[mp], ezn = Ez[mp];
        const T ez_xp = Ez[yat(d, ip, jc, k)], ey_xp = Ey[yat(d, ip, jc, k)];
        const T ez_yp = Ez[yat(d, ic, jp, k)], ex_yp = Ex[yat(d, ic, jp, k)];
        if (active) {
            const T hx = Hx[m], hy = Hy[m], hz = Hz[m];
            T hxn = hx, hyn = hy, hzn = hz;
            if (j < N && k < N) hxn = hx + s * ((eyn - ey) - (ez_yp - ez));
            if (i < N && k < N) hyn = hy + s * ((ez_xp - ez) - (exn - ex));
            if (i < N && j < N) hzn = hz + s * ((ex_yp - ex) - (ey_xp - ey));
            if (j < N && k < N) Hx[m] = hxn;
            if (i < N && k < N) Hy[m] = hyn;
            if (i < N && j < N) Hz[m] = hzn;
            if (epart) {
                const double e = (double)ex * ex + (double)ey * ey + (double)ez * ez;
                const double h = (double)hx * hxn + (double)hy * hyn + (double)hz * hzn;
                esum += 0.5 * e + 0.5 * h;
            }
        }
        ex = exn;
        ey = eyn;
        ez = ezn;
        m = mp;
    }
    if (epart) {
        __shared__ double sh[TX * TY];
        const int t = tx + TX * ty;
        sh[t] = esum;
        __syncthreads();
        for (int w = (TX * TY) / 2; w > 0; w >>= 1) {
            if (t < w) sh[t] += sh[t + w];
            __syncthreads();
        }
        if (t == 0) epart[blockIdx.x] = sh[0];
    }
}

template <typename T, int TX, int TY>
__global__ __launch_bounds__(TX* TY) void yee_e_kernel(YeeDims d, T s, T* __restrict__ f, int zc, int tiles_x,
                                                        int tiles_y)
{
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per = tiles_x * tiles_y;
    const int tz = lb / per, r = lb - tz * per, tyi = r / tiles_x, txi = r - tyi * tiles_x;
    const int N = d.N;
    const int i = txi * TX + tx, j = tyi * TY + ty;
    const bool active = i <= N && j <= N;
    const int ic = min(i, N), jc = min(j, N), im = max(ic - 1, 0), jm = max(jc - 1, 0);
    const int k0 = tz * zc, k1 = min(k0 + zc, N + 1);
    T* __restrict__ Ex = f;
    T* __restrict__ Ey = f + d.nf;
    T* __restrict__ Ez = f + 2 * d.nf;
    const T* __restrict__ Hx = f + 3 * d.nf;
    const T* __restrict__ Hy = f + 4 * d.nf;
    const T* __restrict__ Hz = f + 5 * d.nf;
    // z-1 values of Hx, Hy at this column
    const long m0 = yat(d, ic, jc, max(k0 - 1, 0));
    T hxm = Hx[m0], hym = Hy[m0];
    for (int k = k0; k < k1; ++k) {
        const long m = yat(d, ic, jc, k);
        const T hx = Hx[m], hy = Hy[m], hz = Hz[m];
        const T hz_ym = Hz[yat(d, ic, jm, k)], hx_ym = Hx[yat(d, ic, jm, k)];
        const T hz_xm = Hz[yat(d, im, jc, k)], hy_xm = Hy[yat(d, im, jc, k)];
        if (active) {
            const bool in_k = k >= 1 && k < N, in_j = j >= 1 && j < N, in_i = i >= 1 && i < N;
            if (i < N && in_j && in_k) Ex[m] = Ex[m] + s * ((hz - hz_ym) - (hy - hym));
            if (j < N && in_i && in_k) Ey[m] = Ey[m] + s * ((hx - hxm) - (hz - hz_xm));
            if (k < N && in_i && in_j) Ez[m] = Ez[m] + s * ((hy - hy_xm) - (hx - hx_ym));
        }
        hxm = hx;
        hym = hy;
    }
}

// Buffer access to one z plane of one field (wave-uniform resource + 32-bit
// in-plane byte offset per thread; see heatvar_tb2.hip).
using YRsrc = __amdgpu_buffer_rsrc_t;
constexpr int kYRsrcWord3 = 0x00020000;

__device__ inline YRsrc yrsrc(const char* p, uint32_t range)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<char*>(p), (short)0, (int)range, kYRsrcWord3);
}

template <typename T>
__device__ inline T yld(YRsrc r, uint32_t off)
{
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
    } else {
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
    }
}

template <typename T>
__device__ inline void yst(T v, YRsrc r, uint32_t off)
{
    if constexpr (sizeof(T) == 4) {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), r, (int)off, 0, 0);
    } else {
        typedef int v2i __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2i, v), r, (int)off, 0, 0);
    }
}

// One full leapfrog step in one sweep, out of place (fa -> fb, both holding
// Ex,Ey,Ez,Hx,Hy,Hz).  Per z plane k: H^{n+1/2}(k) from E^n(k) (LDS plane
// with a one-node ring on the +x/+y side) and E^n(k+1) (register queue);
// then, one plane behind, E^{n+1}(k-1) from H^{n+1/2}(k-1) (LDS plane) and
// H^{n+1/2}(k-2) (registers).  Threads on the -x/-y edge of the TX x TY
// workgroup compute H for the neighbouring tile's last column/row (needed
// by the E differences) and store nothing; the (TX-1) x (TY-1) others own
// their nodes and write every field value there (updated or copied), so fb
// is a complete state.  Per cell and step HBM moves 6 reads + 6 writes.
// The arithmetic per value is that of yee_h_kernel / yee_e_kernel, so the
// result is bit-identical to the two-pass in-place step.  EPS: the E half
// step uses the per-node coefficient ce = s / eps (queued along z with the
// fields) in place of s.
template <typename T, int TX, int TY, int D, bool EPS = false>
__global__ __launch_bounds__(TX* TY) void yee_fused_kernel(YeeDims d, T s, const T* __restrict__ fa,
                                                            T* __restrict__ fb, int zc, int tiles_x, int tiles_y,
                                                            double* __restrict__ epart, const T* __restrict__ ce)
{
    static_assert(D >= 2 && D % 2 == 0, "prefetch depth must be even");
    constexpr int OX = TX - 1, OY = TY - 1;
    __shared__ T sEx[2][TY + 1][TX + 1], sEy[2][TY + 1][TX + 1], sEz[2][TY + 1][TX + 1];
    __shared__ T sHx[2][TY][TX], sHy[2][TY][TX], sHz[2][TY][TX];
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per = tiles_x * tiles_y;
    const int tz = lb / per, r = lb - tz * per, tyi = r / tiles_x, txi = r - tyi * tiles_x;
    const int N = d.N;
    const int i = txi * OX - 1 + tx, j = tyi * OY - 1 + ty;
    const int ic = min(max(i, 0), N), jc = min(max(j, 0), N);
    const int ip = min(ic + 1, N), jp = min(jc + 1, N);
    const bool out = tx >= 1 && ty >= 1 && i <= N && j <= N;
    const bool rx = tx == TX - 1, ry = ty == TY - 1;  // +x / +y ring duty
    const int k0 = tz * zc, k1 = min(k0 + zc, N + 1);
    constexpr uint32_t B = sizeof(T);
    const uint32_t oc = (uint32_t)(ic + d.M * jc) * B;
    const uint32_t ox = (uint32_t)(ip + d.M * jc) * B;
    const uint32_t oy = (uint32_t)(ic + d.M * jp) * B;
    const uint32_t range = (uint32_t)(d.M2 * B);
    const size_t fs = (size_t)d.nf * B, pb = (size_t)d.M2 * B;
    const char* a8 = (const char*)fa;
    char* b8 = (char*)fb;
    auto plane = [&](int kk) { return (size_t)min(max(kk, 0), N) * pb; };

    T qex[D], qey[D], qez[D], qhx[D], qhy[D], qhz[D], rxy[D], rxz[D], ryx[D], ryz[D], qce[D];
#pragma unroll
    for (int q = 0; q < D; ++q) qex[q] = qey[q] = qez[q] = qhx[q] = qhy[q] = qhz[q] = rxy[q] = rxz[q] = ryx[q] =
        ryz[q] = qce[q] = T(0);
    auto fetch = [&](const int slot, const int kk) {
        if (kk <= k1 + 1) {
            const char* p = a8 + plane(kk);
            if constexpr (EPS) qce[slot] = yld<T>(yrsrc((const char*)ce + plane(kk), range), oc);
            const YRsrc rEx = yrsrc(p, range), rEy = yrsrc(p + fs, range), rEz = yrsrc(p + 2 * fs, range);
            qex[slot] = yld<T>(rEx, oc);
            qey[slot] = yld<T>(rEy, oc);
            qez[slot] = yld<T>(rEz, oc);
            if (kk <= k1) {
                qhx[slot] = yld<T>(yrsrc(p + 3 * fs, range), oc);
                qhy[slot] = yld<T>(yrsrc(p + 4 * fs, range), oc);
                qhz[slot] = yld<T>(yrsrc(p + 5 * fs, range), oc);
                if (rx) {
                    rxy[slot] = yld<T>(rEy, ox);
                    rxz[slot] = yld<T>(rEz, ox);
                }
                if (ry) {
                    ryx[slot] = yld<T>(rEx, oy);
                    ryz[slot] = yld<T>(rEz, oy);
                }
            }
        }
    };
#pragma unroll
    for (int q = 0; q < D; ++q) fetch(q, k0 - 1 + q);

    T pex = T(0), pey = T(0), pez = T(0);  // E^n at plane k-1
    T phx = T(0), phy = T(0), phz = T(0);  // H^{n+1/2} at plane k-1
    T qqx = T(0), qqy = T(0);              // Hx, Hy^{n+1/2} at plane k-2
    T pc = s;                              // E coefficient at plane k-1
    double esum = 0.0;
    for (int kb = k0 - 1; kb <= k1; kb += D) {
#pragma unroll
        for (int q = 0; q < D; ++q) {
            const int k = kb + q;
            if (k > k1) break;  // uniform over the workgroup
            const int buf = q & 1;
            const int q1 = (q + 1) % D;
            const T ex = qex[q], ey = qey[q], ez = qez[q];
            const T exu = qex[q1], eyu = qey[q1];
            const T hx = qhx[q], hy = qhy[q], hz = qhz[q];
            const T c0 = qce[q];  // slot q is refilled below
            sEx[buf][ty][tx] = ex;
            sEy[buf][ty][tx] = ey;
            sEz[buf][ty][tx] = ez;
            if (rx) {
                sEy[buf][ty][TX] = rxy[q];
                sEz[buf][ty][TX] = rxz[q];
            }
            if (ry) {
                sEx[buf][TY][tx] = ryx[q];
                sEz[buf][TY][tx] = ryz[q];
            }
            fetch(q, k + D);
            __syncthreads();
            // ---- H^{n+1/2} at plane k (every thread position)
            const T ey_xp = sEy[buf][ty][tx + 1], ez_xp = sEz[buf][ty][tx + 1];
            const T ex_yp = sEx[buf][ty + 1][tx], ez_yp = sEz[buf][ty + 1][tx];
            T hxn = hx, hyn = hy, hzn = hz;
            if (j < N && k < N) hxn = hx + s * ((eyu - ey) - (ez_yp - ez));
            if (i < N && k < N) hyn = hy + s * ((ez_xp - ez) - (exu - ex));
            if (i < N && j < N) hzn = hz + s * ((ex_yp - ex) - (ey_xp - ey));
            sHx[buf][ty][tx] = hxn;
            sHy[buf][ty][tx] = hyn;
            sHz[buf][ty][tx] = hzn;
            if (out && k >= k0 && k < k1) {
                const size_t o = (size_t)k * pb;
                yst<T>(hxn, yrsrc(b8 + o + 3 * fs, range), oc);
                yst<T>(hyn, yrsrc(b8 + o + 4 * fs, range), oc);
                yst<T>(hzn, yrsrc(b8 + o + 5 * fs, range), oc);
                if (epart) {
                    const double e = (double)ex * ex + (double)ey * ey + (double)ez * ez;
                    const double h = (double)hx * hxn + (double)hy * hyn + (double)hz * hzn;
                    esum += 0.5 * e + 0.5 * h;
                }
            }
            // ---- E^{n+1} at plane k-1 from the H plane written one iteration ago
            const int km = k - 1;
            if (out && km >= k0) {
                const int pbf = buf ^ 1;
                const T hz_ym = sHz[pbf][ty - 1][tx], hx_ym = sHx[pbf][ty - 1][tx];
                const T hz_xm = sHz[pbf][ty][tx - 1], hy_xm = sHy[pbf][ty][tx - 1];
                const bool in_k = km >= 1 && km < N, in_j = j >= 1 && j < N, in_i = i >= 1 && i < N;
                T exn = pex, eyn = pey, ezn = pez;
                if (i < N && in_j && in_k) exn = pex + pc * ((phz - hz_ym) - (phy - qqy));
                if (j < N && in_i && in_k) eyn = pey + pc * ((phx - qqx) - (phz - hz_xm));
                if (km < N && in_i && in_j) ezn = pez + pc * ((phy - hy_xm) - (phx - hx_ym));
                const size_t o = (size_t)km * pb;
                yst<T>(exn, yrsrc(b8 + o, range), oc);
                yst<T>(eyn, yrsrc(b8 + o + fs, range), oc);
                yst<T>(ezn, yrsrc(b8 + o + 2 * fs, range), oc);
            }
            qqx = phx;
            qqy = phy;
            phx = hxn;
            phy = hyn;
            phz = hzn;
            pex = ex;
            pey = ey;
            pez = ez;
            if constexpr (EPS) pc = c0;
        }
    }
    if (epart) {
        __shared__ double sh[TX * TY];
        const int t = tx + TX * ty;
        sh[t] = esum;
        __syncthreads();
        for (int w = (TX * TY) / 2; w > 0; w >>= 1) {
            if (t < w) sh[t] += sh[t + w];
            __syncthreads();
        }
        if (t == 0) epart[blockIdx.x] = sh[0];
    }
}

__global__ __launch_bounds__(256) void energy_final(const double* partial, int n, double* out)
{
    __shared__ double sh[256];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) acc += partial[i];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

template <typename T>
__global__ void yee_pulse_kernel(YeeDims d, T* f, double sigma)
{
    const int N = d.N;
    const double c = 0.5 * N;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < 6 * d.nf; t += (long)gridDim.x * blockDim.x) {
        const int comp = (int)(t / d.nf);
        const long m = t - comp * d.nf;
        const int i = (int)(m % d.M), j = (int)((m / d.M) % d.M), k = (int)(m / d.M2);
        T v = T(0);
        if (comp == 2 && k < N && i >= 1 && i < N && j >= 1 && j < N) {
            const double dx = i - c, dy = j - c, dz = k + 0.5 - c;
            v = (T)exp(-(dx * dx + dy * dy + dz * dz) / (2.0 * sigma * sigma));
        }
        f[t] = v;
    }
}

constexpr int kYTX = 64, kYTY = 4, kYZC = 64;         // two-pass in-place tiles
constexpr int kFTX = 64, kFTY = 8, kFZC = 64, kFD = 2;  // fused out-of-place tile

// energy: host array of the records (allocating path: the partials and the
// device records come from the stream-ordered allocator and the records are
// copied to the host at the end).  part != nullptr: the no-allocation path:
// the partials live in part (part_slots doubles) and the records are written
// to the device array edev_out; nothing is allocated, copied or waited for.
template <typename T>
static int yee_run_t(Ctx* c, int N, double s, T* f, T* f2, int nsteps, int energy_every, double* energy,
                     int* result_in_second, double* part = nullptr, long part_slots = 0, double* edev_out = nullptr)
{
    YeeDims d;
    d.N = N;
    d.M = N + 1;
    d.M2 = d.M * d.M;
    d.nf = d.M2 * d.M;
    const bool fused = f2 != nullptr;
    if (fused && d.M2 * (long)sizeof(T) >= (1L << 32)) return fail(SCAFES_ERR_ARG, "yee: plane too large");
    const bool tb2 = fused && yee_tb2_supported(d, (int)sizeof(T));
    const int TX = fused ? kFTX : kYTX, TY = fused ? kFTY : kYTY, ZC = fused ? kFZC : kYZC;
    const int OX = fused ? TX - 1 : TX, OY = fused ? TY - 1 : TY;
    const int tiles_x = ceil_div(N + 1, OX), tiles_y = ceil_div(N + 1, OY), tiles_z = ceil_div(N + 1, ZC);
    const long nb = (long)tiles_x * tiles_y * tiles_z;
    if (nb > 0x7fffffffL) return fail(SCAFES_ERR_ARG, "yee: grid too large");
    const Tb2Plan plan = yee_tb2_plan(d, (int)sizeof(T), c->cu_count);
    const long nbp = std::max(nb, tb2 ? plan.nb : 0L);  // partial slots per step
    const int nrec = (energy_every > 0) ? (nsteps + energy_every - 1) / energy_every : 0;
    double* epart = nullptr;
    double* edev = nullptr;
    const bool noalloc = part != nullptr;
    if (nrec > 0 && noalloc) {
        if (2 * nbp > part_slots) {
            char buf[200];
            std::snprintf(buf, sizeof buf,
                          "yee noalloc: %d cells per axis need %ld energy partial slots, the scratch array holds %ld",
                          N, 2 * nbp, part_slots);
            return fail(SCAFES_ERR_ARG, buf);
        }
        epart = part;
        edev = edev_out;
    } else if (nrec > 0) {
        SCAFES_HIP(hipMallocAsync((void**)&epart, sizeof(double) * 2 * nbp, c->stream));
        SCAFES_HIP(hipMallocAsync((void**)&edev, sizeof(double) * nrec, c->stream));
    }
    auto rec = [&](int st) { return nrec > 0 && (st % energy_every) == 0; };
    const dim3 blk(TX, TY);
    int rc = SCAFES_OK;
    T* a = f;
    T* b = f2;
    int st = 0;
    if (tb2) {
        // two steps per sweep; an odd count ends with one fused step
        for (; st + 2 <= nsteps; st += 2) {
            const bool r0 = rec(st), r1 = rec(st + 1);
            rc = yee_tb2_launch<T>(c, d, plan, (T)s, a, b, r0 ? epart : nullptr, r1 ? epart + nbp : nullptr);
            if (rc) break;
            const long nbt = plan.nb;
            if (r0) energy_final<<<1, 256, 0, c->stream>>>(epart, (int)nbt, edev + st / energy_every);
            if (r1) energy_final<<<1, 256, 0, c->stream>>>(epart + nbp, (int)nbt, edev + (st + 1) / energy_every);
            std::swap(a, b);
        }
    }
    for (; rc == SCAFES_OK && st < nsteps; ++st) {
        const bool r = rec(st);
        if (fused) {
            yee_fused_kernel<T, kFTX, kFTY, kFD><<<dim3((unsigned)nb), blk, 0, c->stream>>>(
                d, (T)s, a, b, ZC, tiles_x, tiles_y, r ? epart : nullptr, nullptr);
            std::swap(a, b);
        } else {
            yee_h_kernel<T, kYTX, kYTY><<<dim3((unsigned)nb), blk, 0, c->stream>>>(d, (T)s, f, ZC, tiles_x,
                                                                                  tiles_y, r ? epart : nullptr);
        }
        if (r) energy_final<<<1, 256, 0, c->stream>>>(epart, (int)nb, edev + st / energy_every);
        if (!fused)
            yee_e_kernel<T, kYTX, kYTY><<<dim3((unsigned)nb), blk, 0, c->stream>>>(d, (T)s, f, ZC, tiles_x, tiles_y);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) rc = check_hip(e, "yee kernels");
    }
    // every out-of-place launch ends with a swap: the final state is in a
    if (result_in_second) *result_in_second = (fused && a == f2) ? 1 : 0;
    if (nrec > 0 && !noalloc) {
        if (rc == SCAFES_OK) {
            hipError_t e = hipMemcpyAsync(energy, edev, sizeof(double) * nrec, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) rc = check_hip(e, "yee energy copy");
        }
        (void)hipFreeAsync(epart, c->stream);
        (void)hipFreeAsync(edev, c->stream);
    }
    return rc;
}

template <typename T>
int yee_fused_eps_launch(Ctx* c, const YeeDims& d, T s, const T* a, T* b, const T* ce)
{
    if (d.M2 * (long)sizeof(T) >= (1L << 32)) return fail(SCAFES_ERR_ARG, "yee: plane too large");
    const int tiles_x = ceil_div(d.N + 1, kFTX - 1), tiles_y = ceil_div(d.N + 1, kFTY - 1);
    const long nb = (long)tiles_x * tiles_y * ceil_div(d.N + 1, kFZC);
    if (nb > 0x7fffffffL) return fail(SCAFES_ERR_ARG, "yee: grid too large");
    yee_fused_kernel<T, kFTX, kFTY, kFD, true><<<dim3((unsigned)nb), dim3(kFTX, kFTY), 0, c->stream>>>(
        d, s, a, b, kFZC, tiles_x, tiles_y, nullptr, ce);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

template int yee_fused_eps_launch<float>(Ctx*, const YeeDims&, float, const float*, float*, const float*);
template int yee_fused_eps_launch<double>(Ctx*, const YeeDims&, double, const double*, double*, const double*);

// ---------------------------------------------------------------- no allocation
// Energy partials of scafes_yee_run_noalloc: two slots per workgroup of a
// launch, 2^17 doubles (1 MiB) per GPU, enough for Yee boxes of up to about
// 1150 cells per axis.  Each GPU the library runs on holds its own copy
// until the process ends.  Host code never passes the array's name to a
// kernel (it is a host-side stand-in); it uses the device address looked up
// once per device by yee_noalloc_setup.
constexpr long kScratchSlots = 1L << 17;
static __device__ __align__(256) double scratch[kScratchSlots];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[kMaxDevices] = {};

int yee_noalloc_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return fail(SCAFES_ERR_ARG, "yee noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return SCAFES_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return check_hip(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[device] = (double*)p;
    return SCAFES_OK;
}

static double* yee_scratch(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return g_scratch_ptr[device];
}

}  // namespace scafes

using namespace scafes;

extern "C" {

int scafes_yee_run(scafes_ctx* ctx, int ncell, int dtype, double s, void* f, void* f2, int nsteps,
                   int energy_every, double* energy, int* result_in_second)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !f) return fail(SCAFES_ERR_ARG, "yee: null pointer");
    if (f2 == f) return fail(SCAFES_ERR_ARG, "yee: second state buffer must differ from the first");
    if (ncell < 2 || nsteps < 0) return fail(SCAFES_ERR_ARG, "yee: need ncell >= 2 and nsteps >= 0");
    if (!(s > 0.0) || s > 1.0 / 1.7320508075688772)
        return fail(SCAFES_ERR_ARG, "yee: Courant number must satisfy 0 < s <= 1/sqrt(3)");
    if (energy_every > 0 && !energy) return fail(SCAFES_ERR_ARG, "yee: energy output missing");
    if (dtype == SCAFES_F32)
        return yee_run_t<float>(c, ncell, s, (float*)f, (float*)f2, nsteps, energy_every, energy, result_in_second);
    if (dtype == SCAFES_F64)
        return yee_run_t<double>(c, ncell, s, (double*)f, (double*)f2, nsteps, energy_every, energy,
                                 result_in_second);
    return fail(SCAFES_ERR_ARG, "yee: bad dtype");
}

int scafes_yee_run_noalloc(scafes_ctx* ctx, int ncell, int dtype, double s, void* f, void* f2, int nsteps,
                           int energy_every, double* energy_dev, int* result_in_second)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !f) return fail(SCAFES_ERR_ARG, "yee noalloc: null pointer");
    if (f2 == f) return fail(SCAFES_ERR_ARG, "yee noalloc: second state buffer must differ from the first");
    if (ncell < 2 || nsteps < 0) return fail(SCAFES_ERR_ARG, "yee noalloc: need ncell >= 2 and nsteps >= 0");
    if (!(s > 0.0) || s > 1.0 / 1.7320508075688772)
        return fail(SCAFES_ERR_ARG, "yee noalloc: Courant number must satisfy 0 < s <= 1/sqrt(3)");
    if (energy_every > 0 && !energy_dev) return fail(SCAFES_ERR_ARG, "yee noalloc: energy output missing");
    double* part = yee_scratch(c->device);
    if (!part) return fail(SCAFES_ERR_ARG, "yee noalloc: scratch not set up for this device");
    if (dtype == SCAFES_F32)
        return yee_run_t<float>(c, ncell, s, (float*)f, (float*)f2, nsteps, energy_every, nullptr, result_in_second,
                                part, kScratchSlots, energy_dev);
    if (dtype == SCAFES_F64)
        return yee_run_t<double>(c, ncell, s, (double*)f, (double*)f2, nsteps, energy_every, nullptr,
                                 result_in_second, part, kScratchSlots, energy_dev);
    return fail(SCAFES_ERR_ARG, "yee noalloc: bad dtype");
}

int scafes_yee_init_pulse(scafes_ctx* ctx, int ncell, int dtype, void* f, double sigma)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !f || ncell < 2 || !(sigma > 0.0)) return fail(SCAFES_ERR_ARG, "yee init: bad argument");
    YeeDims d;
    d.N = ncell;
    d.M = ncell + 1;
    d.M2 = d.M * d.M;
    d.nf = d.M2 * d.M;
    const int blocks = (int)std::min<long>((6 * d.nf + 255) / 256, 65536);
    if (dtype == SCAFES_F32) yee_pulse_kernel<float><<<blocks, 256, 0, c->stream>>>(d, (float*)f, sigma);
    else if (dtype == SCAFES_F64) yee_pulse_kernel<double><<<blocks, 256, 0, c->stream>>>(d, (double*)f, sigma);
    else return fail(SCAFES_ERR_ARG, "yee init: bad dtype");
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

}  // extern "C"
