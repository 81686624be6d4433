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
// Each half step is a z-streaming in-place sweep (a half step reads only
// the other field triple, so in place is race free): the centre column
// rides in a register queue, x+1 / y+1 (or x-1 / y-1) neighbours come
// from L1/L2 lines the neighbouring lanes and rows just fetched.  The H
// sweep optionally accumulates the discrete energy
//   W^n = 1/2 sum E^n.E^n + 1/2 sum H^{n-1/2}.H^{n+1/2}
// (conserved by the scheme) in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "scafes_common.hpp"

namespace scafes {

struct YeeDims {
    int N;       // cells per axis; fields hold N+1 nodes per axis
    long M, M2;  // N+1, (N+1)^2
    long nf;     // (N+1)^3
};

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

constexpr int kYTX = 64, kYTY = 4, kYZC = 64;

template <typename T>
static int yee_run_t(Ctx* c, int N, double s, T* f, int nsteps, int energy_every, double* energy)
{
    YeeDims d;
    d.N = N;
    d.M = N + 1;
    d.M2 = d.M * d.M;
    d.nf = d.M2 * d.M;
    const int tiles_x = ceil_div(N + 1, kYTX), tiles_y = ceil_div(N + 1, kYTY), tiles_z = ceil_div(N + 1, kYZC);
    const long nb = (long)tiles_x * tiles_y * tiles_z;
    if (nb > 0x7fffffffL) return fail(SCAFES_ERR_ARG, "yee: grid too large");
    const int nrec = (energy_every > 0) ? (nsteps + energy_every - 1) / energy_every : 0;
    double* epart = nullptr;
    double* edev = nullptr;
    if (nrec > 0) {
        SCAFES_HIP(hipMallocAsync((void**)&epart, sizeof(double) * nb, c->stream));
        SCAFES_HIP(hipMallocAsync((void**)&edev, sizeof(double) * nrec, c->stream));
    }
    const dim3 blk(kYTX, kYTY);
    int rc = SCAFES_OK;
    for (int st = 0; st < nsteps; ++st) {
        const bool rec = nrec > 0 && (st % energy_every) == 0;
        yee_h_kernel<T, kYTX, kYTY><<<dim3((unsigned)nb), blk, 0, c->stream>>>(d, (T)s, f, kYZC, tiles_x, tiles_y,
                                                                              rec ? epart : nullptr);
        if (rec) energy_final<<<1, 256, 0, c->stream>>>(epart, (int)nb, edev + st / energy_every);
        yee_e_kernel<T, kYTX, kYTY><<<dim3((unsigned)nb), blk, 0, c->stream>>>(d, (T)s, f, kYZC, tiles_x, tiles_y);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            rc = check_hip(e, "yee kernels");
            break;
        }
    }
    if (nrec > 0) {
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

}  // namespace scafes

using namespace scafes;

extern "C" {

int scafes_yee_run(scafes_ctx* ctx, int ncell, int dtype, double s, void* f, int nsteps, int energy_every,
                   double* energy)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !f) return fail(SCAFES_ERR_ARG, "yee: null pointer");
    if (ncell < 2 || nsteps < 0) return fail(SCAFES_ERR_ARG, "yee: need ncell >= 2 and nsteps >= 0");
    if (!(s > 0.0) || s > 1.0 / 1.7320508075688772)
        return fail(SCAFES_ERR_ARG, "yee: Courant number must satisfy 0 < s <= 1/sqrt(3)");
    if (energy_every > 0 && !energy) return fail(SCAFES_ERR_ARG, "yee: energy output missing");
    if (dtype == SCAFES_F32) return yee_run_t<float>(c, ncell, s, (float*)f, nsteps, energy_every, energy);
    if (dtype == SCAFES_F64) return yee_run_t<double>(c, ncell, s, (double*)f, nsteps, energy_every, energy);
    return fail(SCAFES_ERR_ARG, "yee: bad dtype");
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
