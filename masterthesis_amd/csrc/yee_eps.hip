// yee_eps.hip -- Yee leapfrog with a relative permittivity field eps and the
// discrete adjoint of it (gradient of a final-time E misfit w.r.t. eps).
//
// Forward: the H half step is the vacuum one; the E half step is
//   D += s curl_b H,  E = D / eps   i.e.   E += ce * curl_b H,  ce = s / eps,
// one eps per node index m shared by Ex(m), Ey(m), Ez(m).  The reference
// scales D to E per node by its material coefficient ga = 1/(1 + sigma dt/eps0)
// (examples/EMFDTD/EMFDTD.hpp:351-356); 1/eps takes that role here.
//
// Adjoint of one step (H' = H + s A E, E' = E + C B H', A = -curl_f,
// B = curl_b, C = diag(ce) on the E entries the step updates; curl_f^T =
// curl_b on those entries):
//   mu_H  = lam_H' - Hform(ce * lam_E')    (H-step stencil, coefficient -1)
//   lam_E = lam_E' - s * Eform(mu_H)       (E-step stencil, coefficient -s)
//   grad(m) -= (ce/eps)(m) * sum_c lam_E'_c(m) * Eform_c(H')(m)
// with H' recomputed from the stored state and lam_E zero on entries the E
// step never updates.  Both halves are z-streaming column sweeps of the same
// shape as the forward half steps; the adjoint fields are updated in place
// (each is written only at the node that reads it there).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "scafes_common.hpp"

namespace scafes {

namespace {

constexpr int kETX = 64, kETY = 4, kEZC = 16;

__device__ inline long ye(const YeeDims& d, int i, int j, int k) { return i + d.M * j + d.M2 * k; }

// E entries the E half step updates (tangential E on the PEC walls and the
// padding entries at index N stay fixed).
__device__ inline bool emask(int N, int c, int i, int j, int k)
{
    if (c == 0) return i < N && j >= 1 && j < N && k >= 1 && k < N;
    if (c == 1) return j < N && i >= 1 && i < N && k >= 1 && k < N;
    return k < N && i >= 1 && i < N && j >= 1 && j < N;
}

template <typename T>
__global__ void ce_kernel(long nf, double s, const T* __restrict__ eps, T* __restrict__ ce)
{
    for (long m = blockIdx.x * (long)blockDim.x + threadIdx.x; m < nf; m += (long)gridDim.x * blockDim.x)
        ce[m] = (T)(s / (double)eps[m]);
}

// H-shaped sweep (forward differences in x, y, z).  Writes H' = H + s A E
// of the state fa into hout (every node; the entries the H step does not
// update are copied).  ADJ: also mu_H = lam_H - Hform(ce * lam_E) in place
// in lam_H on the updated entries.
template <typename T, bool ADJ>
__global__ __launch_bounds__(kETX* kETY) void h_sweep(YeeDims d, T s, const T* __restrict__ fa,
                                                       T* __restrict__ hout, const T* __restrict__ ce,
                                                       T* lam, int zc, int tiles_x, int tiles_y)
{
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per = tiles_x * tiles_y;
    const int tz = lb / per, r = lb - tz * per, tyi = r / tiles_x, txi = r - tyi * tiles_x;
    const int N = d.N;
    const int i = txi * kETX + (int)threadIdx.x, j = tyi * kETY + (int)threadIdx.y;
    if (i > N || j > N) return;
    const int ip = min(i + 1, N), jp = min(j + 1, N);
    const int k0 = tz * zc, k1 = min(k0 + zc, N + 1);
    const T* __restrict__ Ex = fa;
    const T* __restrict__ Ey = fa + d.nf;
    const T* __restrict__ Ez = fa + 2 * d.nf;
    const T* __restrict__ Hx = fa + 3 * d.nf;
    const T* __restrict__ Hy = fa + 4 * d.nf;
    const T* __restrict__ Hz = fa + 5 * d.nf;
    // lam: lam_E (x, y, z) then lam_H (x, y, z), nf values each
    const long nf = d.nf;
    long m = ye(d, i, j, k0);
    T ex = Ex[m], ey = Ey[m], ez = Ez[m];
    T wx = T(0), wy = T(0), wz = T(0);
    if constexpr (ADJ) {
        const T c = ce[m];
        wx = c * lam[m];
        wy = c * lam[nf + m];
        wz = c * lam[2 * nf + m];
    }
    for (int k = k0; k < k1; ++k) {
        const long mz = ye(d, i, j, min(k + 1, N)), mx = ye(d, ip, j, k), my = ye(d, i, jp, k);
        const T exn = Ex[mz], eyn = Ey[mz], ezn = Ez[mz];
        const T ez_xp = Ez[mx], ey_xp = Ey[mx];
        const T ez_yp = Ez[my], ex_yp = Ex[my];
        const bool ux = j < N && k < N, uy = i < N && k < N, uz = i < N && j < N;
        const T hx = Hx[m], hy = Hy[m], hz = Hz[m];
        hout[m] = ux ? hx + s * ((eyn - ey) - (ez_yp - ez)) : hx;
        hout[d.nf + m] = uy ? hy + s * ((ez_xp - ez) - (exn - ex)) : hy;
        hout[2 * d.nf + m] = uz ? hz + s * ((ex_yp - ex) - (ey_xp - ey)) : hz;
        T wxn = T(0), wyn = T(0), wzn = T(0);
        if constexpr (ADJ) {
            const T cz = ce[mz], cx = ce[mx], cy = ce[my];
            wxn = cz * lam[mz];
            wyn = cz * lam[nf + mz];
            wzn = cz * lam[2 * nf + mz];
            const T wz_xp = cx * lam[2 * nf + mx], wy_xp = cx * lam[nf + mx];
            const T wz_yp = cy * lam[2 * nf + my], wx_yp = cy * lam[my];
            if (ux) lam[3 * nf + m] = lam[3 * nf + m] - ((wyn - wy) - (wz_yp - wz));
            if (uy) lam[4 * nf + m] = lam[4 * nf + m] - ((wz_xp - wz) - (wxn - wx));
            if (uz) lam[5 * nf + m] = lam[5 * nf + m] - ((wx_yp - wx) - (wy_xp - wy));
        }
        ex = exn;
        ey = eyn;
        ez = ezn;
        wx = wxn;
        wy = wyn;
        wz = wzn;
        m = mz;
    }
}

// E-shaped sweep (backward differences).  Forward (ADJ false): eout =
// ein + ce * Eform(hin) on the updated entries, ein elsewhere.  Adjoint: e
// holds lam_E' and becomes lam_E = lam_E' - s * Eform(mu) on the updated
// entries; grad(m) -= (ce/eps)(m) * sum_c lam_E'_c Eform_c(hin) in FP64.
template <typename T, bool ADJ>
__global__ __launch_bounds__(kETX* kETY) void e_sweep(YeeDims d, T s, const T* __restrict__ ein,
                                                       const T* __restrict__ hin, T* __restrict__ e,
                                                       const T* __restrict__ ce, const T* __restrict__ eps,
                                                       const T* __restrict__ mu, double* __restrict__ grad, int zc,
                                                       int tiles_x, int tiles_y)
{
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int per = tiles_x * tiles_y;
    const int tz = lb / per, r = lb - tz * per, tyi = r / tiles_x, txi = r - tyi * tiles_x;
    const int N = d.N;
    const int i = txi * kETX + (int)threadIdx.x, j = tyi * kETY + (int)threadIdx.y;
    if (i > N || j > N) return;
    const int im = max(i - 1, 0), jm = max(j - 1, 0);
    const int k0 = tz * zc, k1 = min(k0 + zc, N + 1);
    const T* __restrict__ Px = hin;
    const T* __restrict__ Py = hin + d.nf;
    const T* __restrict__ Pz = hin + 2 * d.nf;
    const long nf = d.nf;
    const long m0 = ye(d, i, j, max(k0 - 1, 0));
    T pxm = Px[m0], pym = Py[m0];
    T mxm = T(0), mym = T(0);
    if constexpr (ADJ) {
        mxm = mu[m0];
        mym = mu[nf + m0];
    }
    for (int k = k0; k < k1; ++k) {
        const long m = ye(d, i, j, k), my = ye(d, i, jm, k), mx = ye(d, im, j, k);
        const T px = Px[m], py = Py[m], pz = Pz[m];
        const T pz_ym = Pz[my], px_ym = Px[my];
        const T pz_xm = Pz[mx], py_xm = Py[mx];
        const bool ux = emask(N, 0, i, j, k), uy = emask(N, 1, i, j, k), uz = emask(N, 2, i, j, k);
        const T dx = (pz - pz_ym) - (py - pym);
        const T dy = (px - pxm) - (pz - pz_xm);
        const T dz = (py - py_xm) - (px - px_ym);
        if constexpr (!ADJ) {
            const T c = ce[m];
            const T ex = ein[m], ey = ein[d.nf + m], ez = ein[2 * d.nf + m];
            e[m] = ux ? ex + c * dx : ex;
            e[d.nf + m] = uy ? ey + c * dy : ey;
            e[2 * d.nf + m] = uz ? ez + c * dz : ez;
        } else {
            const T mx0 = mu[m], my0 = mu[nf + m], mz0 = mu[2 * nf + m];
            const T mz_ym = mu[2 * nf + my], mx_ym = mu[my];
            const T mz_xm = mu[2 * nf + mx], my_xm = mu[nf + mx];
            double acc = 0.0;
            if (ux) {
                const T l = e[m];
                acc = acc + (double)l * (double)dx;
                e[m] = l - s * ((mz0 - mz_ym) - (my0 - mym));
            }
            if (uy) {
                const T l = e[d.nf + m];
                acc = acc + (double)l * (double)dy;
                e[d.nf + m] = l - s * ((mx0 - mxm) - (mz0 - mz_xm));
            }
            if (uz) {
                const T l = e[2 * d.nf + m];
                acc = acc + (double)l * (double)dz;
                e[2 * d.nf + m] = l - s * ((my0 - my_xm) - (mx0 - mx_ym));
            }
            if (ux || uy || uz) grad[m] = grad[m] - ((double)ce[m] / (double)eps[m]) * acc;
            mxm = mx0;
            mym = my0;
        }
        pxm = px;
        pym = py;
    }
}

// lam_E = E^N - target on the updated E entries (0 elsewhere), lam_H = 0;
// per-block FP64 partial sums of (E^N - target)^2 over those entries.
template <typename T>
__global__ __launch_bounds__(256) void lam_init(YeeDims d, const T* __restrict__ fN, const T* __restrict__ target,
                                                T* __restrict__ lam, double* __restrict__ partial)
{
    const long n3 = 3 * d.nf;
    double acc = 0.0;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n3; q += (long)gridDim.x * blockDim.x) {
        const int c = (int)(q / d.nf);
        const long m = q - (long)c * d.nf;
        const int k = (int)(m / d.M2);
        const long rr = m - (long)k * d.M2;
        const int j = (int)(rr / d.M), i = (int)(rr - (long)j * d.M);
        T v = T(0);
        if (emask(d.N, c, i, j, k)) {
            const double rd = (double)fN[q] - (double)target[q];
            v = (T)rd;
            acc += rd * rd;
        }
        lam[q] = v;
        lam[n3 + q] = T(0);
    }
    __shared__ double sh[256];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(256) void sum_partials(const double* __restrict__ partial, int n, double* __restrict__ out)
{
    __shared__ double sh[256];
    double acc = 0.0;
    for (int q = (int)threadIdx.x; q < n; q += 256) acc += partial[q];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = 0.5 * sh[0];
}

constexpr int kInitBlocks = 2048;

struct Sweep {
    dim3 grid, block;
    int tiles_x, tiles_y;
};

Sweep sweep_shape(const YeeDims& d)
{
    Sweep w;
    w.tiles_x = ceil_div(d.M, kETX);
    w.tiles_y = ceil_div(d.M, kETY);
    const long nb = (long)w.tiles_x * w.tiles_y * ceil_div(d.M, kEZC);
    w.grid = dim3((unsigned)nb);
    w.block = dim3(kETX, kETY);
    return w;
}

template <typename T>
int step_t(Ctx* c, const YeeDims& d, double s, const T* ce, const T* a, T* b)
{
    const Sweep w = sweep_shape(d);
    h_sweep<T, false><<<w.grid, w.block, 0, c->stream>>>(d, (T)s, a, b + 3 * d.nf, nullptr, nullptr, kEZC,
                                                           w.tiles_x, w.tiles_y);
    e_sweep<T, false><<<w.grid, w.block, 0, c->stream>>>(d, (T)s, a, b + 3 * d.nf, b, ce, nullptr, nullptr,
                                                           nullptr, kEZC, w.tiles_x, w.tiles_y);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

template <typename T>
int adj_t(Ctx* c, const YeeDims& d, double s, const T* ce, const T* eps, const T* st, T* lam, T* hp, double* grad)
{
    const Sweep w = sweep_shape(d);
    h_sweep<T, true><<<w.grid, w.block, 0, c->stream>>>(d, (T)s, st, hp, ce, lam, kEZC, w.tiles_x, w.tiles_y);
    e_sweep<T, true><<<w.grid, w.block, 0, c->stream>>>(d, (T)s, nullptr, hp, lam, ce, eps, lam + 3 * d.nf, grad,
                                                          kEZC, w.tiles_x, w.tiles_y);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

}  // namespace

bool yee_dims(int ncell, YeeDims* d)
{
    if (ncell < 2 || ncell > 1400) return false;
    d->N = ncell;
    d->M = ncell + 1;
    d->M2 = d->M * d->M;
    d->nf = d->M2 * d->M;
    return true;
}

long yee_eps_partials() { return kInitBlocks; }

int yee_eps_coef(Ctx* c, const YeeDims& d, int dtype, double s, const void* eps, void* ce)
{
    const int blocks = (int)std::min<long>((d.nf + 255) / 256, 65536);
    if (dtype == SCAFES_F64)
        ce_kernel<double><<<blocks, 256, 0, c->stream>>>(d.nf, s, (const double*)eps, (double*)ce);
    else
        ce_kernel<float><<<blocks, 256, 0, c->stream>>>(d.nf, s, (const float*)eps, (float*)ce);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

int yee_eps_step(Ctx* c, const YeeDims& d, int dtype, double s, const void* ce, const void* a, void* b)
{
    if (dtype == SCAFES_F64) return step_t<double>(c, d, s, (const double*)ce, (const double*)a, (double*)b);
    return step_t<float>(c, d, s, (const float*)ce, (const float*)a, (float*)b);
}

int yee_eps_lam_init(Ctx* c, const YeeDims& d, int dtype, const void* fN, const void* target, void* lam,
                     double* partial, double* J)
{
    if (dtype == SCAFES_F64)
        lam_init<double><<<kInitBlocks, 256, 0, c->stream>>>(d, (const double*)fN, (const double*)target,
                                                             (double*)lam, partial);
    else
        lam_init<float><<<kInitBlocks, 256, 0, c->stream>>>(d, (const float*)fN, (const float*)target,
                                                            (float*)lam, partial);
    sum_partials<<<1, 256, 0, c->stream>>>(partial, kInitBlocks, J);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

int yee_eps_adjoint(Ctx* c, const YeeDims& d, int dtype, double s, const void* ce, const void* eps, const void* st,
                    void* lam, void* hp, double* grad)
{
    if (dtype == SCAFES_F64)
        return adj_t<double>(c, d, s, (const double*)ce, (const double*)eps, (const double*)st, (double*)lam,
                             (double*)hp, grad);
    return adj_t<float>(c, d, s, (const float*)ce, (const float*)eps, (const float*)st, (float*)lam, (float*)hp,
                        grad);
}

}  // namespace scafes

using namespace scafes;

extern "C" {

int scafes_yee_eps_run(scafes_ctx* ctx, int ncell, int dtype, double s, const void* eps, void* f, void* work,
                       int nsteps)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !eps || !f || !work) return fail(SCAFES_ERR_ARG, "yee eps: null pointer");
    if (work == f) return fail(SCAFES_ERR_ARG, "yee eps: work buffer must differ from the state");
    YeeDims d;
    if (!yee_dims(ncell, &d) || nsteps < 0) return fail(SCAFES_ERR_ARG, "yee eps: need 2 <= ncell <= 1400, nsteps >= 0");
    if (!(s > 0.0) || s > 1.0 / 1.7320508075688772)
        return fail(SCAFES_ERR_ARG, "yee eps: Courant number must satisfy 0 < s <= 1/sqrt(3)");
    if (dtype != SCAFES_F32 && dtype != SCAFES_F64) return fail(SCAFES_ERR_ARG, "yee eps: bad dtype");
    const size_t es = dtype == SCAFES_F64 ? 8 : 4;
    void* ce = nullptr;
    SCAFES_HIP(hipMallocAsync(&ce, (size_t)d.nf * es, c->stream));
    int rc = yee_eps_coef(c, d, dtype, s, eps, ce);
    void* a = f;
    void* b = work;
    for (int n = 0; rc == SCAFES_OK && n < nsteps; ++n) {
        rc = yee_eps_step(c, d, dtype, s, ce, a, b);
        std::swap(a, b);
    }
    if (rc == SCAFES_OK && a != f) {
        hipError_t e = hipMemcpyAsync(f, a, 6 * (size_t)d.nf * es, hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess) rc = check_hip(e, "yee eps copy");
    }
    (void)hipFreeAsync(ce, c->stream);
    return rc;
}

}  // extern "C"
