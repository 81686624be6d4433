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
// step never updates.  One adjoint step is one out-of-place sweep
// (adj_fused below).  Forward steps use the Yee kernels with the ce field:
// two steps per sweep (yee_tb2.hip) where two remain, else the fused
// one-step sweep (yee.hip); their per-value arithmetic is that of the
// half steps above.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "scafes_common.hpp"

namespace scafes {

namespace {

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

// One adjoint step in one out-of-place sweep: st = state n-1 (E, H), la =
// (lam_E', lam_H') -> lb = (lam_E, lam_H), grad += the step's term.  The
// layout follows the fused forward step (yee.hip): a TX x TY workgroup
// whose -x/-y edge threads only compute; per z plane k
//   level 1 at plane k:   H'(k) = H + s Hform(E) and mu(k) = lam_H' -
//                         Hform(w), w = ce * lam_E' (E and w planes in LDS
//                         with a +x/+y ring; plane k+1 from the z queue);
//                         lam_H(k) = mu(k) is stored;
//   level 2 at plane k-1: lam_E(k-1) = lam_E' - s Eform(mu) and the
//                         gradient term lam_E' . Eform(H') (H' and mu planes
//                         in LDS for the -x/-y neighbours, plane k-2 in
//                         registers).
// Every owned node gets every output value (lam_E = 0 off the E mask,
// lam_H = lam_H' off the H mask), so lb is a complete adjoint state.  The
// per-value arithmetic is that of the oracle's two half-step loops.
template <typename T, int TX, int TY, int D>
__global__ __launch_bounds__(TX* TY) void adj_fused(YeeDims d, T s, const T* __restrict__ st,
                                                     const T* __restrict__ la, T* __restrict__ lb,
                                                     const T* __restrict__ ce, const T* __restrict__ eps,
                                                     double* __restrict__ grad, int zc, int tiles_x, int tiles_y)
{
    static_assert(D >= 2 && D % 2 == 0, "prefetch depth must be even");
    constexpr int OX = TX - 1, OY = TY - 1;
    __shared__ T sE[3][2][TY + 1][TX + 1];  // E^{n-1} plane with the +x/+y ring
    __shared__ T sW[3][2][TY + 1][TX + 1];  // w = ce * lam_E' plane with the ring
    __shared__ T sP[3][2][TY][TX];          // H' plane
    __shared__ T sM[3][2][TY][TX];          // mu plane
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int lbk = xcd_remap(blockIdx.x, gridDim.x);
    const int per = tiles_x * tiles_y;
    const int tz = lbk / per, r = lbk - tz * per, tyi = r / tiles_x, txi = r - tyi * tiles_x;
    const int N = d.N;
    const long nf = d.nf;
    const int i = txi * OX - 1 + tx, j = tyi * OY - 1 + ty;
    const int ic = min(max(i, 0), N), jc = min(max(j, 0), N);
    const int ip = min(ic + 1, N), jp = min(jc + 1, N);
    const bool out = tx >= 1 && ty >= 1 && i <= N && j <= N;
    const bool rx = tx == TX - 1, ry = ty == TY - 1;
    const int k0 = tz * zc, k1 = min(k0 + zc, N + 1);
    // in-plane offsets fit 32 bits (ncell <= 1400); plane bases are uniform
    const int oc = ic + (int)d.M * jc, ox = ip + (int)d.M * jc, oy = ic + (int)d.M * jp;
    auto plane = [&](int kk) { return (long)min(max(kk, 0), N) * d.M2; };

    // z queue: slot q holds plane kb + q
    T qE[3][D], qH[3][D], qL[3][D], qG[3][D], qc[D], qe[D];
    T xE[2][D], xL[2][D], xc[D], yE[2][D], yL[2][D], yc[D];  // +x ring: y, z; +y ring: x, z
#pragma unroll
    for (int q = 0; q < D; ++q) {
        for (int a = 0; a < 3; ++a) qE[a][q] = qH[a][q] = qL[a][q] = qG[a][q] = T(0);
        for (int a = 0; a < 2; ++a) xE[a][q] = xL[a][q] = yE[a][q] = yL[a][q] = T(0);
        qc[q] = qe[q] = xc[q] = yc[q] = T(0);
    }
    auto fetch = [&](const int slot, const int kk) {
        if (kk <= k1 + 1) {
            const long p = plane(kk);
            const T* sp = st + p;
            const T* lp = la + p;
            const T* cp = ce + p;
            for (int a = 0; a < 3; ++a) {
                qE[a][slot] = sp[a * nf + oc];
                qL[a][slot] = lp[a * nf + oc];
            }
            qc[slot] = cp[oc];
            if (kk <= k1) {
                for (int a = 0; a < 3; ++a) {
                    qH[a][slot] = sp[(3 + a) * nf + oc];
                    qG[a][slot] = lp[(3 + a) * nf + oc];
                }
                qe[slot] = eps[p + oc];
                if (rx) {
                    xE[0][slot] = sp[nf + ox];
                    xE[1][slot] = sp[2 * nf + ox];
                    xL[0][slot] = lp[nf + ox];
                    xL[1][slot] = lp[2 * nf + ox];
                    xc[slot] = cp[ox];
                }
                if (ry) {
                    yE[0][slot] = sp[oy];
                    yE[1][slot] = sp[2 * nf + oy];
                    yL[0][slot] = lp[oy];
                    yL[1][slot] = lp[2 * nf + oy];
                    yc[slot] = cp[oy];
                }
            }
        }
    };
#pragma unroll
    for (int q = 0; q < D; ++q) fetch(q, k0 - 1 + q);

    // level-2 inputs of plane k-1 (and k-2 for the z differences)
    T pPx = T(0), pPy = T(0), pPz = T(0), pMx = T(0), pMy = T(0), pMz = T(0);
    T qPx = T(0), qPy = T(0), qMx = T(0), qMy = T(0);
    T pLx = T(0), pLy = T(0), pLz = T(0), pc = T(0), pe = T(1);
    for (int kb = k0 - 1; kb <= k1; kb += D) {
#pragma unroll
        for (int q = 0; q < D; ++q) {
            const int k = kb + q;
            if (k > k1) break;  // uniform over the workgroup
            const int buf = q & 1, pbf = buf ^ 1;
            const int q1 = (q + 1) % D;
            const T ex = qE[0][q], ey = qE[1][q], ez = qE[2][q];
            const T exu = qE[0][q1], eyu = qE[1][q1];
            const T hx = qH[0][q], hy = qH[1][q], hz = qH[2][q];
            const T lx = qL[0][q], ly = qL[1][q], lz = qL[2][q];
            const T gx = qG[0][q], gy = qG[1][q], gz = qG[2][q];
            const T c0 = qc[q], e0 = qe[q];
            const T wx = c0 * lx, wy = c0 * ly, wz = c0 * lz;
            const T wxu = qc[q1] * qL[0][q1], wyu = qc[q1] * qL[1][q1];
            sE[0][buf][ty][tx] = ex;
            sE[1][buf][ty][tx] = ey;
            sE[2][buf][ty][tx] = ez;
            sW[0][buf][ty][tx] = wx;
            sW[1][buf][ty][tx] = wy;
            sW[2][buf][ty][tx] = wz;
            if (rx) {
                sE[1][buf][ty][TX] = xE[0][q];
                sE[2][buf][ty][TX] = xE[1][q];
                sW[1][buf][ty][TX] = xc[q] * xL[0][q];
                sW[2][buf][ty][TX] = xc[q] * xL[1][q];
            }
            if (ry) {
                sE[0][buf][TY][tx] = yE[0][q];
                sE[2][buf][TY][tx] = yE[1][q];
                sW[0][buf][TY][tx] = yc[q] * yL[0][q];
                sW[2][buf][TY][tx] = yc[q] * yL[1][q];
            }
            fetch(q, k + D);
            __syncthreads();
            // ---- level 1 at plane k
            const bool ux = j < N && k < N, uy = i < N && k < N, uz = i < N && j < N;
            const T ey_xp = sE[1][buf][ty][tx + 1], ez_xp = sE[2][buf][ty][tx + 1];
            const T ex_yp = sE[0][buf][ty + 1][tx], ez_yp = sE[2][buf][ty + 1][tx];
            const T wy_xp = sW[1][buf][ty][tx + 1], wz_xp = sW[2][buf][ty][tx + 1];
            const T wx_yp = sW[0][buf][ty + 1][tx], wz_yp = sW[2][buf][ty + 1][tx];
            const T Px = ux ? hx + s * ((eyu - ey) - (ez_yp - ez)) : hx;
            const T Py = uy ? hy + s * ((ez_xp - ez) - (exu - ex)) : hy;
            const T Pz = uz ? hz + s * ((ex_yp - ex) - (ey_xp - ey)) : hz;
            const T Mx = ux ? gx - ((wyu - wy) - (wz_yp - wz)) : gx;
            const T My = uy ? gy - ((wz_xp - wz) - (wxu - wx)) : gy;
            const T Mz = uz ? gz - ((wx_yp - wx) - (wy_xp - wy)) : gz;
            sP[0][buf][ty][tx] = Px;
            sP[1][buf][ty][tx] = Py;
            sP[2][buf][ty][tx] = Pz;
            sM[0][buf][ty][tx] = Mx;
            sM[1][buf][ty][tx] = My;
            sM[2][buf][ty][tx] = Mz;
            if (out && k >= k0 && k < k1) {
                T* o = lb + (long)k * d.M2;
                o[3 * nf + oc] = Mx;
                o[4 * nf + oc] = My;
                o[5 * nf + oc] = Mz;
            }
            // ---- level 2 at plane k-1 (H' and mu planes written one iteration ago)
            const int km = k - 1;
            if (out && km >= k0) {
                const T pz_ym = sP[2][pbf][ty - 1][tx], px_ym = sP[0][pbf][ty - 1][tx];
                const T pz_xm = sP[2][pbf][ty][tx - 1], py_xm = sP[1][pbf][ty][tx - 1];
                const T mz_ym = sM[2][pbf][ty - 1][tx], mx_ym = sM[0][pbf][ty - 1][tx];
                const T mz_xm = sM[2][pbf][ty][tx - 1], my_xm = sM[1][pbf][ty][tx - 1];
                const bool vx = emask(N, 0, i, j, km), vy = emask(N, 1, i, j, km), vz = emask(N, 2, i, j, km);
                const T dx = (pPz - pz_ym) - (pPy - qPy);
                const T dy = (pPx - qPx) - (pPz - pz_xm);
                const T dz = (pPy - py_xm) - (pPx - px_ym);
                double acc = 0.0;
                T nx = T(0), ny = T(0), nz = T(0);
                if (vx) {
                    acc = acc + (double)pLx * (double)dx;
                    nx = pLx - s * ((pMz - mz_ym) - (pMy - qMy));
                }
                if (vy) {
                    acc = acc + (double)pLy * (double)dy;
                    ny = pLy - s * ((pMx - qMx) - (pMz - mz_xm));
                }
                if (vz) {
                    acc = acc + (double)pLz * (double)dz;
                    nz = pLz - s * ((pMy - my_xm) - (pMx - mx_ym));
                }
                const long pk = (long)km * d.M2;
                T* o = lb + pk;
                o[oc] = nx;
                o[nf + oc] = ny;
                o[2 * nf + oc] = nz;
                double* go = grad + pk;
                if (vx || vy || vz) go[oc] = go[oc] - ((double)pc / (double)pe) * acc;
            }
            qPx = pPx;
            qPy = pPy;
            qMx = pMx;
            qMy = pMy;
            pPx = Px;
            pPy = Py;
            pPz = Pz;
            pMx = Mx;
            pMy = My;
            pMz = Mz;
            pLx = lx;
            pLy = ly;
            pLz = lz;
            pc = c0;
            pe = e0;
        }
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

// FP32: 64 x 16 workgroups with 32-plane z chunks (measured at 384^3 on an
// MI355X: 1.27-1.48 ms per adjoint step against 1.51-1.62 for 64 x 8 with
// 32 or 64 planes, 64 x 4 and a four-plane queue; the 200-step gradient with
// 100 snapshots took 434 ms with 32-plane chunks, 441, 436 and 458 with 48,
// 64 and 128); FP64: 64 x 4 (its four LDS planes take twice the space) with
// 64-plane chunks.
template <typename T>
struct AdjTile {
    static constexpr int TX = 64, TY = sizeof(T) == 4 ? 16 : 4, D = 2, ZC = sizeof(T) == 4 ? 32 : 64;
};

template <typename T>
int adj_t(Ctx* c, const YeeDims& d, double s, const T* ce, const T* eps, const T* st, const T* la, T* lb,
          double* grad)
{
    typedef AdjTile<T> A;
    const int tiles_x = ceil_div(d.M, A::TX - 1), tiles_y = ceil_div(d.M, A::TY - 1);
    const long nb = (long)tiles_x * tiles_y * ceil_div(d.M, A::ZC);
    if (nb > 0x7fffffffL) return fail(SCAFES_ERR_ARG, "yee eps: grid too large");
    adj_fused<T, A::TX, A::TY, A::D><<<dim3((unsigned)nb), dim3(A::TX, A::TY), 0, c->stream>>>(
        d, (T)s, st, la, lb, ce, eps, grad, A::ZC, tiles_x, tiles_y);
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
    if (dtype == SCAFES_F64)
        return yee_fused_eps_launch<double>(c, d, s, (const double*)a, (double*)b, (const double*)ce);
    return yee_fused_eps_launch<float>(c, d, (float)s, (const float*)a, (float*)b, (const float*)ce);
}

bool yee_eps_two_step_supported(const YeeDims& d, int dtype)
{
    return yee_tb2_supported(d, dtype == SCAFES_F64 ? 8 : 4);
}

int yee_eps_step2(Ctx* c, const YeeDims& d, int dtype, double s, const void* ce, const void* a, void* b)
{
    const int es = dtype == SCAFES_F64 ? 8 : 4;
    const Tb2Plan pl = yee_tb2_plan(d, es, c->cu_count);
    if (dtype == SCAFES_F64)
        return yee_tb2_launch<double>(c, d, pl, s, (const double*)a, (double*)b, nullptr, nullptr,
                                      (const double*)ce);
    return yee_tb2_launch<float>(c, d, pl, (float)s, (const float*)a, (float*)b, nullptr, nullptr, (const float*)ce);
}

int yee_eps_step2_states(Ctx* c, const YeeDims& d, int dtype, double s, const void* ce, const void* a, void* mid,
                         void* b)
{
    const int es = dtype == SCAFES_F64 ? 8 : 4;
    const Tb2Plan pl = yee_tb2_plan(d, es, c->cu_count);
    if (dtype == SCAFES_F64)
        return yee_tb2_launch<double>(c, d, pl, s, (const double*)a, (double*)b, nullptr, nullptr,
                                      (const double*)ce, (double*)mid);
    return yee_tb2_launch<float>(c, d, pl, (float)s, (const float*)a, (float*)b, nullptr, nullptr, (const float*)ce,
                                 (float*)mid);
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
                    const void* la, void* lb, double* grad)
{
    if (dtype == SCAFES_F64)
        return adj_t<double>(c, d, s, (const double*)ce, (const double*)eps, (const double*)st, (const double*)la,
                             (double*)lb, grad);
    return adj_t<float>(c, d, s, (const float*)ce, (const float*)eps, (const float*)st, (const float*)la, (float*)lb,
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
    // pairs of steps in one sweep, an odd remainder by the two half-step sweeps
    const bool two = yee_eps_two_step_supported(d, dtype);
    void* a = f;
    void* b = work;
    for (int n = 0; rc == SCAFES_OK && n < nsteps;) {
        const int k = (two && nsteps - n >= 2) ? 2 : 1;
        rc = k == 2 ? yee_eps_step2(c, d, dtype, s, ce, a, b) : yee_eps_step(c, d, dtype, s, ce, a, b);
        std::swap(a, b);
        n += k;
    }
    if (rc == SCAFES_OK && a != f) rc = scafes_memcpy_d2d(reinterpret_cast<scafes_ctx*>(c), f, a, 6 * (size_t)d.nf * es);
    (void)hipFreeAsync(ce, c->stream);
    return rc;
}

}  // extern "C"
