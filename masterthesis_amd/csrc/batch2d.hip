// batch2d.hip -- ensemble of independent 2-D heat problems in ONE persistent
// launch (the reference runs one problem per process; this batches a whole
// parameter ensemble on one GPU).
//
// Problem i: n_i x n_i nodes on the unit square (h = 1/(n_i-1)), zero
// Dirichlet border, scalar kappa_i and tau_i, explicit Euler with the
// 5-point stencil (the 2-D case of HeatEqnFDM::updateInner,
// tests/system/HeatEqnFDM/HeatEqnFDM.hpp:204-218, coefficient kappa_i):
//   u' = u + (tau kappa / h^2) ((W + E - 2u) + (N + S - 2u))   on inner nodes.
//
// Mapping: one 32x16-thread workgroup per CU owns a whole problem at a
// time.  Thread (tx, ty) keeps a BX x BY block of nodes (BX = ceil(n/32),
// BY = ceil(n/16), so BX*BY <= 128 for n <= 256) in registers for all steps
// (2 waves per SIMD leave 256 registers per lane).  Per step it publishes
// its block's four edges to LDS (component-major, bank-conflict free) and
// reads its neighbours' edges back: LDS traffic is one value per edge node
// instead of five per node, and HBM is touched only to load and store each
// problem once.  Problems are taken in descending BX*BY cost in a snake
// order over workgroups (longest processing time first; no atomics and no
// counters to reset between graph replays).
//
// Gradient: with kappa scalar and tau fixed, the step operator
// A = I + tau kappa L (L symmetric on inner nodes with zero border) commutes
// with L, so the discrete adjoint sum collapses:
//   dJ/dkappa = sum_n <A^{N-n} r, tau L A^{n-1} u0> = N tau <r, L u^{N-1}>,
// r = u^N - target.  The last step already forms L u^{N-1} for every inner
// node, so the gradient costs one FP64 dot product, no reverse sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
#include <vector>

#include "scafes_common.hpp"

namespace scafes {

constexpr int kTX = 32, kTY = 16;  // workgroup shape
constexpr int kNmax = 256;         // largest problem edge
constexpr int kBXmax = kNmax / kTX, kBYmax = kNmax / kTY;

struct BatchArgs {
    const int* n;
    const long* off;
    const float* kappa;
    const float* tau;
    const int* order;
    float* u;
    const float* target;
    double* J;
    double* grad;
    int nprob;
    int nsteps;
};

// Block edges published per step (LDS): top / bottom rows indexed
// [column][ty][tx], left / right columns [row][ty][tx].
__shared__ float eT[kBXmax][kTY][kTX];
__shared__ float eB[kBXmax][kTY][kTX];
__shared__ float eL[kBYmax][kTY][kTX];
__shared__ float eR[kBYmax][kTY][kTX];
__shared__ double red[2 * kTX * kTY];

template <int BX, int BY>
__device__ void run_problem(const BatchArgs& a, int p)
{
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int n = a.n[p];
    const double h = 1.0 / (n - 1);
    const float ck = (float)((double)a.tau[p] * (double)a.kappa[p] / (h * h));
    float* __restrict__ u = a.u + a.off[p];
    const int ox = tx * BX, oy = ty * BY;
    // bit c of inx / iny: column ox+c / row oy+c is inner (1..n-2); vx / vy: is a node (< n)
    unsigned inx = 0, iny = 0, vx = 0, vy = 0;
#pragma unroll
    for (int c = 0; c < BX; ++c) {
        inx |= (ox + c >= 1 && ox + c <= n - 2) ? (1u << c) : 0u;
        vx |= (ox + c < n) ? (1u << c) : 0u;
    }
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        iny |= (oy + r >= 1 && oy + r <= n - 2) ? (1u << r) : 0u;
        vy |= (oy + r < n) ? (1u << r) : 0u;
    }
    float s[BY][BX];
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        const float* __restrict__ row = u + (long)(oy + r) * n + ox;
#pragma unroll
        for (int c = 0; c < BX; ++c) s[r][c] = ((vy >> r) & (vx >> c) & 1u) ? row[c] : 0.0f;
    }
    double gsum = 0.0, jsum = 0.0;
    const bool want_grad = a.target != nullptr;
    const float* __restrict__ tgt = want_grad ? a.target + a.off[p] : nullptr;
    // One explicit step on the register block; GRAD also forms the gradient
    // and misfit terms of the final step.  Rows are updated top-down in
    // place: prev[] keeps the old row above, `wold` the old left neighbour.
    auto step = [&](auto grad_tag) {
        constexpr bool GRAD = decltype(grad_tag)::value;
        // Re-derive the per-node inner predicates from the bit masks every
        // step: hoisted out of the time loop they would pin one mask per
        // node in registers for the whole problem.
        asm volatile("" : "+v"(inx), "+v"(iny));
#pragma unroll
        for (int c = 0; c < BX; ++c) {
            eT[c][ty][tx] = s[0][c];
            eB[c][ty][tx] = s[BY - 1][c];
        }
#pragma unroll
        for (int r = 0; r < BY; ++r) {
            eL[r][ty][tx] = s[r][0];
            eR[r][ty][tx] = s[r][BX - 1];
        }
        __syncthreads();
        float prev[BX];
#pragma unroll
        for (int c = 0; c < BX; ++c) prev[c] = ty > 0 ? eB[c][ty - 1][tx] : 0.0f;
#pragma unroll
        for (int r = 0; r < BY; ++r) {
            float wold = tx > 0 ? eR[r][ty][tx - 1] : 0.0f;
            const bool rin = (iny >> r) & 1u;
#pragma unroll
            for (int c = 0; c < BX; ++c) {
                const float uc = s[r][c];
                const float e = c == BX - 1 ? (tx < kTX - 1 ? eL[r][ty][tx + 1] : 0.0f) : s[r][c + 1];
                const float sn = r == BY - 1 ? (ty < kTY - 1 ? eT[c][ty + 1][tx] : 0.0f) : s[r + 1][c];
                if (rin && ((inx >> c) & 1u)) {
                    const float lap = ((wold + e) - 2.0f * uc) + ((prev[c] + sn) - 2.0f * uc);
                    const float v = __builtin_fmaf(ck, lap, uc);
                    s[r][c] = v;
                    if constexpr (GRAD) {
                        const double res = (double)v - (double)tgt[(long)(oy + r) * n + ox + c];
                        gsum += res * (double)lap;
                        jsum += res * res;
                    }
                }
                prev[c] = uc;
                wold = uc;
            }
        }
        __syncthreads();
    };
    for (int st = 0; st + 1 < a.nsteps; ++st) step(std::false_type{});
    if (want_grad) step(std::true_type{});
    else step(std::false_type{});
    asm volatile("" : "+v"(vx), "+v"(vy));
#pragma unroll
    for (int r = 0; r < BY; ++r) {
        float* __restrict__ row = u + (long)(oy + r) * n + ox;
#pragma unroll
        for (int c = 0; c < BX; ++c)
            if ((vy >> r) & (vx >> c) & 1u) row[c] = s[r][c];
    }
    if (want_grad) {
        constexpr int NT = kTX * kTY;
        const int t = tx + kTX * ty;
        red[t] = gsum;
        red[NT + t] = jsum;
        __syncthreads();
        for (int w = NT / 2; w > 0; w >>= 1) {
            if (t < w) {
                red[t] += red[t + w];
                red[NT + t] += red[NT + t + w];
            }
            __syncthreads();
        }
        if (t == 0) {
            // dJ/dkappa = N tau <r, L u^{N-1}>,  L u = lap / h^2
            a.grad[p] = (double)a.nsteps * (double)a.tau[p] * red[0] / (h * h);
            a.J[p] = 0.5 * red[NT];
        }
        __syncthreads();
    }
}

// Per-thread block shape of a problem edge n: BX = ceil(n/32), BY = ceil(n/16).
__host__ __device__ inline int batch_bx(int n) { return (n + kTX - 1) / kTX; }
__host__ __device__ inline int batch_by(int n) { return (n + kTY - 1) / kTY; }

template <int BX>
__device__ void dispatch_by(const BatchArgs& a, int p, int by)
{
    // BY is 2*BX - 1 or 2*BX for every n (a single row block when n <= 16)
    if (by == 2 * BX) run_problem<BX, 2 * BX>(a, p);
    else if constexpr (BX > 1) run_problem<BX, 2 * BX - 1>(a, p);
    else run_problem<1, 1>(a, p);
}

__global__ __launch_bounds__(kTX* kTY) void batch2d_kernel(BatchArgs a)
{
    const int G = gridDim.x, w = blockIdx.x;
    for (int round = 0;; ++round) {
        const long idx = (long)round * G + ((round & 1) ? (G - 1 - w) : w);
        if (idx >= a.nprob) break;  // uniform over the workgroup
        const int p = a.order[idx];
        const int n = a.n[p];
        const int bx = batch_bx(n), by = batch_by(n);
        switch (bx) {
        case 1: dispatch_by<1>(a, p, by); break;
        case 2: dispatch_by<2>(a, p, by); break;
        case 3: dispatch_by<3>(a, p, by); break;
        case 4: dispatch_by<4>(a, p, by); break;
        case 5: dispatch_by<5>(a, p, by); break;
        case 6: dispatch_by<6>(a, p, by); break;
        case 7: dispatch_by<7>(a, p, by); break;
        default: dispatch_by<8>(a, p, by); break;
        }
    }
}

}  // namespace scafes

using namespace scafes;

extern "C" {

int scafes_batch2d_lds_limit(int* max_n)
{
    if (!max_n) return fail(SCAFES_ERR_ARG, "null pointer");
    *max_n = kNmax;
    return SCAFES_OK;
}

int scafes_batch2d_run(scafes_ctx* ctx, int nprob, const int* n, const long* off, const float* kappa,
                       const float* tau, const int* order, float* u, const float* target, double* J,
                       double* grad, int nsteps, int use_graph)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !n || !off || !kappa || !tau || !order || !u) return fail(SCAFES_ERR_ARG, "batch2d: null pointer");
    if (nprob < 0 || nsteps < 0) return fail(SCAFES_ERR_ARG, "batch2d: negative count");
    if (target && (!J || !grad)) return fail(SCAFES_ERR_ARG, "batch2d: gradient outputs missing");
    if (nprob == 0 || nsteps == 0) return SCAFES_OK;
    // sizes are device data; check them on the host once before launching
    std::vector<int> hn((size_t)nprob);
    SCAFES_HIP(hipMemcpyAsync(hn.data(), n, sizeof(int) * (size_t)nprob, hipMemcpyDeviceToHost, c->stream));
    SCAFES_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < nprob; ++i)
        if (hn[i] < 2 || hn[i] > kNmax) return fail(SCAFES_ERR_ARG, "batch2d: problem edge outside 2..256");
    BatchArgs a;
    a.n = n;
    a.off = off;
    a.kappa = kappa;
    a.tau = tau;
    a.order = order;
    a.u = u;
    a.target = target;
    a.J = J;
    a.grad = grad;
    a.nprob = nprob;
    a.nsteps = nsteps;
    const int G = std::max(1, std::min(c->cu_count, nprob));
    if (use_graph) {
        hipGraph_t graph;
        hipGraphExec_t exec;
        SCAFES_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        batch2d_kernel<<<G, dim3(kTX, kTY), 0, c->stream>>>(a);
        hipError_t le = hipGetLastError();
        hipError_t ce = hipStreamEndCapture(c->stream, &graph);
        if (le != hipSuccess) return check_hip(le, "batch2d launch");
        if (ce != hipSuccess) return check_hip(ce, "batch2d capture");
        hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            (void)hipGraphDestroy(graph);
            return check_hip(ie, "batch2d graph instantiate");
        }
        hipError_t ge = hipGraphLaunch(exec, c->stream);
        if (ge == hipSuccess) ge = hipStreamSynchronize(c->stream);
        (void)hipGraphExecDestroy(exec);
        (void)hipGraphDestroy(graph);
        if (ge != hipSuccess) return check_hip(ge, "batch2d graph launch");
        return SCAFES_OK;
    }
    batch2d_kernel<<<G, dim3(kTX, kTY), 0, c->stream>>>(a);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

}  // extern "C"
