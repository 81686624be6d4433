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
// Mapping: one 32x32-thread workgroup per CU owns a whole problem at a
// time.  Thread (tx, ty) keeps a B x B block of nodes (B = ceil(n/32) <= 8
// for n <= 256) in registers for all steps (4 waves per SIMD leave 128
// registers per lane).  Per step it publishes its block's four edges to LDS
// (component-major, bank-conflict free) and reads its neighbours' edges
// back: LDS traffic is one value per edge node instead of five per node,
// and HBM is touched only to load and store each problem once.  Border
// nodes take the update with a zero coefficient (fma(0, lap, u) = u).
// Each block size B is a separate non-inlined function, so each gets its
// own register allocation: inlined into one kernel, the B = 8 body's
// pressure would spill in every smaller body too.  Problems are taken in
// descending B in a snake order over workgroups (longest processing time
// first; no atomics and no counters to reset between graph replays).
//
// Gradient: with kappa scalar and tau fixed, the step operator
// A = I + tau kappa L (L symmetric on inner nodes with zero border) commutes
// with L, so the discrete adjoint sum collapses:
//   dJ/dkappa = sum_n <A^{N-n} r, tau L A^{n-1} u0> = N tau <r, L u^{N-1}>,
// r = u^N - target.  The last step already forms L u^{N-1} for every inner
// node, so the gradient costs one FP64 dot product, no reverse sweep.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "scafes_common.hpp"

namespace scafes {

constexpr int kTX = 32, kTY = 32;  // workgroup shape
constexpr int kNmax = 256;         // largest problem edge
constexpr int kBmax = kNmax / kTX;

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
__shared__ float eT[kBmax][kTY][kTX];
__shared__ float eB[kBmax][kTY][kTX];
__shared__ float eL[kBmax][kTY][kTX];
__shared__ float eR[kBmax][kTY][kTX];
__shared__ double red[2][kTX * kTY / 64];

template <int B>
__device__ __attribute__((noinline)) void run_problem(const BatchArgs& a, int p)
{
    const int tx = threadIdx.x, ty = threadIdx.y;
    const int n = a.n[p];
    const double h = 1.0 / (n - 1);
    const float ck = (float)((double)a.tau[p] * (double)a.kappa[p] / (h * h));
    float* __restrict__ u = a.u + a.off[p];
    const int ox = tx * B, oy = ty * B;
    // bit c of inx / iny: column ox+c / row oy+c is inner (1..n-2); vx / vy: is a node (< n)
    unsigned inx = 0, iny = 0, vx = 0, vy = 0;
#pragma unroll
    for (int c = 0; c < B; ++c) {
        inx |= (ox + c >= 1 && ox + c <= n - 2) ? (1u << c) : 0u;
        vx |= (ox + c < n) ? (1u << c) : 0u;
        iny |= (oy + c >= 1 && oy + c <= n - 2) ? (1u << c) : 0u;
        vy |= (oy + c < n) ? (1u << c) : 0u;
    }
    float s[B][B];
#pragma unroll
    for (int r = 0; r < B; ++r) {
        const float* __restrict__ row = u + (long)(oy + r) * n + ox;
#pragma unroll
        for (int c = 0; c < B; ++c) s[r][c] = ((vy >> r) & (vx >> c) & 1u) ? row[c] : 0.0f;
    }
    const bool want_grad = a.target != nullptr;
    // One explicit step on the register block, rows top-down in place:
    // prev[] keeps the old row above, `wold` the old left neighbour.  `last`
    // (the final step when a gradient is wanted) stores u^N to HBM and keeps
    // the inner nodes' L u^{N-1} terms (zero elsewhere) in the block.
    auto step = [&](const bool last) {
        // the masks live in VGPRs: scalar copies would be re-derived per use
        asm volatile("" : "+v"(inx), "+v"(iny));
#pragma unroll
        for (int c = 0; c < B; ++c) {
            eT[c][ty][tx] = s[0][c];
            eB[c][ty][tx] = s[B - 1][c];
        }
#pragma unroll
        for (int r = 0; r < B; ++r) {
            eL[r][ty][tx] = s[r][0];
            eR[r][ty][tx] = s[r][B - 1];
        }
        __syncthreads();
        float prev[B];
#pragma unroll
        for (int c = 0; c < B; ++c) prev[c] = ty > 0 ? eB[c][ty - 1][tx] : 0.0f;
#pragma unroll
        for (int r = 0; r < B; ++r) {
            float wold = tx > 0 ? eR[r][ty][tx - 1] : 0.0f;
            const float ckr = ((iny >> r) & 1u) ? ck : 0.0f;
            float* __restrict__ urow = u + (long)(oy + r) * n + ox;
#pragma unroll
            for (int c = 0; c < B; ++c) {
                const float uc = s[r][c];
                const float e = c == B - 1 ? (tx < kTX - 1 ? eL[r][ty][tx + 1] : 0.0f) : s[r][c + 1];
                const float sn = r == B - 1 ? (ty < kTY - 1 ? eT[c][ty + 1][tx] : 0.0f) : s[r + 1][c];
                const float lap = ((wold + e) - 2.0f * uc) + ((prev[c] + sn) - 2.0f * uc);
                const float k = ((inx >> c) & 1u) ? ckr : 0.0f;
                const float v = __builtin_fmaf(k, lap, uc);
                if (last && ((vy >> r) & (vx >> c) & 1u)) urow[c] = v;
                s[r][c] = last ? (k != 0.0f ? lap : 0.0f) : v;
                prev[c] = uc;
                wold = uc;
            }
        }
        __syncthreads();
    };
    for (int st = 0; st + 1 < a.nsteps; ++st) step(false);
    // The final step is the same code with a run-time flag: a second copy of
    // the unrolled step (compile-time flag) would raise register pressure.
    int lastg = want_grad ? 1 : 0;
    asm volatile("" : "+v"(lastg));
    step(lastg != 0);
    if (want_grad) {
        // the final state was stored by this thread; the barrier at the end
        // of the step made the stores visible to its own reloads
        const float* __restrict__ tgt = a.target + a.off[p];
        double gsum = 0.0, jsum = 0.0;
#pragma unroll
        for (int r = 0; r < B; ++r) {
            const long ro = (long)(oy + r) * n + ox;
#pragma unroll
            for (int c = 0; c < B; ++c) {
                if ((iny >> r) & (inx >> c) & 1u) {
                    const double res = (double)u[ro + c] - (double)tgt[ro + c];
                    gsum += res * (double)s[r][c];
                    jsum += res * res;
                }
            }
        }
        // wave sums by lane shuffles, then the 16 waves' sums in order
        for (int w = 32; w > 0; w >>= 1) {
            gsum += __shfl_down(gsum, w, 64);
            jsum += __shfl_down(jsum, w, 64);
        }
        const int t = tx + kTX * ty;
        if ((t & 63) == 0) {
            red[0][t >> 6] = gsum;
            red[1][t >> 6] = jsum;
        }
        __syncthreads();
        if (t == 0) {
            double g = 0.0, jj = 0.0;
            for (int w = 0; w < kTX * kTY / 64; ++w) {
                g += red[0][w];
                jj += red[1][w];
            }
            // dJ/dkappa = N tau <r, L u^{N-1}>,  L u = lap / h^2
            a.grad[p] = (double)a.nsteps * (double)a.tau[p] * g / (h * h);
            a.J[p] = 0.5 * jj;
        }
        __syncthreads();
    } else {
        asm volatile("" : "+v"(vx), "+v"(vy));
#pragma unroll
        for (int r = 0; r < B; ++r) {
            float* __restrict__ row = u + (long)(oy + r) * n + ox;
#pragma unroll
            for (int c = 0; c < B; ++c)
                if ((vy >> r) & (vx >> c) & 1u) row[c] = s[r][c];
        }
    }
}

// Per-thread block edge of a problem edge n: B = ceil(n/32).
__host__ __device__ inline int batch_b(int n) { return (n + kTX - 1) / kTX; }

__global__ __launch_bounds__(kTX* kTY) void batch2d_kernel(BatchArgs a)
{
    const int G = gridDim.x, w = blockIdx.x;
    for (int round = 0;; ++round) {
        const long idx = (long)round * G + ((round & 1) ? (G - 1 - w) : w);
        if (idx >= a.nprob) break;  // uniform over the workgroup
        const int p = a.order[idx];
        switch (batch_b(a.n[p])) {
        case 1: run_problem<1>(a, p); break;
        case 2: run_problem<2>(a, p); break;
        case 3: run_problem<3>(a, p); break;
        case 4: run_problem<4>(a, p); break;
        case 5: run_problem<5>(a, p); break;
        case 6: run_problem<6>(a, p); break;
        case 7: run_problem<7>(a, p); break;
        default: run_problem<8>(a, p); break;
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
    std::memset(&a, 0, sizeof a);  // the graph key compares the bytes
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
        // capture once per argument set; later calls with the same arguments
        // replay the instantiated graph (an optimisation loop re-running the
        // ensemble pays the capture and instantiation once)
        unsigned char key[sizeof(BatchArgs) + sizeof(int)];
        static_assert(sizeof key <= sizeof c->batch_key, "graph key too large");
        std::memcpy(key, &a, sizeof a);
        std::memcpy(key + sizeof a, &G, sizeof G);
        if (!c->batch_exec || std::memcmp(key, c->batch_key, sizeof key) != 0) {
            if (c->batch_exec) {
                SCAFES_HIP(hipStreamSynchronize(c->stream));
                (void)hipGraphExecDestroy(c->batch_exec);
                c->batch_exec = nullptr;
            }
            hipGraph_t graph;
            SCAFES_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            batch2d_kernel<<<G, dim3(kTX, kTY), 0, c->stream>>>(a);
            hipError_t le = hipGetLastError();
            hipError_t ce = hipStreamEndCapture(c->stream, &graph);
            if (le != hipSuccess) return check_hip(le, "batch2d launch");
            if (ce != hipSuccess) return check_hip(ce, "batch2d capture");
            hipGraphExec_t exec;
            hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ie != hipSuccess) return check_hip(ie, "batch2d graph instantiate");
            c->batch_exec = exec;
            std::memcpy(c->batch_key, key, sizeof key);
        }
        SCAFES_HIP(hipGraphLaunch(c->batch_exec, c->stream));
        return SCAFES_OK;
    }
    batch2d_kernel<<<G, dim3(kTX, kTY), 0, c->stream>>>(a);
    SCAFES_HIP(hipGetLastError());
    return SCAFES_OK;
}

}  // extern "C"
