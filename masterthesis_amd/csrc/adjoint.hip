// adjoint.hip -- gradient of a final-time misfit w.r.t. a coefficient field
// (the Pennes perfusion w, the variable-kappa heat's kappa, or the Yee
// permittivity eps) by the
// reverse-time discrete adjoint with binomial checkpointing.
//
// The reference differentiates with ADOL-C: one tape of the update phase
// replayed in forward vector mode w.r.t. a few geometry parameters
// (src/ScaFES_Problem.hpp:5841-5967 tape, :6632-6747 fov_forward).  A
// gradient w.r.t. a coefficient FIELD needs reverse mode: J = 0.5 sum_inner
// (T^N - target)^2, lam^N = T^N - target, then for n = N..1
//   grad += lam^n * dT^n/dw (needs T^{n-1}),  lam^{n-1} = (dT^n/dT^{n-1})^T lam^n.
// T^{n-1} is needed in reverse order; it is regenerated from at most
// `snaps` stored states with Griewank's binomial (treeverse) schedule:
// l steps and s snapshots are split so that the right part runs with s-1
// snapshots and the left part later with s, taking the right part as
// beta(s-1, r) = C(s-1+r, s-1) steps where r is the smallest repetition
// number with C(s+r, s) >= l.  Forward cost ~ r*N - C(s+1, r-1) steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "scafes_common.hpp"

namespace scafes {

// beta(s, r) = C(s + r, s), saturating.
static long beta(int s, int r)
{
    if (s < 0 || r < 0) return 0;
    double v = 1.0;
    for (int i = 1; i <= s; ++i) {
        v = v * (r + i) / i;
        if (v > 4e18) return (long)4e18;
    }
    return (long)(v + 0.5);
}

// Number of steps the right sub-range takes when l steps are reversed with
// s spare snapshots.
static long right_part(long l, int s)
{
    if (s <= 0 || l <= 1) return 0;
    int r = 0;
    while (beta(s, r) < l) ++r;
    return std::min(beta(s - 1, r), l - 1);
}

// lam^N = T^N - target on inner nodes (T^N = hi [+ lo]), or on every node
// (every_node), 0 elsewhere; the second adjoint buffer starts at zero (the
// gradient is cleared before).  partial[block] = the block's FP64 sum of
// lam^N squared (2 J), to be summed by reduce_partials.
template <typename T>
__global__ __launch_bounds__(256) void lam_init_kernel(Geo g, const T* __restrict__ t, const T* __restrict__ tlo,
                                                       const T* __restrict__ target, T* __restrict__ lam,
                                                       T* __restrict__ lam2, int every_node,
                                                       double* __restrict__ partial)
{
    __shared__ double sh[256];
    double acc = 0.0;
    // work items: 256-node chunks of x-rows (indices formed once per chunk)
    const int cpr = (g.nx + 255) / 256;
    const long items = (long)g.ny * g.nz * cpr;
    for (long w = blockIdx.x; w < items; w += gridDim.x) {
        const long r = w / cpr;
        const int i = (int)(w - r * cpr) * 256 + (int)threadIdx.x;
        const int j = (int)(r % g.ny), k = (int)(r / g.ny);
        if (i >= g.nx) continue;
        const long m = at(g, i, j, k);
        double v = (double)t[m] - (double)target[m];
        if (tlo) v += (double)tlo[m];
        const T l = (every_node || !is_border(g, i, j, k)) ? (T)v : T(0);
        lam[m] = l;
        lam2[m] = T(0);
        acc += (double)l * (double)l;
    }
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// A forward state is one array (F32, F64) or a (hi, lo) pair stored as two
// consecutive arrays of ne elements (F32_KAHAN).
enum { FAMILY_PENNES = 0, FAMILY_HEATVAR = 1, FAMILY_YEE_EPS = 2 };

struct Rev {
    int family = FAMILY_PENNES;
    // FAMILY_HEATVAR: SCAFES_HEATVAR_* mode of the forward, kappa field,
    // time step and border of the variable-kappa step
    int heat_mode = 0;
    const void* kappa = nullptr;
    double tau = 0.0, border_value = 0.0;
    int border_mode = SCAFES_BORDER_KEEP;
    // FAMILY_YEE_EPS: box, Courant number, ce = s/eps and eps fields, J
    // partials and J on the device
    YeeDims yd{};
    double yee_s = 0.0;
    const void* ce = nullptr;
    const void* eps = nullptr;
    double* partial = nullptr;
    double* jdev = nullptr;
    Ctx* c;
    Geo geo;
    const scafes_geom* g;
    int dtype;     // state arithmetic
    int lam_dtype; // adjoint field arithmetic (F32 or F64)
    long ne;       // elements per array
    size_t sbytes; // bytes per state
    const void* w;
    const void* q;
    const void* target;
    const scafes_pennes_params* p;
    double* grad;
    std::vector<void*> slots;  // snapshot states (free list)
    void* tmp[2];              // forward ping-pong states
    void* work;                // recomputed state when no snapshot is free
    void* lam[2];
    int lcur;
    int nsteps;
    bool pair;  // several adjoint steps per sweep where their states are at hand
    bool lam_ready = false;  // lam^N and J set (first_leaf)
    long fwd_steps, adj_steps;
    double J;

    bool kahan() const { return dtype == SCAFES_F32_KAHAN; }
    const void* lo(const void* s) const { return kahan() ? (const void*)((const float*)s + ne) : nullptr; }
    void* lo(void* s) const { return kahan() ? (void*)((float*)s + ne) : nullptr; }

    int step(const void* a, void* b)
    {
        ++fwd_steps;
        if (family == FAMILY_YEE_EPS) return yee_eps_step(c, yd, dtype, yee_s, ce, a, b);
        if (family == FAMILY_HEATVAR)
            return heat7var_step_one(c, geo, g, heat_mode, a, b, lo(a), lo(b), kappa, tau, border_mode,
                                     border_value);
        return pennes_step_launch(c, geo, g, dtype, a, b, w, q, p, lo(a), lo(b));
    }
    // the library's device copy kernel (about 1.8x the runtime's copy rate)
    int copy(void* dst, const void* src)
    {
        return scafes_memcpy_d2d(reinterpret_cast<scafes_ctx*>(c), dst, src, sbytes);
    }
    // State after n steps from src, written to dst (src is never written).
    // The Yee and variable-kappa families advance several steps per sweep
    // where that many remain (bit-identical to single steps); the
    // variable-kappa family splits n into sweeps as its forward run does.
    int advance(const void* src, long n, void* dst)
    {
        if (n == 0) return src == dst ? SCAFES_OK : copy(dst, src);
        const bool yee2 = family == FAMILY_YEE_EPS && yee_eps_two_step_supported(yd, dtype);
        const void* x = src;
        for (long done = 0; done < n;) {
            int k = 1;
            if (yee2 && n - done >= 2) k = 2;
            if (family == FAMILY_HEATVAR) k = heat7var_next_sweep(c, geo, heat_mode, n - done);
            done += k;
            void* out = (done == n) ? dst : (x == tmp[0] ? tmp[1] : tmp[0]);
            if (out == x) out = (x == tmp[0]) ? tmp[1] : tmp[0];
            int r;
            if (k > 1) {
                fwd_steps += k;
                if (family == FAMILY_YEE_EPS)
                    r = yee_eps_step2(c, yd, dtype, yee_s, ce, x, out);
                else
                    r = heat7var_sweep_k(c, geo, g, heat_mode, k, x, out, lo(x), lo(out), kappa, tau, border_mode,
                                         border_value);
            } else {
                r = step(x, out);
            }
            if (r) return r;
            x = out;
        }
        return x == dst ? SCAFES_OK : copy(dst, x);
    }
    // Adjoint step n (lam^n -> lam^{n-1}) with T^{n-1} = tprev.
    int adjoint(int n, const void* tprev)
    {
        if (family == FAMILY_YEE_EPS) {
            // J^N goes to the device and is read at the end
            if (n == nsteps) {
                int r = step(tprev, tmp[0]);
                if (!r) r = yee_eps_lam_init(c, yd, dtype, tmp[0], target, lam[lcur], partial, jdev);
                if (r) return r;
            }
            ++adj_steps;
            const int r = yee_eps_adjoint(c, yd, dtype, yee_s, ce, eps, tprev, lam[lcur], lam[1 - lcur], grad);
            lcur = 1 - lcur;
            return r;
        }
        if (n == nsteps && !lam_ready) {
            const int r = first_leaf(tprev);
            if (r) return r;
        }
        ++adj_steps;
        // the var-kappa adjoint reads the forward state's hi part only
        int r = family == FAMILY_HEATVAR
                    ? heatvar_adj_launch(c, geo, g, lam_dtype, lam[lcur], lam[1 - lcur], kappa, tprev, grad, tau)
                    : pennes_adj_launch(c, geo, g, lam_dtype, lam[lcur], lam[1 - lcur], w, tprev, grad, p,
                                        lo(tprev));
        lcur = 1 - lcur;
        return r;
    }
    // T^N from T^{N-1} = tprev (tprev = nullptr: T^N is in tmp[0] already),
    // then lam^N and J (the Pennes and variable-kappa families): after it the
    // reverse steps from N on may run one or several per sweep.
    int first_leaf(const void* tprev)
    {
        {
            int r = tprev ? step(tprev, tmp[0]) : SCAFES_OK;
            if (r) return r;
            const long total = (long)geo.nx * geo.ny * geo.nz;
            const int blocks = (int)std::min<long>((total + 255) / 256, c->red_capacity);
            // J over the nodes the step updates: every node under the
            // Tumor border, the inner nodes otherwise
            const int every = family == FAMILY_PENNES && p->border_mode == SCAFES_BORDER_TUMOR;
            if (lam_dtype == SCAFES_F64)
                lam_init_kernel<double><<<blocks, 256, 0, c->stream>>>(geo, (const double*)tmp[0], nullptr,
                                                                       (const double*)target, (double*)lam[0],
                                                                       (double*)lam[1], every, c->red_partial);
            else
                lam_init_kernel<float><<<blocks, 256, 0, c->stream>>>(geo, (const float*)tmp[0],
                                                                      (const float*)lo(tmp[0]), (const float*)target,
                                                                      (float*)lam[0], (float*)lam[1], every,
                                                                      c->red_partial);
            SCAFES_HIP(hipGetLastError());
            r = reduce_partials(c, blocks, &J);
            if (r) return r;
            J *= 0.5;
        }
        lam_ready = true;
        return SCAFES_OK;
    }
    // Reverse steps the next sweep takes when `left` (>= 2) remain with
    // their states at hand: as many as the kernel allows, but one fewer where
    // that would leave a single step (a one-step launch costs about two steps
    // of a multi-step sweep), so 5 -> 3 + 2 instead of 4 + 1.
    int sweep_k(long left) const
    {
        int kmax = family == FAMILY_HEATVAR && !heatvar_adjk_supported(geo) ? 2 : kAdjMaxK;
        // Under the Tumor border the four-step sweep needs more than the 80
        // VGPRs of two workgroups per CU and spills (267 us per launch at
        // 256^3 FP32, 67 us per step), the three-step one does not (161 us,
        // 54 per step): the 2000-step gradient with 500 snapshots took 293
        // ms with at most three reverse steps per sweep against 312 with four
        // (355 with four at one workgroup per CU).
        if (family == FAMILY_PENNES && p && p->border_mode == SCAFES_BORDER_TUMOR) kmax = 3;
        int K = (int)std::min<long>(kmax, left);
        if (left - K == 1 && K >= 3) --K;
        return K;
    }
    // Adjoint steps n, n-1, ..., n-K+1 in one sweep (lam^n -> lam^{n-K})
    // with st[l] = T^{n-1-l}; lam^n is set (first_leaf for n = nsteps).
    int adjointk(int K, const void* const* st)
    {
        adj_steps += K;
        if (family == FAMILY_HEATVAR) {
            // the var-kappa adjoint reads the states' hi parts
            int r = heatvar_adjk_supported(geo)
                        ? heatvar_adjk_launch(c, geo, g, lam_dtype, K, lam[lcur], lam[1 - lcur], kappa, st, grad,
                                              tau)
                        : heatvar_adj2_launch(c, geo, g, lam_dtype, lam[lcur], lam[1 - lcur], kappa, st[0], st[1],
                                              grad, tau);
            lcur = 1 - lcur;
            return r;
        }
        const void* los[kAdjMaxK];
        for (int l = 0; l < K; ++l) los[l] = lo(st[l]);
        int r = pennes_adjk_launch(c, geo, g, lam_dtype, K, lam[lcur], lam[1 - lcur], w, st, kahan() ? los : nullptr,
                                   grad, p);
        lcur = 1 - lcur;
        return r;
    }
    // Reverse steps a+1..b given T^a in `base`; s snapshots may be taken.
    int treeverse(long a, long b, int s, const void* base);

    // The Yee permittivity family's run of one-step binomial splits: the
    // peeled states T^{a+1}, T^{a+2}, ... are formed two per sweep where two
    // remain (the two-step sweep also writing the state after one step;
    // bit-identical to one-step sweeps, at about 0.6 instead of 0.95 of a
    // two-step sweep per state), the right part is reversed, then the peeled
    // steps one by one: the same forward steps, snapshots and adjoint order
    // as splitting one step at a time.
    int yee_peel_run(long a, long b, int s, const void* base)
    {
        long len = 0;
        {
            long aa = a;
            int ss = s;
            size_t free_slots = slots.size();
            while (b - aa > 1 && ss > 0 && free_slots > 0 && right_part(b - aa, ss) == b - aa - 1) {
                ++aa;
                --ss;
                --free_slots;
                ++len;
            }
        }
        std::vector<const void*> chain(1, base);  // chain[q] = T^{a+q}
        int r = SCAFES_OK;
        for (long q = 0; q < len && !r;) {
            if (len - q >= 2) {
                void* s1 = slots.back();
                slots.pop_back();
                void* s2 = slots.back();
                slots.pop_back();
                r = yee_eps_step2_states(c, yd, dtype, yee_s, ce, chain.back(), s1, s2);
                fwd_steps += 2;
                chain.push_back(s1);
                chain.push_back(s2);
                q += 2;
            } else {
                void* sl = slots.back();
                slots.pop_back();
                r = advance(chain.back(), 1, sl);
                chain.push_back(sl);
                q += 1;
            }
        }
        if (!r) r = treeverse(a + len, b, s - (int)len, chain.back());
        for (long n = a + len; n > a && !r; --n) r = adjoint((int)n, chain[n - 1 - a]);
        for (size_t q = chain.size() - 1; q >= 1; --q) slots.push_back(const_cast<void*>(chain[q]));
        return r;
    }

    // How the variable-kappa family forms `total` consecutive states:
    // `single` one-step sweeps first, then `q4` four-step and the rest
    // three-step state-writing sweeps.  Four-step sweeps (where the mode has
    // them) replace the one or two single steps a total that is not a
    // multiple of three would leave (4 = 4, 8 = 4 + 4, 10 = 4 + 3 + 3).
    void states_plan(long total, long* single, long* q4) const
    {
        *single = total;
        *q4 = 0;
        if (total < 3 || !heat7var_states_sweep_ok(c, geo, heat_mode)) return;
        *single = total % 3;
        if (!heat7var_states4_ok(c, geo, heat_mode)) return;
        if (*single == 1 && total >= 4) {
            *q4 = 1;
            *single = 0;
        } else if (*single == 2 && total >= 8) {
            *q4 = 2;
            *single = 0;
        }
    }

    // Where snapshots are plentiful the binomial split takes one step off
    // the left again and again: each time T^{a+1} goes into a snapshot, the
    // right part is reversed with one snapshot less, then step a+1 alone.
    // The same run of splits is taken here in one go (the same forward steps
    // and snapshots); the peeled states stay in their snapshots while the
    // right part is reversed, and their steps are then reversed together,
    // up to kAdjMaxK per sweep (bit-identical to single steps).
    int peel_run(long a, long b, int s, const void* base)
    {
        long len = 0;  // splits in the run
        {
            long x = a;
            int sx = s;
            long free_slots = (long)slots.size();
            while (b - x > 1 && sx > 0 && free_slots > 0 && right_part(b - x, sx) == b - x - 1 &&
                   !(sx >= b - x - 1 && free_slots >= b - x - 1)) {
                ++x;
                --sx;
                --free_slots;
                ++len;
            }
        }
        // The variable-kappa family forms the peeled states three or four per
        // sweep where the grid's sweeps can write the states in between
        // (value parts, all the adjoint reads), the rest first one step at a
        // time, so the run's last state (the right part's base) is whole.
        std::vector<const void*> chain(1, base);  // chain[q] = T^{a+q}
        long single = len, q4 = 0;
        if (family == FAMILY_HEATVAR) states_plan(len, &single, &q4);
        int r = SCAFES_OK;
        for (long q = 0; q < single && !r; ++q) {
            void* sl = slots.back();
            slots.pop_back();
            r = advance(chain.back(), 1, sl);
            chain.push_back(sl);
        }
        while (!r && (long)chain.size() - 1 < len) {
            const int k = q4 > 0 ? 4 : 3;
            if (q4 > 0) --q4;
            void* sk[4];
            for (int q = 0; q < k; ++q) {
                sk[q] = slots.back();
                slots.pop_back();
            }
            const void* x = chain.back();
            r = heat7var_sweep_states(c, geo, g, heat_mode, k, x, sk[k - 1], lo(x), lo(sk[k - 1]), sk, kappa, tau,
                                      border_mode, border_value);
            fwd_steps += k;
            for (int q = 0; q < k; ++q) chain.push_back(sk[q]);
        }
        const long aa = a + len;
        const int ss = s - (int)len;
        if (!r) r = treeverse(aa, b, ss, chain.back());
        for (long n = aa; n > a && !r;) {
            if (n == nsteps && !lam_ready) {
                r = first_leaf(chain[n - 1 - a]);
                if (r) break;
            }
            if (n - a < 2) {
                r = adjoint((int)n, chain[n - 1 - a]);
                n -= 1;
            } else {
                const int K = sweep_k(n - a);
                const void* sk[kAdjMaxK];
                for (int l = 0; l < K; ++l) sk[l] = chain[n - 1 - l - a];
                r = adjointk(K, sk);
                n -= K;
            }
        }
        for (size_t q = chain.size() - 1; q >= 1; --q) slots.push_back(const_cast<void*>(chain[q]));
        return r;
    }
};

int Rev::treeverse(long a, long b, int s, const void* base)
{
    {
        if (b - a == 1) return adjoint((int)b, base);
        if (pair && s >= b - a - 1 && (long)slots.size() >= b - a - 1) {
            // Repetition number 1: every intermediate state T^{a+1..b-1} gets
            // a snapshot (b-a-1 forward steps, as the recursive split would
            // take), then the steps are reversed up to kAdjMaxK per sweep.
            std::vector<const void*> st(1, base);  // st[q] = T^{a+q}
            int r = SCAFES_OK;
            // The variable-kappa family takes the first states one step at a
            // time, the rest three or four per sweep with the states in
            // between written by the same sweep (value parts only: its
            // adjoint reads no more), so that the last state of each sweep is
            // whole.  Where the stretch ends at N and lam^N is not set yet,
            // T^N is the last sweep's output (into tmp[0]): the m + 1 states
            // are split instead of m and one more step for T^N.
            const long m = b - a - 1;
            const bool sweeps = family == FAMILY_HEATVAR && heat7var_states_sweep_ok(c, geo, heat_mode);
            const bool leaf = sweeps && b == nsteps && !lam_ready && m + 1 >= 3;
            const long total = leaf ? m + 1 : m;  // states this stretch forms
            long single = m, q4 = 0;
            if (family == FAMILY_HEATVAR) states_plan(total, &single, &q4);
            for (long q = 1; q <= single && !r; ++q) {
                void* sl = slots.back();
                slots.pop_back();
                r = step(st.back(), sl);
                st.push_back(sl);
            }
            while (!r && (long)st.size() - 1 < total) {
                const int k = q4 > 0 ? 4 : 3;
                if (q4 > 0) --q4;
                const bool to_leaf = leaf && (long)st.size() - 1 + k == total;
                void* sk[4];
                for (int q = 0; q < k; ++q) {
                    if (q == k - 1 && to_leaf) {
                        sk[q] = tmp[0];
                    } else {
                        sk[q] = slots.back();
                        slots.pop_back();
                    }
                }
                const void* x = st.back();
                r = heat7var_sweep_states(c, geo, g, heat_mode, k, x, sk[k - 1], lo(x), lo(sk[k - 1]), sk, kappa,
                                          tau, border_mode, border_value);
                fwd_steps += k;
                for (int q = 0; q < k - 1; ++q) st.push_back(sk[q]);
                if (to_leaf) break;
                st.push_back(sk[k - 1]);
            }
            if (!r && leaf) r = first_leaf(nullptr);
            for (long n = b; n > a && !r;) {
                if (n == nsteps && !lam_ready) {
                    r = first_leaf(st[n - 1 - a]);
                    if (r) break;
                }
                if (n - a < 2) {
                    r = adjoint((int)n, st[n - 1 - a]);
                    n -= 1;
                } else {
                    const int K = sweep_k(n - a);
                    const void* sk[kAdjMaxK];
                    for (int l = 0; l < K; ++l) sk[l] = st[n - 1 - l - a];
                    r = adjointk(K, sk);
                    n -= K;
                }
            }
            for (size_t q = st.size() - 1; q >= 1; --q) slots.push_back(const_cast<void*>(st[q]));
            return r;
        }
        if (s <= 0 || slots.empty()) {
            for (long n = b; n > a; --n) {
                const void* tp = base;
                if (n - 1 > a) {
                    int r = advance(base, n - 1 - a, work);
                    if (r) return r;
                    tp = work;
                }
                int r = adjoint((int)n, tp);
                if (r) return r;
            }
            return SCAFES_OK;
        }
        if (pair && right_part(b - a, s) == b - a - 1) return peel_run(a, b, s, base);
        if (family == FAMILY_YEE_EPS && s >= 2 && slots.size() >= 2 && right_part(b - a, s) == b - a - 1 &&
            right_part(b - a - 1, s - 1) == b - a - 2 && yee_eps_two_step_supported(yd, dtype))
            return yee_peel_run(a, b, s, base);
        const long right = right_part(b - a, s);
        const long m = b - right;
        void* slot = slots.back();
        slots.pop_back();
        int r = advance(base, m - a, slot);
        if (!r) r = treeverse(m, b, s - 1, slot);
        slots.push_back(slot);
        if (r) return r;
        return treeverse(a, m, s, base);
    }
}

// Counts the forward steps of the schedule without running it.
static long count_forward(long a, long b, int s, int nsteps)
{
    if (b - a == 1) return (b == nsteps) ? 1 : 0;
    if (s <= 0) {
        long n = 0;
        for (long k = b; k > a; --k) n += (k - 1 - a) + ((k == nsteps) ? 1 : 0);
        return n;
    }
    const long right = right_part(b - a, s);
    const long m = b - right;
    return (m - a) + count_forward(m, b, s - 1, nsteps) + count_forward(a, m, s, nsteps);
}

}  // namespace scafes

using namespace scafes;

// Reverse sweep with binomial checkpointing: R carries the family's fields;
// t0 is the initial state (FP32 fields in the FP32 modes); a state is one
// array of abytes (two in the compensated mode), an adjoint field abytes,
// the gradient ngrad FP64 values.
static int run_schedule(Rev& R, Ctx* c, int dtype, const void* t0, size_t abytes, long ngrad, const void* target,
                        int nsteps, int max_snapshots, double* grad, double* J, double* stats)
{
    int rc = SCAFES_OK;
    const int snaps = std::max(0, std::min(max_snapshots, nsteps - 1));
    R.c = c;
    R.dtype = dtype;
    R.lam_dtype = dtype == SCAFES_F64 ? SCAFES_F64 : SCAFES_F32;
    R.sbytes = dtype == SCAFES_F32_KAHAN ? 2 * abytes : abytes;
    R.target = target;
    R.grad = grad;
    R.lcur = 0;
    R.nsteps = nsteps;
    // several reverse steps per sweep where the states are at hand (3-D grids)
    R.pair = (R.family == FAMILY_PENNES && pennes_adjk_supported(R.geo)) ||
             (R.family == FAMILY_HEATVAR && heatvar_adj2_supported(R.geo));
    R.fwd_steps = R.adj_steps = 0;
    R.J = 0.0;
    R.work = nullptr;
    // Workspace: [compensated initial state] tmp[0] tmp[1] work lam[0]
    // lam[1] and the snapshot pool (a store-all budget of a 256^3 x
    // 2000-step FP32 run is 134 GB; HBM holds 288 GB), carved from one
    // buffer the context keeps for the next gradient call: re-allocating
    // freed memory makes the driver clear it first (about 0.5 s per 137 GB).
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const bool kahan0 = dtype == SCAFES_F32_KAHAN;
    const size_t sb = up(R.sbytes), lb = up(abytes);
    const size_t need = (kahan0 ? sb : 0) + 3 * sb + 2 * lb + (size_t)snaps * sb;
    if (c->grad_ws_bytes < need) {
        if (c->grad_ws) (void)hipFree(c->grad_ws);
        c->grad_ws = nullptr;
        c->grad_ws_bytes = 0;
        hipError_t e = hipMalloc(&c->grad_ws, need);
        if (e != hipSuccess) {
            c->grad_ws = nullptr;
            return check_hip(e, "gradient workspace");
        }
        c->grad_ws_bytes = need;
    }
    char* ws = (char*)c->grad_ws;
    // the initial state as a state (a (t0, 0) pair in the compensated mode)
    const void* base = t0;
    if (kahan0) {
        void* base_buf = ws;
        ws += sb;
        rc = scafes_memcpy_d2d(reinterpret_cast<scafes_ctx*>(c), base_buf, t0, abytes);
        if (!rc) {
            hipError_t e = hipMemsetAsync((char*)base_buf + abytes, 0, abytes, c->stream);
            if (e != hipSuccess) rc = check_hip(e, "gradient init");
        }
        base = base_buf;
    }
    R.tmp[0] = ws;
    R.tmp[1] = ws + sb;
    R.work = ws + 2 * sb;
    R.lam[0] = ws + 3 * sb;
    R.lam[1] = ws + 3 * sb + lb;
    ws += 3 * sb + 2 * lb;
    for (int s = 0; s < snaps; ++s) R.slots.push_back(ws + (size_t)s * sb);
    if (!rc) {
        hipError_t e = hipMemsetAsync(grad, 0, (size_t)ngrad * sizeof(double), c->stream);
        if (e != hipSuccess) rc = check_hip(e, "gradient zero");
    }
    if (!rc) rc = R.treeverse(0, nsteps, (int)R.slots.size(), base);
    if (!rc) {
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = check_hip(e, "gradient");
    }
    if (rc) return rc;
    *J = R.J;
    if (stats) {
        stats[0] = (double)R.fwd_steps;
        stats[1] = (double)R.adj_steps;
        stats[2] = (double)snaps;
        stats[3] = (double)snaps * (double)R.sbytes;
    }
    return SCAFES_OK;
}

// Node-field families (Pennes, variable kappa): one state array per field
// of the grid g with halo hw.
static int run_gradient(Rev& R, Ctx* c, const scafes_geom* g, int hw, int dtype, const void* t0,
                        const void* target, int nsteps, int max_snapshots, double* grad, double* J, double* stats)
{
    Geo geo;
    int rc = make_geo(g, hw, &geo);
    if (rc) return rc;
    long ne = 1;
    for (int p2 = 0; p2 < 3; ++p2) ne *= (p2 < g->dim) ? (g->n[p2] + 2 * hw) : 1;
    const size_t es = dtype == SCAFES_F64 ? 8 : 4;
    R.geo = geo;
    R.g = g;
    R.ne = ne;
    return run_schedule(R, c, dtype, t0, (size_t)ne * es, ne, target, nsteps, max_snapshots, grad, J, stats);
}

extern "C" {

long scafes_checkpoint_forward_steps(int nsteps, int snaps)
{
    if (nsteps < 1) return 0;
    return count_forward(0, nsteps, std::max(0, std::min(snaps, nsteps - 1)), nsteps);
}

int scafes_pennes_gradient(scafes_ctx* ctx, const scafes_geom* g, int hw, int dtype, const void* t0, const void* w,
                           const void* q, const void* target, const scafes_pennes_params* p, int nsteps,
                           int max_snapshots, double* grad, double* J, double* stats)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !g || !t0 || !w || !q || !target || !p || !grad || !J)
        return fail(SCAFES_ERR_ARG, "pennes gradient: null pointer");
    if (nsteps < 1) return fail(SCAFES_ERR_ARG, "pennes gradient: need nsteps >= 1");
    if (dtype != SCAFES_F32 && dtype != SCAFES_F64 && dtype != SCAFES_F32_KAHAN)
        return fail(SCAFES_ERR_ARG, "pennes gradient: bad dtype");
    if (hw < 1 || g->nlayers < 1) return fail(SCAFES_ERR_ARG, "pennes gradient: need hw >= 1 and nlayers >= 1");
    if (p->border_mode != SCAFES_BORDER_KEEP && p->border_mode != SCAFES_BORDER_TUMOR)
        return fail(SCAFES_ERR_ARG, "pennes gradient: border_mode must be KEEP or TUMOR");
    if (p->border_mode == SCAFES_BORDER_TUMOR && p->k_cond == 0.0)
        return fail(SCAFES_ERR_ARG, "pennes gradient: the TUMOR border needs k_cond != 0");
    {
        Geo geo;
        int rc = make_geo(g, hw, &geo);
        if (rc) return rc;
        // the perfusion field changes between the calls of an optimisation: no cache
        rc = pennes_check_tau(c, geo, g, hw, dtype == SCAFES_F64 ? SCAFES_F64 : SCAFES_F32, w, p, false,
                              "pennes gradient");
        if (rc) return rc;
    }
    Rev R;
    R.family = FAMILY_PENNES;
    R.w = w;
    R.q = q;
    R.p = p;
    return run_gradient(R, c, g, hw, dtype, t0, target, nsteps, max_snapshots, grad, J, stats);
}

int scafes_heat7var_gradient(scafes_ctx* ctx, const scafes_geom* g, int hw, int mode, const void* u0,
                             const void* kappa, const void* target, double tau, int border_mode,
                             double border_value, int nsteps, int max_snapshots, double* grad, double* J,
                             double* stats)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !g || !u0 || !kappa || !target || !grad || !J)
        return fail(SCAFES_ERR_ARG, "heat7var gradient: null pointer");
    if (nsteps < 1) return fail(SCAFES_ERR_ARG, "heat7var gradient: need nsteps >= 1");
    if (mode < SCAFES_HEATVAR_F64 || mode > SCAFES_HEATVAR_F32_KAHAN)
        return fail(SCAFES_ERR_ARG, "heat7var gradient: bad mode");
    if (hw < 1 || g->nlayers < 1) return fail(SCAFES_ERR_ARG, "heat7var gradient: need hw >= 1 and nlayers >= 1");
    if (border_mode != SCAFES_BORDER_KEEP && border_mode != SCAFES_BORDER_VALUE)
        return fail(SCAFES_ERR_UNSUPPORTED, "heat7var gradient: border must be KEEP or VALUE");
    static const int kDtype[3] = {SCAFES_F64, SCAFES_F32, SCAFES_F32_KAHAN};
    {
        Geo geo;
        int rc = make_geo(g, hw, &geo);
        if (rc) return rc;
        double kmax = 0.0;
        rc = coef_abs_max(c, geo, hw, mode == SCAFES_HEATVAR_F64 ? SCAFES_F64 : SCAFES_F32, kappa, false, &kmax);
        if (rc) return rc;
        rc = check_explicit_tau(tau, diffusion_rate(g, kmax), "heat7var gradient");
        if (rc) return rc;
    }
    Rev R;
    R.family = FAMILY_HEATVAR;
    R.heat_mode = mode;
    R.kappa = kappa;
    R.tau = tau;
    R.border_mode = border_mode;
    R.border_value = border_value;
    R.w = R.q = nullptr;
    R.p = nullptr;
    return run_gradient(R, c, g, hw, kDtype[mode], u0, target, nsteps, max_snapshots, grad, J, stats);
}

int scafes_yee_eps_gradient(scafes_ctx* ctx, int ncell, int dtype, double s, const void* f0, const void* eps,
                            const void* target, int nsteps, int max_snapshots, double* grad, double* J,
                            double* stats)
{
    Ctx* c = as_ctx(ctx);
    if (!c || !f0 || !eps || !target || !grad || !J) return fail(SCAFES_ERR_ARG, "yee eps gradient: null pointer");
    YeeDims d;
    if (!yee_dims(ncell, &d)) return fail(SCAFES_ERR_ARG, "yee eps gradient: need 2 <= ncell <= 1400");
    if (nsteps < 1) return fail(SCAFES_ERR_ARG, "yee eps gradient: need nsteps >= 1");
    if (!(s > 0.0) || s > 1.0 / 1.7320508075688772)
        return fail(SCAFES_ERR_ARG, "yee eps gradient: Courant number must satisfy 0 < s <= 1/sqrt(3)");
    if (dtype != SCAFES_F32 && dtype != SCAFES_F64) return fail(SCAFES_ERR_ARG, "yee eps gradient: bad dtype");
    const size_t es = dtype == SCAFES_F64 ? 8 : 4;
    Rev R;
    R.family = FAMILY_YEE_EPS;
    R.yd = d;
    R.yee_s = s;
    R.eps = eps;
    R.ne = 6 * d.nf;
    R.w = R.q = nullptr;
    R.p = nullptr;
    void* ce = nullptr;
    void* aux = nullptr;  // J partials, then J
    const size_t naux = (size_t)yee_eps_partials() + 1;
    hipError_t e = hipMalloc(&ce, (size_t)d.nf * es);
    if (e == hipSuccess) e = hipMalloc(&aux, naux * sizeof(double));
    int rc = e == hipSuccess ? SCAFES_OK : check_hip(e, "yee eps gradient buffers");
    if (!rc) rc = yee_eps_coef(c, d, dtype, s, eps, ce);
    if (!rc) {
        R.ce = ce;
        R.partial = (double*)aux;
        R.jdev = (double*)aux + yee_eps_partials();
        rc = run_schedule(R, c, dtype, f0, 6 * (size_t)d.nf * es, d.nf, target, nsteps, max_snapshots, grad, J,
                          stats);
    }
    if (!rc) {
        e = hipMemcpy(J, R.jdev, sizeof(double), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = check_hip(e, "yee eps gradient J");
    }
    (void)hipFree(ce);
    (void)hipFree(aux);
    return rc;
}

}  // extern "C"
