// system_problems.hpp -- the reference's exact-arithmetic system tests as
// problem classes written against the host API: their update hooks are
// generic in TT, with no DeviceTraits, so a hipcc-compiled driver runs them
// through the generic device update phase (the hooks themselves on the GPU).
// Each known solution is exact in floating point, so every "Local Linf"
// line must read 0.000000e+00 (scripts/scafesrun.sh:940-963).
//
//   StencilOne   (tests/system/HeatEqnFdmStencilWidthOne): u^0 = sum_p i_p;
//                inner nodes average themselves and their 2*DIM neighbours
//                and double it; a border node adds the opposite neighbour
//                and corrects by 2^(n-1); exact u^n = 2^n sum_p i_p.  The
//                powers of two are ldexp(1, k), exact on host and device
//                alike (the reference's hook writes pow(2.0, n-1)).
//   StencilVar   (tests/system/HeatEqnFdmStencilWidthVar): fields with
//                stencil widths 0, 1, 2; every node averages the values up to
//                its field's stencil width away along each axis (as far as
//                the grid reaches), corrected by the offsets, plus 2;
//                exact u^n = sum_p i_p + 2n.
//   Propagation  (tests/system/PropagationFromLeftLowerCorner): a front
//                starting at the first node; a node not yet reached takes the
//                time step's value once a minus-neighbour is reached; exact
//                u^n = sum_p i_p where sum_p i_p <= n, else -25.
#ifndef SCAFES_EXAMPLES_SYSTEM_PROBLEMS_HPP_
#define SCAFES_EXAMPLES_SYSTEM_PROBLEMS_HPP_

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "ScaFES.hpp"

namespace systests {

template <typename CT, std::size_t DIM>
using VDf = std::vector<ScaFES::DataField<CT, DIM>>;
template <std::size_t DIM>
using Idx = ScaFES::Ntuple<int, DIM>;

// Common constructor plumbing of the three classes.
#define SYSTEST_PROBLEM_CTOR(NAME)                                                                        \
    NAME(ScaFES::Parameters const& params, ScaFES::GridGlobal<DIM> const& gg, bool useLeapfrog,          \
         std::vector<std::string> const& names, std::vector<int> const& stencilWidth,                    \
         std::vector<bool> const& isKnownDf, std::vector<int> const& nLayers,                            \
         std::vector<CT> const& defaultValue, std::vector<ScaFES::WriteHowOften> const& writeToFile,     \
         std::vector<bool> const& computeError)                                                          \
        : ScaFES::Problem<NAME<CT, DIM>, CT, DIM>(params, gg, useLeapfrog, names, stencilWidth, isKnownDf,  \
                                                  nLayers, defaultValue, writeToFile, computeError)      \
    {                                                                                                     \
    }

template <std::size_t DIM>
inline int indexSum(const Idx<DIM>& x)
{
    int s = 0;
    for (std::size_t p = 0; p < DIM; ++p) s += x[p];
    return s;
}

// ------------------------------------------------------------ StencilOne
template <typename CT, std::size_t DIM>
class StencilOne : public ScaFES::Problem<StencilOne<CT, DIM>, CT, DIM> {
public:
    SYSTEST_PROBLEM_CTOR(StencilOne)

    void evalInner(VDf<CT, DIM>& v, Idx<DIM> const& x, int const& n)
    {
        for (std::size_t k = 0; k < v.size(); ++k) v[k](x) = indexSum(x) * std::ldexp(1.0, n);
    }
    void evalBorder(VDf<CT, DIM>& v, Idx<DIM> const& x, int const& n) { evalInner(v, x, n); }
    template <typename TT>
    void initInner(VDf<TT, DIM>& v, std::vector<TT> const&, Idx<DIM> const& x, int const&)
    {
        for (std::size_t k = 0; k < v.size(); ++k) v[k](x) = indexSum(x);
    }
    template <typename TT>
    void initBorder(VDf<TT, DIM>& v, std::vector<TT> const& p, Idx<DIM> const& x, int const& n)
    {
        initInner<TT>(v, p, x, n);
    }
    template <typename TT>
    void updateInner(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const&)
    {
        for (std::size_t k = 0; k < vNew.size(); ++k) {
            TT sum = vOld[k](x);
            for (int dir = 0; dir < 2 * (int)DIM; ++dir) sum += vOld[k](this->connect(x, dir));
            vNew[k](x) = sum * (2.0 / (2 * DIM + 1));
        }
    }
    template <typename TT>
    void updateBorder(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const& n)
    {
        // the step's value of one index unit, halved: 2^(n-1), written with
        // ldexp, which is exact on the device too (the device math library's
        // pow is within an ulp but not correctly rounded: pow(2.0, 11) gives
        // 2048 - 2^-41 there)
        const double half = std::ldexp(1.0, n - 1);
        for (std::size_t k = 0; k < vNew.size(); ++k)
            for (int dir = 0; dir < 2 * (int)DIM; ++dir) {
                if (!(this->connect(x, dir) == Idx<DIM>(-(dir + 1)))) continue;  // neighbour exists
                const int opposite = (dir % 2 == 0) ? dir + 1 : dir - 1;
                const double sign = (dir % 2 == 0) ? 1.0 : -1.0;
                vNew[k](x) = vOld[k](x) + vOld[k](this->connect(x, opposite)) - sign * half;
            }
    }
    template <typename TT>
    void updateInner2(VDf<TT, DIM>&, VDf<TT, DIM> const&, Idx<DIM> const&, int const&)
    {
    }
    template <typename TT>
    void updateBorder2(VDf<TT, DIM>&, VDf<TT, DIM> const&, Idx<DIM> const&, int const&)
    {
    }
};

// ------------------------------------------------------------ StencilVar
template <typename CT, std::size_t DIM>
class StencilVar : public ScaFES::Problem<StencilVar<CT, DIM>, CT, DIM> {
public:
    SYSTEST_PROBLEM_CTOR(StencilVar)

    void evalInner(VDf<CT, DIM>& v, Idx<DIM> const& x, int const& n)
    {
        for (std::size_t k = 0; k < v.size(); ++k) v[k](x) = indexSum(x) + 2.0 * n;
    }
    void evalBorder(VDf<CT, DIM>& v, Idx<DIM> const& x, int const& n) { evalInner(v, x, n); }
    template <typename TT>
    void initInner(VDf<TT, DIM>& v, std::vector<TT> const&, Idx<DIM> const& x, int const&)
    {
        for (std::size_t k = 0; k < v.size(); ++k) v[k](x) = indexSum(x);
    }
    template <typename TT>
    void initBorder(VDf<TT, DIM>& v, std::vector<TT> const& p, Idx<DIM> const& x, int const& n)
    {
        initInner<TT>(v, p, x, n);
    }
    template <typename TT>
    void updateInner(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const&)
    {
        for (std::size_t k = 0; k < vNew.size(); ++k) {
            const ScaFES::Grid<DIM>& dom = vOld[k].gridGlobal().discreteDomain();
            const Idx<DIM> below = x - dom.idxNodeFirst(), above = dom.idxNodeLast() - x;
            const int reach = vOld[k].stencilWidth();
            TT sum = 0.0;
            int offsets = 0, count = 1;
            for (std::size_t p = 0; p < DIM; ++p) {
                const int lo = std::min(below[p], reach), hi = std::min(above[p], reach);
                for (int r = -lo; r <= hi; ++r) {
                    if (r == 0) continue;
                    Idx<DIM> y = x;
                    y[p] += r;
                    sum += vOld[k](y);
                    offsets += r;
                }
                count += lo + hi;
            }
            vNew[k](x) = 2.0 + (sum + vOld[k](x) - offsets) / count;
        }
    }
    template <typename TT>
    void updateBorder(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const& n)
    {
        updateInner<TT>(vNew, vOld, x, n);
    }
    template <typename TT>
    void updateInner2(VDf<TT, DIM>&, VDf<TT, DIM> const&, Idx<DIM> const&, int const&)
    {
    }
    template <typename TT>
    void updateBorder2(VDf<TT, DIM>&, VDf<TT, DIM> const&, Idx<DIM> const&, int const&)
    {
    }
};

// ------------------------------------------------------------ Propagation
const double kUnreached = -25.0;
const double kTolerance = 2.2e-2;

template <typename CT, std::size_t DIM>
class Propagation : public ScaFES::Problem<Propagation<CT, DIM>, CT, DIM> {
public:
    SYSTEST_PROBLEM_CTOR(Propagation)

    void evalInner(VDf<CT, DIM>& v, Idx<DIM> const& x, int const& n)
    {
        const int s = indexSum(x);
        for (std::size_t k = 0; k < v.size(); ++k) v[k](x) = (s > n) ? kUnreached : (double)s;
    }
    void evalBorder(VDf<CT, DIM>& v, Idx<DIM> const& x, int const& n) { evalInner(v, x, n); }
    // The front starts where all coordinates are below the tolerance (the
    // first node of a domain starting at 0).
    template <typename TT>
    void initInner(VDf<TT, DIM>& v, std::vector<TT> const&, Idx<DIM> const& x, int const&)
    {
        const bool start = this->coordinates(x) < ScaFES::Ntuple<double, DIM>(kTolerance);
        for (std::size_t k = 0; k < v.size(); ++k) v[k](x) = start ? 0.0 : kUnreached;
    }
    template <typename TT>
    void initBorder(VDf<TT, DIM>& v, std::vector<TT> const& p, Idx<DIM> const& x, int const& n)
    {
        initInner<TT>(v, p, x, n);
    }
    template <typename TT>
    void updateInner(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const& n)
    {
        advance<TT>(vNew, vOld, x, n, false);
    }
    template <typename TT>
    void updateBorder(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const& n)
    {
        advance<TT>(vNew, vOld, x, n, true);
    }
    template <typename TT>
    void updateInner2(VDf<TT, DIM>&, VDf<TT, DIM> const&, Idx<DIM> const&, int const&)
    {
    }
    template <typename TT>
    void updateBorder2(VDf<TT, DIM>&, VDf<TT, DIM> const&, Idx<DIM> const&, int const&)
    {
    }

private:
    template <typename TT>
    void advance(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const& n, bool border)
    {
        for (std::size_t k = 0; k < vNew.size(); ++k) {
            bool reached = false;
            for (std::size_t p = 0; p < DIM && !reached; ++p) {
                if (border && !(vOld[k].gridGlobal().discreteDomain().idxNodeFirst(p) < x[p])) continue;
                const TT below = vOld[k](this->connect(x, 2 * (int)p));
                reached = fabs(below - kUnreached) >= kTolerance;
            }
            const bool waiting = fabs(vOld[k](x) - kUnreached) < kTolerance;
            vNew[k](x) = (reached && waiting) ? TT((double)n) : vOld[k](x);
        }
    }
};

#undef SYSTEST_PROBLEM_CTOR

}  // namespace systests
#endif
