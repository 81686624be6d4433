// families.cpp -- one problem class per device kernel family, written
// against the host API the way a ScaFES user writes a problem: the hooks
// state the per-node update (the reference runs them per node on the CPU,
// src/ScaFES_Problem.hpp:5397-5476), and a DeviceTraits specialisation names
// the kernel family the update phase runs as on the GPU.
//
// Build variants (-DEXAMPLE_FAMILY=...; -DEXAMPLE_NO_TRAITS with hipcc: no
// DeviceTraits, the generic device update runs the hooks below):
//   HEATVAR  conservative variable-kappa heat, kappa a known field with the
//            unknown's stencil width, zero Dirichlet border
//            (-DEXAMPLE_PRECISION=F64 or F32_KAHAN)
//   PENNES   Pennes bioheat with perfusion and source fields in the
//            TumorHeatEqnFDM form (examples/TumorHeatEqnFDM/TumorHeatEqnFDM.hpp:233-269),
//            border nodes keep their value (-DEXAMPLE_PRECISION=F64 or F32)
//   YEE      leapfrog Yee FDTD in a PEC box (updateInner = H half step,
//            updateInner2 = E half step, as examples/EMFDTD/EMFDTD.hpp:295-386
//            splits them) (-DEXAMPLE_PRECISION=F64 or F32)
// Initial and coefficient values are small integer patterns divided by
// powers of two, so another program can rebuild them bit for bit.  The
// final unknown fields are dumped to $SCAFES_DUMP (suffix = field index).
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <stdexcept>

#include "ScaFES.hpp"
#include "example_util.hpp"

#define FAMILY_HEATVAR 1
#define FAMILY_PENNES 2
#define FAMILY_YEE 3

#ifndef EXAMPLE_FAMILY
#define EXAMPLE_FAMILY FAMILY_HEATVAR
#endif
#ifndef EXAMPLE_PRECISION
#define EXAMPLE_PRECISION F64
#endif

// The problem classes are host+device code in a hipcc build (the
// EXAMPLE_NO_TRAITS variants run their own hooks on the GPU).
SCAFES_DEVICE_HOOKS_BEGIN

template <typename CT, std::size_t DIM>
using VDf = std::vector<ScaFES::DataField<CT, DIM>>;
template <std::size_t DIM>
using Idx = ScaFES::Ntuple<int, DIM>;

// Pattern ((a i + b j + c k) mod m) / d.
template <std::size_t DIM>
inline double pattern(const Idx<DIM>& x, int a, int b, int c, int m, double d)
{
    return (double)((a * x[0] + b * x[1] + c * x[2]) % m) / d;
}

// ------------------------------------------------------------------ HEATVAR
template <typename CT, std::size_t DIM>
class HeatVar : public ScaFES::Problem<HeatVar<CT, DIM>, CT, DIM> {
public:
    typedef ScaFES::Problem<HeatVar<CT, DIM>, CT, DIM> Base;
    using Base::Base;
    // field 0: kappa (known), field 1: U
    void evalInner(VDf<CT, DIM>& v, Idx<DIM> const& x, int const&) { v[0](x) = 1.0 + pattern(x, 1, 2, 3, 5, 8.0); }
    void evalBorder(VDf<CT, DIM>& v, Idx<DIM> const& x, int const& n) { evalInner(v, x, n); }
    template <typename TT>
    void initInner(VDf<TT, DIM>& v, std::vector<TT> const&, Idx<DIM> const& x, int const&)
    {
        v[0](x) = pattern(x, 3, 5, 7, 13, 16.0);
    }
    template <typename TT>
    void initBorder(VDf<TT, DIM>& v, std::vector<TT> const&, Idx<DIM> const& x, int const&)
    {
        v[0](x) = 0.0;
    }
    template <typename TT>
    void updateInner(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const&)
    {
        const CT kc = this->knownDf(0, x);
        vNew[0](x) = vOld[0](x);
        for (std::size_t p = 0; p < DIM; ++p) {
            const Idx<DIM> w = this->connect(x, 2 * p), e = this->connect(x, 2 * p + 1);
            vNew[0](x) += this->tau() *
                          (0.5 * (this->knownDf(0, e) + kc) * (vOld[0](e) - vOld[0](x)) -
                           0.5 * (kc + this->knownDf(0, w)) * (vOld[0](x) - vOld[0](w))) /
                          (this->gridsize(p) * this->gridsize(p));
        }
    }
    template <typename TT>
    void updateBorder(VDf<TT, DIM>& vNew, VDf<TT, DIM> const&, Idx<DIM> const& x, int const&)
    {
        vNew[0](x) = 0.0;
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

// ------------------------------------------------------------------ PENNES
template <typename CT, std::size_t DIM>
class PennesField : public ScaFES::Problem<PennesField<CT, DIM>, CT, DIM> {
public:
    typedef ScaFES::Problem<PennesField<CT, DIM>, CT, DIM> Base;
    using Base::Base;
    const double K = 0.5, PERF = 1.0, SRC = 1.0, T_A = 37.0;  // k/(rho c), rho_b c_b/(rho c), 1/(rho c)
    // field 0: w (known), field 1: Q (known), field 2: T
    void evalInner(VDf<CT, DIM>& v, Idx<DIM> const& x, int const&)
    {
        v[0](x) = 0.5 + pattern(x, 1, 1, 1, 4, 8.0);
        v[1](x) = pattern(x, 2, 1, 3, 7, 4.0);
    }
    void evalBorder(VDf<CT, DIM>& v, Idx<DIM> const& x, int const& n) { evalInner(v, x, n); }
    template <typename TT>
    void initInner(VDf<TT, DIM>& v, std::vector<TT> const&, Idx<DIM> const& x, int const&)
    {
        v[0](x) = T_A + pattern(x, 1, 2, 1, 5, 4.0);
    }
    template <typename TT>
    void initBorder(VDf<TT, DIM>& v, std::vector<TT> const& p, Idx<DIM> const& x, int const& n)
    {
        initInner<TT>(v, p, x, n);
    }
    template <typename TT>
    void updateInner(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const&)
    {
        vNew[0](x) = vOld[0](x);
        for (std::size_t p = 0; p < DIM; ++p)
            vNew[0](x) += (this->tau() * K *
                           (vOld[0](this->connect(x, 2 * p)) + vOld[0](this->connect(x, 2 * p + 1)) - 2.0 * vOld[0](x))) /
                          (this->gridsize(p) * this->gridsize(p));
        vNew[0](x) += (this->tau() * PERF * this->knownDf(0, x)) * (T_A - vOld[0](x));
        vNew[0](x) += (this->tau() * SRC) * this->knownDf(1, x);
    }
    template <typename TT>
    void updateBorder(VDf<TT, DIM>&, VDf<TT, DIM> const&, Idx<DIM> const&, int const&)
    {
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

// ------------------------------------------------------------------ YEE
template <typename CT, std::size_t DIM>
class YeePec : public ScaFES::Problem<YeePec<CT, DIM>, CT, DIM> {
public:
    typedef ScaFES::Problem<YeePec<CT, DIM>, CT, DIM> Base;
    using Base::Base;
    // fields 0..5: Ex, Ey, Ez, Hx, Hy, Hz on (N+1)^3 nodes; Ex(i,j,k) sits at (i+1/2, j, k) ...
    int N() const { return this->nNodes(0) - 1; }
    double courant() const { return this->tau() / this->gridsize(0); }
    void evalInner(VDf<CT, DIM>&, Idx<DIM> const&, int const&) {}
    void evalBorder(VDf<CT, DIM>&, Idx<DIM> const&, int const&) {}
    template <typename TT>
    void initInner(VDf<TT, DIM>& v, std::vector<TT> const&, Idx<DIM> const& x, int const&)
    {
        const int n = N(), i = x[0], j = x[1], k = x[2];
        const bool ii = i >= 1 && i < n, jj = j >= 1 && j < n, kk = k >= 1 && k < n;
        v[0](x) = (i < n && jj && kk) ? pattern(x, 3, 5, 7, 11, 16.0) - 0.25 : 0.0;
        v[1](x) = (j < n && ii && kk) ? pattern(x, 5, 7, 3, 13, 16.0) - 0.25 : 0.0;
        v[2](x) = (k < n && ii && jj) ? pattern(x, 7, 3, 5, 9, 16.0) - 0.25 : 0.0;
        v[3](x) = (j < n && k < n) ? pattern(x, 1, 2, 3, 5, 64.0) : 0.0;
        v[4](x) = (i < n && k < n) ? pattern(x, 2, 3, 1, 7, 64.0) : 0.0;
        v[5](x) = (i < n && j < n) ? pattern(x, 3, 1, 2, 6, 64.0) : 0.0;
    }
    template <typename TT>
    void initBorder(VDf<TT, DIM>& v, std::vector<TT> const& p, Idx<DIM> const& x, int const& n)
    {
        initInner<TT>(v, p, x, n);
    }
    // H half step: H += s * (forward differences of E) where H is defined
    template <typename TT>
    void updateInner(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const&)
    {
        const int n = N();
        const double s = courant();
        const Idx<DIM> xp = this->connect(x, 1), yp = this->connect(x, 3), zp = this->connect(x, 5);
        if (x[1] < n && x[2] < n)
            vNew[3](x) = vOld[3](x) + s * ((vOld[1](zp) - vOld[1](x)) - (vOld[2](yp) - vOld[2](x)));
        if (x[0] < n && x[2] < n)
            vNew[4](x) = vOld[4](x) + s * ((vOld[2](xp) - vOld[2](x)) - (vOld[0](zp) - vOld[0](x)));
        if (x[0] < n && x[1] < n)
            vNew[5](x) = vOld[5](x) + s * ((vOld[0](yp) - vOld[0](x)) - (vOld[1](xp) - vOld[1](x)));
    }
    // E half step: E += s * (backward differences of H); tangential E on the walls stays
    template <typename TT>
    void updateInner2(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const&)
    {
        const int n = N();
        const double s = courant();
        const Idx<DIM> xm = this->connect(x, 0), ym = this->connect(x, 2), zm = this->connect(x, 4);
        const bool ii = x[0] >= 1 && x[0] < n, jj = x[1] >= 1 && x[1] < n, kk = x[2] >= 1 && x[2] < n;
        if (x[0] < n && jj && kk)
            vNew[0](x) = vOld[0](x) + s * ((vOld[5](x) - vOld[5](ym)) - (vOld[4](x) - vOld[4](zm)));
        if (x[1] < n && ii && kk)
            vNew[1](x) = vOld[1](x) + s * ((vOld[3](x) - vOld[3](zm)) - (vOld[5](x) - vOld[5](xm)));
        if (x[2] < n && ii && jj)
            vNew[2](x) = vOld[2](x) + s * ((vOld[4](x) - vOld[4](xm)) - (vOld[3](x) - vOld[3](ym)));
    }
    template <typename TT>
    void updateBorder(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const& n)
    {
        updateInner<TT>(vNew, vOld, x, n);
    }
    template <typename TT>
    void updateBorder2(VDf<TT, DIM>& vNew, VDf<TT, DIM> const& vOld, Idx<DIM> const& x, int const& n)
    {
        updateInner2<TT>(vNew, vOld, x, n);
    }
};

SCAFES_DEVICE_HOOKS_END

// ------------------------------------------------------------ device traits
// (left out with -DEXAMPLE_NO_TRAITS: the update phase then runs the classes'
// own hooks, which needs a hipcc build)
#ifndef EXAMPLE_NO_TRAITS
namespace ScaFES {
template <typename CT, std::size_t DIM>
struct DeviceTraits<HeatVar<CT, DIM>> {
    static const bool available = true;
    static void describe(const HeatVar<CT, DIM>&, DeviceUpdate& d)
    {
        d.kernel = DeviceKernel::HEAT7VAR;
        d.precision = DevicePrecision::EXAMPLE_PRECISION;
        d.unknown = 1;
        d.kappaKnown = 0;
        d.border = DeviceBorder::VALUE;
#ifdef EXAMPLE_BAD_TRAITS
        d.borderValue = 0.5;  // a wrong border value: the hooks keep the border at 0
#else
        d.borderValue = 0.0;
#endif
        d.knownTimeDependent = false;
    }
};
template <typename CT, std::size_t DIM>
struct DeviceTraits<PennesField<CT, DIM>> {
    static const bool available = true;
    static void describe(const PennesField<CT, DIM>& p, DeviceUpdate& d)
    {
        d.kernel = DeviceKernel::PENNES;
        d.precision = DevicePrecision::EXAMPLE_PRECISION;
        d.unknown = 2;
        d.perfusionKnown = 0;
        d.sourceFieldKnown = 1;
        d.cdiff = p.K;
        d.cperf = p.PERF;
        d.csrc = p.SRC;
        d.tA = p.T_A;
        d.knownTimeDependent = false;
    }
};
template <typename CT, std::size_t DIM>
struct DeviceTraits<YeePec<CT, DIM>> {
    static const bool available = true;
    static void describe(const YeePec<CT, DIM>& p, DeviceUpdate& d)
    {
        d.kernel = DeviceKernel::YEE;
        d.precision = DevicePrecision::EXAMPLE_PRECISION;
        d.courant = p.courant();
        d.energyEvery = 1;
        d.knownTimeDependent = false;
    }
};
}  // namespace ScaFES
#endif

static int runFamily(int argc, char* argv[]);

// Errors of the engine (a refused DeviceTraits, a device failure) end the
// program with status 1 and the message on stderr.
int main(int argc, char* argv[])
{
    try {
        return runFamily(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
}

static int runFamily(int argc, char* argv[])
{
    const std::size_t DIM = 3;
    ScaFES::Parameters params(argc, argv);
    ScaFES::GridGlobal<DIM> gg(params);
#if EXAMPLE_FAMILY == FAMILY_HEATVAR
    const std::vector<std::string> names = {"kappa", "U"};
    const std::vector<int> sw = {1, 1};
    const std::vector<bool> known = {true, false};
    const bool leapfrog = false;
    typedef HeatVar<double, DIM> Prb;
#elif EXAMPLE_FAMILY == FAMILY_PENNES
    const std::vector<std::string> names = {"w", "Q", "T"};
    const std::vector<int> sw = {1, 1, 1};
    const std::vector<bool> known = {true, true, false};
    const bool leapfrog = false;
    typedef PennesField<double, DIM> Prb;
#else
    const std::vector<std::string> names = {"Ex", "Ey", "Ez", "Hx", "Hy", "Hz"};
    const std::vector<int> sw(6, 1);
    const std::vector<bool> known(6, false);
    const bool leapfrog = true;
    typedef YeePec<double, DIM> Prb;
#endif
    const std::size_t nDf = names.size();
    // EXAMPLE_WRITE_ALL=1: every field written at the --nSnapshots intervals
    // (time-independent known fields included) to --outputfile
    const char* wa = std::getenv("EXAMPLE_WRITE_ALL");
    const ScaFES::WriteHowOften pol =
        (wa && wa[0] == '1') ? ScaFES::WriteHowOften::LIKE_GIVEN_AT_CL : ScaFES::WriteHowOften::NEVER;
    Prb ppp(params, gg, leapfrog, names, sw, known, std::vector<int>(nDf, 0), std::vector<double>(nDf, 0.0),
            std::vector<ScaFES::WriteHowOften>(nDf, pol), std::vector<bool>(nDf, false));
    ppp.iterateOverTime();
    for (std::size_t u = 0; u < ppp.vectUnknownDfsDomOld().size(); ++u)
        dumpNodes(ppp.vectUnknownDfsDomOld()[u], "." + std::to_string(u));
    for (double e : ppp.energyHistory())
        std::cout << "ENERGY " << std::scientific << std::setprecision(17) << e << std::endl;
    return 0;
}
