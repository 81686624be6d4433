// heat_timelin.cpp -- the TimeLinSpaceLin system problems of the reference
// (heat: tests/system/HeatEqnFDM/Dirichlet/TimeLinSpaceLin, Pennes:
// tests/system/PennesBioheatEqnFDM/Dirichlet/TimeLinSpaceLin) written as
// problem classes against the host API, in the reference's CRTP layering
// (equation class -> boundary class -> concrete problem).  Exact discrete
// solution y = (1+t)(1+sum x); every "Local Linf" line must be round-off.
//
// Build variants: -DSCAFES_DIM=1|2|3, -DEXAMPLE_PENNES_FORM for the Pennes form,
// -DEXAMPLE_BORDER=1|2|3 for the Neumann / Cauchy / radiation border classes
// (tests/system/*/{Neumann,Cauchy,Radiation}/TimeLinSpaceLin; default 0 =
// Dirichlet), -DEXAMPLE_DEVICE_KNOWN to have the engine evaluate F, G and
// the exact solution on the device after step 0 (the hooks below then run
// only at n = 0), -DEXAMPLE_WRITE for snapshot output policies (--outputfile,
// --nSnapshots).
#include <iostream>

#include "ScaFES.hpp"
#include "example_util.hpp"

#ifndef SCAFES_DIM
#define SCAFES_DIM 3
#endif
#ifndef EXAMPLE_BORDER
#define EXAMPLE_BORDER 0
#endif

// Equation layer: explicit Euler + 7-point stencil (HeatEqnFDM.hpp:204-218;
// with PENNES also the perfusion term of PennesBioheatEqnFDM.hpp:153-168).
template <typename CT, std::size_t DIM, typename Class>
class Equation : public ScaFES::Problem<Equation<CT, DIM, Class>, CT, DIM> {
public:
    typedef ScaFES::Problem<Equation<CT, DIM, Class>, CT, DIM> Base;
    const double LAMBDA = 1.0, RHO = 1.0, C = 1.0, RHO_BLOOD = 1.0, C_BLOOD = 1.0, W = 1.0;
    const double ALPHA = 1.0, EPSILON = 1.0, SIGMA = 1.0;
    const double COEFF_A = LAMBDA / (RHO * C);

    Equation(ScaFES::Parameters const& params, ScaFES::GridGlobal<DIM> const& gg, bool useLeapfrog,
             std::vector<std::string> const& names, std::vector<int> const& sw, std::vector<bool> const& known,
             std::vector<int> const& nLayers, std::vector<CT> const& defaults,
             std::vector<ScaFES::WriteHowOften> const& write, std::vector<bool> const& err)
        : Base(params, gg, useLeapfrog, names, sw, known, nLayers, defaults, write, err)
    {
    }
    void evalInner(std::vector<ScaFES::DataField<CT, DIM>>& v, ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->evalInner(v, x, n);
    }
    void evalBorder(std::vector<ScaFES::DataField<CT, DIM>>& v, ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->evalBorder(v, x, n);
    }
    template <typename TT>
    void initInner(std::vector<ScaFES::DataField<TT, DIM>>& v, std::vector<TT> const& p,
                   ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->template initInner<TT>(v, p, x, n);
    }
    template <typename TT>
    void initBorder(std::vector<ScaFES::DataField<TT, DIM>>& v, std::vector<TT> const& p,
                    ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->template initBorder<TT>(v, p, x, n);
    }
    // The per-node form of the update the device kernel HEAT7 performs.
    template <typename TT>
    void updateInner(std::vector<ScaFES::DataField<TT, DIM>>& vNew,
                     std::vector<ScaFES::DataField<TT, DIM>> const& vOld, ScaFES::Ntuple<int, DIM> const& x,
                     int const&)
    {
        vNew[0](x) = vOld[0](x);
        for (std::size_t p = 0; p < DIM; ++p)
            vNew[0](x) += this->tau() * COEFF_A *
                          (vOld[0](this->connect(x, 2 * p)) + vOld[0](this->connect(x, 2 * p + 1)) - 2.0 * vOld[0](x)) /
                          (this->gridsize(p) * this->gridsize(p));
#ifdef EXAMPLE_PENNES_FORM
        vNew[0](x) -= this->tau() * ((RHO_BLOOD * C_BLOOD) / (RHO * C)) * W * vOld[0](x);
#endif
        vNew[0](x) += this->tau() * (1.0 / (RHO * C)) * this->knownDfOld(0, x);
    }
    template <typename TT>
    void updateBorder(std::vector<ScaFES::DataField<TT, DIM>>& vNew, std::vector<ScaFES::DataField<TT, DIM>> const& vOld,
                      ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->template updateBorder<TT>(vNew, vOld, x, n);
    }
    template <typename TT>
    void updateInner2(std::vector<ScaFES::DataField<TT, DIM>>&, std::vector<ScaFES::DataField<TT, DIM>> const&,
                      ScaFES::Ntuple<int, DIM> const&, int const&)
    {
    }
    template <typename TT>
    void updateBorder2(std::vector<ScaFES::DataField<TT, DIM>>&, std::vector<ScaFES::DataField<TT, DIM>> const&,
                       ScaFES::Ntuple<int, DIM> const&, int const&)
    {
    }
};

// Boundary layer: Dirichlet, u = G at border nodes (Dirichlet.hpp:124-130).
template <typename CT, std::size_t DIM, typename Class>
class DirichletBorder : public Equation<CT, DIM, DirichletBorder<CT, DIM, Class>> {
public:
    typedef Equation<CT, DIM, DirichletBorder<CT, DIM, Class>> Base;
    using Base::Base;
    void evalInner(std::vector<ScaFES::DataField<CT, DIM>>& v, ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->evalInner(v, x, n);
    }
    void evalBorder(std::vector<ScaFES::DataField<CT, DIM>>& v, ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->evalBorder(v, x, n);
    }
    template <typename TT>
    void initInner(std::vector<ScaFES::DataField<TT, DIM>>& v, std::vector<TT> const& p,
                   ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->template initInner<TT>(v, p, x, n);
    }
    template <typename TT>
    void initBorder(std::vector<ScaFES::DataField<TT, DIM>>& v, std::vector<TT> const& p,
                    ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->template initBorder<TT>(v, p, x, n);
    }
    template <typename TT>
    void updateBorder(std::vector<ScaFES::DataField<TT, DIM>>& vNew, std::vector<ScaFES::DataField<TT, DIM>> const&,
                      ScaFES::Ntuple<int, DIM> const& x, int const&)
    {
        vNew[0](x) = this->knownDf(1, x);
    }
};

// Boundary layer of the flux classes (Neumann.hpp / Cauchy.hpp /
// Radiation.hpp:126-175): G_new on faces with an index 0, otherwise the
// outer neighbour on far faces replaced by the inner one plus the flux term
// built from G_old.
template <typename CT, std::size_t DIM, typename Class>
class FluxBorder : public Equation<CT, DIM, FluxBorder<CT, DIM, Class>> {
public:
    typedef Equation<CT, DIM, FluxBorder<CT, DIM, Class>> Base;
    using Base::Base;
    void evalInner(std::vector<ScaFES::DataField<CT, DIM>>& v, ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->evalInner(v, x, n);
    }
    void evalBorder(std::vector<ScaFES::DataField<CT, DIM>>& v, ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->evalBorder(v, x, n);
    }
    template <typename TT>
    void initInner(std::vector<ScaFES::DataField<TT, DIM>>& v, std::vector<TT> const& p,
                   ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->template initInner<TT>(v, p, x, n);
    }
    template <typename TT>
    void initBorder(std::vector<ScaFES::DataField<TT, DIM>>& v, std::vector<TT> const& p,
                    ScaFES::Ntuple<int, DIM> const& x, int const& n)
    {
        static_cast<Class*>(this)->template initBorder<TT>(v, p, x, n);
    }
    template <typename TT>
    void updateBorder(std::vector<ScaFES::DataField<TT, DIM>>& vNew, std::vector<ScaFES::DataField<TT, DIM>> const& vOld,
                      ScaFES::Ntuple<int, DIM> const& x, int const&)
    {
        for (std::size_t p = 0; p < DIM; ++p)
            if (x.elem(p) == 0) {
                vNew[0](x) = this->knownDf(1, x);
                return;
            }
        const TT u = vOld[0](x);
        vNew[0](x) = u;
        for (std::size_t p = 0; p < DIM; ++p) {
            const TT w = vOld[0](this->connect(x, 2 * p));
            const double a = 2.0 * this->gridsize(p) / this->LAMBDA;
            TT t;
            if (x.elem(p) == this->nNodes(p) - 1) {
#if EXAMPLE_BORDER == 1
                t = w + w - (a * this->knownDfOld(1, x)) - 2.0 * u;
#elif EXAMPLE_BORDER == 2
                t = w + w - (a * this->ALPHA * u) + (a * this->knownDfOld(1, x)) - 2.0 * u;
#else
                t = w + w - (a * this->EPSILON * this->SIGMA * u * u * u * u) + (a * this->knownDfOld(1, x)) - 2.0 * u;
#endif
            } else {
                t = w + vOld[0](this->connect(x, 2 * p + 1)) - 2.0 * u;
            }
            vNew[0](x) += this->tau() * this->COEFF_A * t / (this->gridsize(p) * this->gridsize(p));
        }
#ifdef EXAMPLE_PENNES_FORM
        vNew[0](x) -= this->tau() * ((this->RHO_BLOOD * this->C_BLOOD) / (this->RHO * this->C)) * this->W * u;
#endif
        vNew[0](x) += this->tau() * (1.0 / (this->RHO * this->C)) * this->knownDfOld(0, x);
    }
};

#if EXAMPLE_BORDER == 0
template <typename CT, std::size_t DIM, typename Class>
using BorderLayer = DirichletBorder<CT, DIM, Class>;
#else
template <typename CT, std::size_t DIM, typename Class>
using BorderLayer = FluxBorder<CT, DIM, Class>;
#endif

// Concrete problem: y = (1+t)(1+sum x).
template <typename CT, std::size_t DIM>
class TimeLinSpaceLin : public BorderLayer<CT, DIM, TimeLinSpaceLin<CT, DIM>> {
public:
    typedef BorderLayer<CT, DIM, TimeLinSpaceLin<CT, DIM>> Base;
    using Base::Base;
    void evalInner(std::vector<ScaFES::DataField<CT, DIM>>& v, ScaFES::Ntuple<int, DIM> const& idx, int const& n)
    {
        ScaFES::Ntuple<double, DIM> x = this->coordinates(idx);
        const double t = this->time(n);
        v[0](idx) = 1.0;
        for (std::size_t p = 0; p < DIM; ++p) v[0](idx) += x[p];
        v[0](idx) *= this->RHO * this->C;
#ifdef EXAMPLE_PENNES_FORM
        double tmp = 1.0;
        for (std::size_t p = 0; p < DIM; ++p) tmp += x[p];
        tmp *= (1.0 + t);
        v[0](idx) += this->RHO_BLOOD * this->C_BLOOD * this->W * tmp;
#endif
        v[1](idx) = 0.0;
        v[2](idx) = 1.0;
        for (std::size_t p = 0; p < DIM; ++p) v[2](idx) += x[p];
        v[2](idx) *= (1.0 + t);
    }
    void evalBorder(std::vector<ScaFES::DataField<CT, DIM>>& v, ScaFES::Ntuple<int, DIM> const& idx, int const& n)
    {
        ScaFES::Ntuple<double, DIM> x = this->coordinates(idx);
        const double t = this->time(n);
#if EXAMPLE_BORDER == 0
        v[0](idx) = 0.0;
        v[1](idx) = 1.0;
        for (std::size_t p = 0; p < DIM; ++p) v[1](idx) += x[p];
        v[1](idx) *= (1.0 + t);
        v[2](idx) = v[1](idx);
#else
        // F as on inner nodes; G = Dirichlet value on faces with an index 0,
        // else the flux datum (.../{Neumann,Cauchy,Radiation}/TimeLinSpaceLin)
        evalInner(v, idx, n);
        bool near = false;
        for (std::size_t p = 0; p < DIM; ++p) near = near || idx.elem(p) == 0;
        if (near) {
            v[1](idx) = v[2](idx);
            return;
        }
#if EXAMPLE_BORDER == 1
        (void)x;
        v[1](idx) = -1.0 * this->LAMBDA * (1.0 + t);
#elif EXAMPLE_BORDER == 2
        v[1](idx) = 1.0;
        for (std::size_t p = 0; p < DIM; ++p) v[1](idx) += x[p];
        v[1](idx) *= this->ALPHA * (1.0 + t);
        v[1](idx) += this->LAMBDA * (1.0 + t);
#else
        v[1](idx) = 1.0;
        for (std::size_t p = 0; p < DIM; ++p) v[1](idx) += x[p];
        v[1](idx) *= (1.0 + t);
        v[1](idx) = v[1](idx) * v[1](idx) * v[1](idx) * v[1](idx);
        v[1](idx) *= this->SIGMA * this->EPSILON;
        v[1](idx) += this->LAMBDA * (1.0 + t);
#endif
#endif
    }
    template <typename TT>
    void initInner(std::vector<ScaFES::DataField<TT, DIM>>& v, std::vector<TT> const&, ScaFES::Ntuple<int, DIM> const& idx,
                   int const& n)
    {
        ScaFES::Ntuple<double, DIM> x = this->coordinates(idx);
        v[0](idx) = 1.0;
        for (std::size_t p = 0; p < DIM; ++p) v[0](idx) += x[p];
        v[0](idx) *= (1.0 + this->time(n));
    }
    template <typename TT>
    void initBorder(std::vector<ScaFES::DataField<TT, DIM>>& v, std::vector<TT> const& p,
                    ScaFES::Ntuple<int, DIM> const& idx, int const& n)
    {
        this->template initInner<TT>(v, p, idx, n);
    }
};

// Device kernel of the equation class (the OWNPRBLM of Problem): HEAT7 with
// source F (field 0, old buffer) and Dirichlet values G (field 1, new
// buffer) as the equation/boundary layers read them; the unknown is field 2.
namespace ScaFES {
template <typename CT, std::size_t DIM, typename Class>
struct DeviceTraits<Equation<CT, DIM, Class>> {
    static const bool available = true;
    static void describe(const Equation<CT, DIM, Class>& p, ScaFES::DeviceUpdate& d)
    {
        d.kernel = ScaFES::DeviceKernel::HEAT7;
        d.unknown = 2;
        d.coeffA = p.COEFF_A;
        d.rhoCInv = 1.0 / (p.RHO * p.C);
#ifdef EXAMPLE_PENNES_FORM
        d.perfusion = true;
        d.perfC = (p.RHO_BLOOD * p.C_BLOOD) / (p.RHO * p.C);
        d.perfW = p.W;
#endif
        d.sourceKnown = 0;
#if EXAMPLE_BORDER == 0
        d.border = ScaFES::DeviceBorder::KNOWN_FIELD;
#elif EXAMPLE_BORDER == 1
        d.border = ScaFES::DeviceBorder::NEUMANN;
#elif EXAMPLE_BORDER == 2
        d.border = ScaFES::DeviceBorder::CAUCHY;
#else
        d.border = ScaFES::DeviceBorder::RADIATION;
#endif
        d.lambda = p.LAMBDA;
        d.alpha = p.ALPHA;
        d.epsilon = p.EPSILON;
        d.sigma = p.SIGMA;
        d.borderKnown = 1;
        d.knownTimeDependent = true;
#ifdef EXAMPLE_DEVICE_KNOWN
#ifdef EXAMPLE_PENNES_FORM
        d.knownGen = ScaFES::DeviceKnownGen::PENNES_TIMELIN;
#else
        d.knownGen = ScaFES::DeviceKnownGen::HEAT_TIMELIN;
#endif
        d.rhoC = p.RHO * p.C;
        d.rhobCbW = p.RHO_BLOOD * p.C_BLOOD * p.W;
#endif
    }
};
}  // namespace ScaFES

int main(int argc, char* argv[])
{
    const std::size_t DIM = SCAFES_DIM;
    ScaFES::Parameters params(argc, argv);
    ScaFES::GridGlobal<DIM> gg(params);
    std::vector<std::string> names = {"F", "G", "U"};
    std::vector<int> sw = {0, 0, 1};
    std::vector<bool> known = {true, true, false};
    std::vector<int> nLayers(3, 0);
    std::vector<double> defaults(3, 0.0);
    std::vector<ScaFES::WriteHowOften> write(3, ScaFES::WriteHowOften::NEVER);
#ifdef EXAMPLE_WRITE
    // snapshots: G at the start, U at the --nSnapshots intervals (F never)
    write[1] = ScaFES::WriteHowOften::AT_START;
    write[2] = ScaFES::WriteHowOften::LIKE_GIVEN_AT_CL;
#endif
    std::vector<bool> err = {false, false, true};
    TimeLinSpaceLin<double, DIM> ppp(params, gg, false, names, sw, known, nLayers, defaults, write, err);
    double sumH = 0.0;
    for (std::size_t p = 0; p < DIM; ++p) sumH += 1.0 / (ppp.gridsize(p) * ppp.gridsize(p));
    if ((1.0 / sumH) < 2.0 * ppp.tau()) std::cerr << "\nERROR: Stability condition is not fulfilled." << std::endl;
    ppp.iterateOverTime();
    dumpNodes(ppp.vectUnknownDfsDomOld()[0]);
    double emax = 0.0;
    for (double e : ppp.errorHistory()) emax = e > emax ? e : emax;
    std::cout << "MAX_LINF " << std::scientific << emax << std::endl;
    return 0;
}
