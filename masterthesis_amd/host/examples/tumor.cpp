// tumor.cpp -- the reference's TumorHeatEqnFDM example (examples/
// TumorHeatEqnFDM: TumorHeatEqnFDM.hpp, main1D/2D/3D.cpp) as a problem class
// against the host API: Pennes bioheat in brain tissue with a spherical
// tumor (its own perfusion and heat source), Neumann heat-flux borders and a
// convective (Cauchy) border on the far face of the last axis.  T is written
// to <--outputfile>.nc at the end and checked for convergence, as in main3D.
//
// The hooks below are the reference class's CPU form; DeviceTraits hands the
// update phase to the PENNES kernel with the TUMOR border, and the engine
// selects perfusion and source per node from the tumor sphere the same way.
// Parameters come from the [Parameters] section of --configfile.  Build
// variants: -DSCAFES_DIM=1|2|3.  SCAFES_DUMP=<path> dumps the final T.
#include <cmath>
#include <iostream>

#include "ScaFES.hpp"
#include "example_util.hpp"
#include "ini_reader.hpp"

#ifndef SCAFES_DIM
#define SCAFES_DIM 3
#endif

template <typename CT, std::size_t DIM>
class TumorHeatEqnFDM : public ScaFES::Problem<TumorHeatEqnFDM<CT, DIM>, CT, DIM> {
public:
    typedef ScaFES::Problem<TumorHeatEqnFDM<CT, DIM>, CT, DIM> Super;
    typedef ScaFES::Ntuple<int, DIM> Index;
    template <typename TT>
    using VDf = std::vector<ScaFES::DataField<TT, DIM>>;

    const CT RHO, C, K, Q_BRAIN, Q_TUMOR, RHO_B, C_PB, OMEGA_B_BRAIN, OMEGA_B_TUMOR, T_I, H, T_INF, Q_BC,
        DIAMETER, DEPTH, T_A;
    CT RADIUS;
    ScaFES::Ntuple<CT, DIM> tumorCenter;

    TumorHeatEqnFDM(ScaFES::Parameters const& params, ScaFES::GridGlobal<DIM> const& gg, bool useLeapfrog,
                    std::vector<std::string> const& nameDatafield, std::vector<int> const& stencilWidth,
                    std::vector<bool> const& isKnownDf, const Ini& ini, std::vector<int> const& nLayers,
                    std::vector<CT> const& defaultValue, std::vector<ScaFES::WriteHowOften> const& writeToFile,
                    std::vector<bool> const& computeError, std::vector<CT> const& geomparamsInit,
                    std::vector<bool> const& checkConvergence)
        : Super(params, gg, useLeapfrog, nameDatafield, stencilWidth, isKnownDf, nLayers, defaultValue, writeToFile,
                computeError, geomparamsInit, checkConvergence),
          RHO(ini.real("Parameters.RHO")), C(ini.real("Parameters.C")), K(ini.real("Parameters.K")),
          Q_BRAIN(ini.real("Parameters.Q_BRAIN")), Q_TUMOR(ini.real("Parameters.Q_TUMOR")),
          RHO_B(ini.real("Parameters.RHO_B")), C_PB(ini.real("Parameters.C_PB")),
          OMEGA_B_BRAIN(ini.real("Parameters.OMEGA_B_BRAIN")), OMEGA_B_TUMOR(ini.real("Parameters.OMEGA_B_TUMOR")),
          T_I(ini.real("Parameters.T_I")), H(ini.real("Parameters.H")), T_INF(ini.real("Parameters.T_INF")),
          Q_BC(ini.real("Parameters.Q_BC")), DIAMETER(ini.real("Parameters.DIAMETER")),
          DEPTH(ini.real("Parameters.DEPTH")), T_A(ini.real("Parameters.T_A"))
    {
        // tumor center (TumorHeatEqnFDM.hpp:150-185): centred on the other
        // axes, DEPTH below the far face of the last axis
        RADIUS = DIAMETER / 2.0;
        const std::vector<double>& last = this->params().coordNodeLast();
        for (std::size_t pp = 0; pp < DIM; ++pp) {
            const double lp = pp < last.size() ? last[pp] : 1.0;
            if (lp <= DIAMETER)
                std::cerr << "WARNING: Diameter of tumor is bigger than grid in dimension: " << pp << "." << std::endl;
            tumorCenter[pp] = pp == DIM - 1 ? lp - DEPTH : lp / 2.0;
        }
    }

    void evalInner(VDf<CT>&, Index const&, int const&) {}
    void evalBorder(VDf<CT>&, Index const&, int const&) {}
    template <typename TT>
    void initInner(VDf<TT>& vNew, std::vector<TT> const&, Index const& idx, int const&)
    {
        vNew[0](idx) = T_I;  // every node starts at T_I
    }
    template <typename TT>
    void initBorder(VDf<TT>& vNew, std::vector<TT> const& p, Index const& idx, int const& n)
    {
        initInner<TT>(vNew, p, idx, n);
    }
    // omega, Q of a node (TumorHeatEqnFDM.hpp:239-254)
    void material(Index const& idx, CT& omega, CT& Q) const
    {
        ScaFES::Ntuple<double, DIM> x = this->coordinates(idx);
        CT distance = 0.0;
        for (std::size_t pp = 0; pp < DIM; ++pp) distance += (x[pp] - tumorCenter[pp]) * (x[pp] - tumorCenter[pp]);
        const bool in = distance <= (RADIUS * RADIUS);
        omega = in ? OMEGA_B_TUMOR : OMEGA_B_BRAIN;
        Q = in ? Q_TUMOR : Q_BRAIN;
    }
    // TumorHeatEqnFDM.hpp:233-269
    template <typename TT>
    void updateInner(VDf<TT>& vNew, VDf<TT> const& vOld, Index const& idx, int const&)
    {
        CT omega, Q;
        material(idx, omega, Q);
        vNew[0](idx) = vOld[0](idx);
        for (std::size_t pp = 0; pp < DIM; ++pp)
            vNew[0](idx) += this->tau() * (K / (RHO * C)) *
                            (vOld[0](this->connect(idx, 2 * pp)) + vOld[0](this->connect(idx, 2 * pp + 1)) -
                             2.0 * vOld[0](idx)) /
                            (this->gridsize(pp) * this->gridsize(pp));
        vNew[0](idx) += this->tau() * ((RHO_B * C_PB) / (RHO * C)) * omega * (T_A - vOld[0](idx));
        vNew[0](idx) += this->tau() * (1.0 / (RHO * C)) * Q;
    }
    // TumorHeatEqnFDM.hpp:272-349
    template <typename TT>
    void updateBorder(VDf<TT>& vNew, VDf<TT> const& vOld, Index const& idx, int const&)
    {
        CT omega, Q;
        material(idx, omega, Q);
        const TT u = vOld[0](idx);
        vNew[0](idx) = u;
        for (std::size_t pp = 0; pp < DIM; ++pp) {
            const double a = 2.0 * this->gridsize(pp) / K;
            TT t;
            if (idx.elem(pp) == this->nNodes(pp) - 1) {
                const TT w = vOld[0](this->connect(idx, 2 * pp));
                if (pp == DIM - 1) t = w + w - (a * H * (u - T_INF)) - 2.0 * u;
                else t = w + w - (a * (-1.0 * Q_BC)) - 2.0 * u;
            } else if (idx.elem(pp) == 0) {
                const TT e = vOld[0](this->connect(idx, 2 * pp + 1));
                t = e + (a * Q_BC) + e - 2.0 * u;
            } else {
                t = vOld[0](this->connect(idx, 2 * pp)) + vOld[0](this->connect(idx, 2 * pp + 1)) - 2.0 * u;
            }
            vNew[0](idx) += this->tau() * (K / (RHO * C)) * t / (this->gridsize(pp) * this->gridsize(pp));
        }
        vNew[0](idx) += this->tau() * ((RHO_B * C_PB) / (RHO * C)) * omega * (T_A - u);
        vNew[0](idx) += this->tau() * (1.0 / (RHO * C)) * Q;
    }
    template <typename TT>
    void updateInner2(VDf<TT>&, VDf<TT> const&, Index const&, int const&)
    {
    }
    template <typename TT>
    void updateBorder2(VDf<TT>&, VDf<TT> const&, Index const&, int const&)
    {
    }
};

namespace ScaFES {
template <typename CT, std::size_t DIM>
struct DeviceTraits<TumorHeatEqnFDM<CT, DIM>> {
    static const bool available = true;
    static void describe(const TumorHeatEqnFDM<CT, DIM>& p, DeviceUpdate& d)
    {
        d.kernel = DeviceKernel::PENNES;
        d.unknown = 0;
        d.cdiff = p.K / (p.RHO * p.C);
        d.cperf = (p.RHO_B * p.C_PB) / (p.RHO * p.C);
        d.csrc = 1.0 / (p.RHO * p.C);
        d.tA = p.T_A;
        d.border = DeviceBorder::TUMOR;
        d.kCond = p.K;
        d.hConv = p.H;
        d.tInf = p.T_INF;
        d.qBc = p.Q_BC;
        d.tumorFields = true;
        for (std::size_t q = 0; q < DIM; ++q) d.tumorCenter[q] = p.tumorCenter[q];
        d.tumorRadius = p.RADIUS;
        d.omegaIn = p.OMEGA_B_TUMOR;
        d.omegaOut = p.OMEGA_B_BRAIN;
        d.qIn = p.Q_TUMOR;
        d.qOut = p.Q_BRAIN;
        d.knownTimeDependent = false;
    }
};
}  // namespace ScaFES

// main1D/2D/3D.cpp: one field T, written at the end, convergence checked
int main(int argc, char* argv[])
{
    const std::size_t DIM = SCAFES_DIM;
    ScaFES::Parameters paramsCl(argc, argv);
    ScaFES::GridGlobal<DIM> gg(paramsCl);
    std::vector<std::string> nameDatafield(1, "T");
    std::vector<int> stencilWidth(1, 1);
    std::vector<bool> isKnownDf(1, false);
    std::vector<int> nLayers(1, 0);
    std::vector<double> defaultValue(1, 0.0);
    std::vector<ScaFES::WriteHowOften> writeToFile(1, ScaFES::WriteHowOften::AT_END);
    std::vector<bool> computeError(1, false);
    std::vector<bool> checkConvergence(1, true);
    std::vector<double> geomparamsInit;
    Ini ini(paramsCl.nameConfigFile());
    TumorHeatEqnFDM<double, DIM> ppp(paramsCl, gg, false, nameDatafield, stencilWidth, isKnownDf, ini, nLayers,
                                     defaultValue, writeToFile, computeError, geomparamsInit, checkConvergence);
    ppp.iterateOverTime();
    dumpNodes(ppp.vectUnknownDfsDomOld()[0]);
    std::cout << "LAST_ITER " << ppp.lastTimeIter() << std::endl;
    return 0;
}
