// mri.cpp -- the reference's MRIDataPBHEqnFDM example (examples/
// MRIDataPBHEqnFDM: MRIDataPBHEqnFDM.hpp, main3D.cpp) as a problem class
// against the host API: Pennes bioheat in the brain with nine per-node
// material fields (rho, c, lambda, rho_blood, c_blood, omega, T_blood, q,
// surface) read with the temperature T from --initfile (--initKnownDfs=1),
// Neumann heat-flux borders, and on the far face of the last axis an open
// skull (surface == 1: convection and radiation) or a closed one (flux
// Q_SKULL).  T is written to <--outputfile>.nc at the end and checked for
// convergence, as in main3D.
//
// The hooks below are the reference class's CPU form; DeviceTraits hands the
// update phase to the MRI kernel.  Constants come from the [Parameters]
// section of --configfile.  Build variants: -DSCAFES_DIM=1|2|3 (main3D is
// three-dimensional).  SCAFES_DUMP=<path> dumps the final T.
#include <iostream>

#include "ScaFES.hpp"
#include "example_util.hpp"
#include "ini_reader.hpp"

#ifndef SCAFES_DIM
#define SCAFES_DIM 3
#endif

template <typename CT, std::size_t DIM>
class MRIDataPBHEqnFDM : public ScaFES::Problem<MRIDataPBHEqnFDM<CT, DIM>, CT, DIM> {
public:
    typedef ScaFES::Problem<MRIDataPBHEqnFDM<CT, DIM>, CT, DIM> Super;
    typedef ScaFES::Ntuple<int, DIM> Index;
    template <typename TT>
    using VDf = std::vector<ScaFES::DataField<TT, DIM>>;

    const CT H, T_AMB, Q_BC, Q_SKULL, EPSILON;
    const CT SIGMA = 5.670373e-8;  // MRIDataPBHEqnFDM.hpp:50

    MRIDataPBHEqnFDM(ScaFES::Parameters const& params, ScaFES::GridGlobal<DIM> const& gg, bool useLeapfrog,
                     std::vector<std::string> const& nameDatafield, std::vector<int> const& stencilWidth,
                     std::vector<bool> const& isKnownDf, const Ini& ini, std::vector<int> const& nLayers,
                     std::vector<CT> const& defaultValue, std::vector<ScaFES::WriteHowOften> const& writeToFile,
                     std::vector<bool> const& computeError, std::vector<CT> const& geomparamsInit,
                     std::vector<bool> const& checkConvergence)
        : Super(params, gg, useLeapfrog, nameDatafield, stencilWidth, isKnownDf, nLayers, defaultValue, writeToFile,
                computeError, geomparamsInit, checkConvergence),
          H(ini.real("Parameters.H")), T_AMB(ini.real("Parameters.T_INF")), Q_BC(ini.real("Parameters.Q_BC")),
          Q_SKULL(ini.real("Parameters.Q_SKULL")), EPSILON(ini.real("Parameters.EPSILON"))
    {
    }

    // known fields come from the init file; the hooks leave them alone
    void evalInner(VDf<CT>&, Index const&, int const&) {}
    void evalBorder(VDf<CT>&, Index const&, int const&) {}
    template <typename TT>
    void initInner(VDf<TT>&, std::vector<TT> const&, Index const&, int const&)
    {
    }
    template <typename TT>
    void initBorder(VDf<TT>&, std::vector<TT> const&, Index const&, int const&)
    {
    }
    // MRIDataPBHEqnFDM.hpp:143-168
    template <typename TT>
    void updateInner(VDf<TT>& vNew, VDf<TT> const& vOld, Index const& idx, int const&)
    {
        const CT rho = this->knownDf(0, idx), c = this->knownDf(1, idx), k = this->knownDf(2, idx);
        const CT rho_blood = this->knownDf(3, idx), c_blood = this->knownDf(4, idx), omega = this->knownDf(5, idx);
        const CT T_blood = this->knownDf(6, idx), q = this->knownDf(7, idx);
        vNew[0](idx) = vOld[0](idx);
        for (std::size_t pp = 0; pp < DIM; ++pp)
            vNew[0](idx) += this->tau() * (k / (rho * c)) *
                            (vOld[0](this->connect(idx, 2 * pp)) + vOld[0](this->connect(idx, 2 * pp + 1)) -
                             2.0 * vOld[0](idx)) /
                            (this->gridsize(pp) * this->gridsize(pp));
        vNew[0](idx) += this->tau() * ((rho_blood * c_blood) / (rho * c)) * omega * (T_blood - vOld[0](idx));
        vNew[0](idx) += this->tau() * (1.0 / (rho * c)) * q;
    }
    // MRIDataPBHEqnFDM.hpp:176-273
    template <typename TT>
    void updateBorder(VDf<TT>& vNew, VDf<TT> const& vOld, Index const& idx, int const&)
    {
        const CT rho = this->knownDf(0, idx), c = this->knownDf(1, idx), k = this->knownDf(2, idx);
        const CT rho_blood = this->knownDf(3, idx), c_blood = this->knownDf(4, idx), omega = this->knownDf(5, idx);
        const CT T_blood = this->knownDf(6, idx), q = this->knownDf(7, idx);
        const int trepanationArea = this->knownDf(8, idx);
        const TT u = vOld[0](idx);
        vNew[0](idx) = u;
        for (std::size_t pp = 0; pp < DIM; ++pp) {
            const double a = 2.0 * this->gridsize(pp) / k;
            TT t;
            if (idx.elem(pp) == this->nNodes(pp) - 1) {
                const TT w = vOld[0](this->connect(idx, 2 * pp));
                if (pp == DIM - 1) {
                    if (trepanationArea == 1) {
                        // open skull: the radiated T^4 at the value accumulated so far
                        const CT tempOld = vNew[0](idx) + 273.15;
                        const CT tempOldPow4 = tempOld * tempOld * tempOld * tempOld;
                        const CT tempAmb = T_AMB + 273.15;
                        const CT tempAmbPow4 = tempAmb * tempAmb * tempAmb * tempAmb;
                        t = w + w - (a * H * (u - T_AMB)) - (a * EPSILON * SIGMA * (tempOldPow4 - tempAmbPow4)) -
                            2.0 * u;
                    } else {
                        t = w + w - (a * (-1.0 * Q_SKULL)) - 2.0 * u;
                    }
                } else {
                    t = w + w - (a * (-1.0 * Q_BC)) - 2.0 * u;
                }
            } else if (idx.elem(pp) == 0) {
                const TT e = vOld[0](this->connect(idx, 2 * pp + 1));
                t = e + (a * Q_BC) + e - 2.0 * u;
            } else {
                t = vOld[0](this->connect(idx, 2 * pp)) + vOld[0](this->connect(idx, 2 * pp + 1)) - 2.0 * u;
            }
            vNew[0](idx) += this->tau() * (k / (rho * c)) * t / (this->gridsize(pp) * this->gridsize(pp));
        }
        vNew[0](idx) += this->tau() * ((rho_blood * c_blood) / (rho * c)) * omega * (T_blood - u);
        vNew[0](idx) += this->tau() * (1.0 / (rho * c)) * q;
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
struct DeviceTraits<MRIDataPBHEqnFDM<CT, DIM>> {
    static const bool available = true;
    static void describe(const MRIDataPBHEqnFDM<CT, DIM>& p, DeviceUpdate& d)
    {
        d.kernel = DeviceKernel::MRI;
        d.unknown = 9;
        for (int q = 0; q < 9; ++q) d.mri[q] = q;
        d.hConv = p.H;
        d.tInf = p.T_AMB;
        d.qBc = p.Q_BC;
        d.qSkull = p.Q_SKULL;
        d.epsilon = p.EPSILON;
        d.sigma = p.SIGMA;
        d.knownTimeDependent = false;
    }
};
}  // namespace ScaFES

// main3D.cpp: nine known material fields and T; T written at the end,
// convergence checked on T
int main(int argc, char* argv[])
{
    const std::size_t DIM = SCAFES_DIM;
    ScaFES::Parameters paramsCl(argc, argv);
    ScaFES::GridGlobal<DIM> gg(paramsCl);
    std::vector<std::string> nameDatafield = {"rho",   "c",       "lambda", "rho_blood", "c_blood",
                                              "omega", "T_blood", "q",      "surface",   "T"};
    std::vector<int> stencilWidth(10, 0);
    stencilWidth[9] = 1;
    std::vector<bool> isKnownDf(10, true);
    isKnownDf[9] = false;
    std::vector<int> nLayers(10, 0);
    std::vector<double> defaultValue(10, 0.0);
    std::vector<ScaFES::WriteHowOften> writeToFile(10, ScaFES::WriteHowOften::NEVER);
    writeToFile[9] = ScaFES::WriteHowOften::AT_END;
    std::vector<bool> computeError(10, false);
    std::vector<bool> checkConvergence(10, false);
    checkConvergence[9] = true;
    std::vector<double> geomparamsInit;
    Ini ini(paramsCl.nameConfigFile());
    MRIDataPBHEqnFDM<double, DIM> ppp(paramsCl, gg, false, nameDatafield, stencilWidth, isKnownDf, ini, nLayers,
                                      defaultValue, writeToFile, computeError, geomparamsInit, checkConvergence);
    ppp.iterateOverTime();
    dumpNodes(ppp.vectUnknownDfsDomOld()[0]);
    std::cout << "LAST_ITER " << ppp.lastTimeIter() << std::endl;
    return 0;
}
