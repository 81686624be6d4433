// system_tests.cpp -- driver of the exact-arithmetic system tests
// (system_problems.hpp) in the reference's main1D/2D/3D form
// (tests/system/*/main3D.cpp): every field unknown, computeError on, no
// file output; --dim picks the dimension.  Built with hipcc: the problem
// header sits between SCAFES_DEVICE_HOOKS_BEGIN / _END, so without a
// DeviceTraits specialisation the update phase runs the classes' own hooks
// on the GPU.
//
// -DEXAMPLE_SYSTEM=1 StencilOne (one field, stencil width 1),
//                  2 StencilVar (three fields, stencil widths 0, 1, 2),
//                  3 Propagation (three fields, stencil widths 0, 1, 2).
// The final fields are dumped to $SCAFES_DUMP (suffix = field index).
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ScaFES.hpp"
#include "example_util.hpp"

SCAFES_DEVICE_HOOKS_BEGIN
#include "system_problems.hpp"
SCAFES_DEVICE_HOOKS_END

#ifndef EXAMPLE_SYSTEM
#define EXAMPLE_SYSTEM 1
#endif

template <std::size_t DIM>
static int runDim(ScaFES::Parameters& params)
{
    ScaFES::GridGlobal<DIM> gg(params);
    const int nDf = EXAMPLE_SYSTEM == 1 ? 1 : 3;
    std::vector<std::string> names;
    std::vector<int> sw;
    for (int k = 0; k < nDf; ++k) {
        names.push_back(std::string(1, (char)('A' + k)));
        sw.push_back(EXAMPLE_SYSTEM == 1 ? 1 : k);
    }
    const std::vector<bool> known(nDf, false), err(nDf, true);
    const std::vector<int> layers(nDf, 0);
    const std::vector<double> defaults(nDf, 0.0);
    const std::vector<ScaFES::WriteHowOften> write(nDf, ScaFES::WriteHowOften::NEVER);
#if EXAMPLE_SYSTEM == 1
    systests::StencilOne<double, DIM> ppp(params, gg, false, names, sw, known, layers, defaults, write, err);
#elif EXAMPLE_SYSTEM == 2
    systests::StencilVar<double, DIM> ppp(params, gg, false, names, sw, known, layers, defaults, write, err);
#else
    systests::Propagation<double, DIM> ppp(params, gg, false, names, sw, known, layers, defaults, write, err);
#endif
    ppp.iterateOverTime();
    for (std::size_t u = 0; u < ppp.vectUnknownDfsDomOld().size(); ++u)
        dumpNodes(ppp.vectUnknownDfsDomOld()[u], "." + std::to_string(u));
    return 0;
}

int main(int argc, char* argv[])
{
    ScaFES::Parameters params(argc, argv);
    try {
        switch (params.dim()) {
        case 1: return runDim<1>(params);
        case 2: return runDim<2>(params);
        default: return runDim<3>(params);
        }
    } catch (const std::exception& e) {
        std::cerr << e.what() << std::endl;
        return 1;
    }
}
