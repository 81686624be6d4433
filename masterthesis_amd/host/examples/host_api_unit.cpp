// host_api_unit.cpp -- CPU checks of the host API's index math (no GPU):
// the numbering and neighbour semantics the reference's unit tests pin
// (tests/unit/ScaFES_GridTest*.cpp, ScaFES_NtupleTest.cpp, ScaFES_GridSubTest.cpp).
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "ScaFES_DataField.hpp"
#include "ScaFES_GridGlobal.hpp"

static int g_checks = 0, g_fail = 0;
#define CHECK(c)                                                          \
    do {                                                                  \
        ++g_checks;                                                       \
        if (!(c)) {                                                       \
            ++g_fail;                                                     \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
        }                                                                 \
    } while (0)

int main()
{
    typedef ScaFES::Ntuple<int, 3> I3;
    // Ntuple: component-wise comparisons, size() = product
    CHECK(I3(1, 2, 3) < I3(2, 3, 4));
    CHECK(!(I3(1, 5, 3) < I3(2, 3, 4)));
    CHECK(!(I3(1, 2, 3) < I3(1, 3, 4)));
    CHECK(I3(1, 2, 3) <= I3(1, 3, 4));
    CHECK(I3(2, 3, 4).size() == 24);
    CHECK(I3(1, 2, 3) + I3(1) == I3(2, 3, 4));

    // Grid 3-D: Fortran numbering with x fastest, offset first index
    ScaFES::Grid<3> g(I3(-1, -1, -1), I3(4, 3, 2), ScaFES::Ntuple<double, 3>(0.0), ScaFES::Ntuple<double, 3>(1.0));
    CHECK(g.nNodes() == I3(6, 5, 4));
    CHECK(g.idxNodeTuple2ScalarStyleFortran(I3(-1, -1, -1)) == 0UL);
    CHECK(g.idxNodeTuple2ScalarStyleFortran(I3(0, -1, -1)) == 1UL);
    CHECK(g.idxNodeTuple2ScalarStyleFortran(I3(-1, 0, -1)) == 6UL);
    CHECK(g.idxNodeTuple2ScalarStyleFortran(I3(-1, -1, 0)) == 30UL);
    CHECK(g.idxNodeTuple2ScalarStyleFortran(I3(4, 3, 2)) == 119UL);
    bool roundTrip = true;
    for (unsigned long s = 0; s < 120; ++s)
        roundTrip = roundTrip && g.idxNodeTuple2ScalarStyleFortran(g.idxNodeScalar2TupleStyleFortran(s)) == s;
    CHECK(roundTrip);
    // gridsize (x_last - x_first)/(n-1) and coordinates
    CHECK(std::fabs(g.gridsize(0) - 1.0 / 5.0) < 1e-15);
    CHECK(std::fabs(g.coordinates(I3(4, 3, 2))[1] - 1.0) < 1e-15);
    CHECK(std::fabs(g.coordinates(I3(0, 0, 0))[2] - 1.0 / 3.0) < 1e-15);
    // neighbours: dir 2p = minus, 2p+1 = plus; sentinel -(dir+1) at the edge
    CHECK(g.neigh(I3(0, 0, 0), 0) == I3(-1, 0, 0));
    CHECK(g.neigh(I3(0, 0, 0), 5) == I3(0, 0, 1));
    CHECK(g.neigh(I3(-1, 0, 0), 0) == I3(-1, -1, -1));
    CHECK(g.neigh(I3(4, 0, 0), 1) == I3(-2, -2, -2));
    CHECK(g.neigh(I3(0, 0, 2), 5) == I3(-6, -6, -6));

    // GridSub iterator walks Fortran order
    ScaFES::GridSub<3> sub(I3(0, 0, 0), I3(1, 2, 1), g);
    int cnt = 0;
    I3 prev(-9);
    bool order = true;
    for (ScaFES::GridSub<3>::iterator it = sub.begin(); it != sub.end(); ++it, ++cnt) {
        if (cnt == 1) order = order && (*it == I3(1, 0, 0));
        if (cnt == 2) order = order && (*it == I3(0, 1, 0));
        prev = *it;
    }
    CHECK(cnt == 12 && order && prev == I3(1, 2, 1));

    // GridGlobal from parameters and DataField addressing with ghosts
    const char* argv[] = {"prog", "--dim=3", "--nNodes=5x4x3", "--nTimesteps=10", "--endtime=2"};
    ScaFES::Parameters params(5, const_cast<char**>(argv));
    CHECK(std::fabs(params.tau() - 0.2) < 1e-15);
    ScaFES::GridGlobal<3> gg(params);
    CHECK(gg.discreteDomain().nNodes() == I3(5, 4, 3));
    ScaFES::Grid<3> memAll(I3(-1), I3(5, 4, 3), ScaFES::Ntuple<double, 3>(0.0), ScaFES::Ntuple<double, 3>(1.0));
    std::vector<double> lump((std::size_t)memAll.nNodes().size() + 1, 0.0);
    ScaFES::DataField<double, 3> df("U", &params, gg, memAll, 1, lump.data());
    df(I3(0, 0, 0)) = 7.0;
    CHECK(lump[1 + 7 * (1 + 6 * 1)] == 7.0);
    df.time() = 3.5;
    CHECK(lump.back() == 3.5);
    df(I3(2, 1, 0)) = 1.0;
    CHECK(df(2, 1, 0) == 1.0);

    std::printf("host_api_unit: %d checks, %d failed\n", g_checks, g_fail);
    return g_fail ? 1 : 0;
}
