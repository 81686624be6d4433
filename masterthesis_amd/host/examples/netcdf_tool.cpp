// netcdf_tool.cpp -- exercises ScaFES::DataFile (classic NetCDF snapshot
// files) on the CPU, for the tests:
//   netcdf_tool write <base> <nrec> n0 [n1 [n2]]
//       variables A (written every record: A = i + 1000 r over the x-fastest
//       node index i) and B (written on even records: B = -(i + r)); odd
//       records of B hold the fill value
//   netcdf_tool read <base> n0 [n1 [n2]] -- <var> [<var> ...]
//       prints record 0 of each variable, one value per line (%.17g)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ScaFES_DataFile.hpp"

template <std::size_t DIM>
int run(const std::string& mode, const std::string& base, int nrec, const std::vector<int>& n,
        const std::vector<std::string>& vars)
{
    ScaFES::Ntuple<int, DIM> nn;
    for (std::size_t p = 0; p < DIM; ++p) nn[p] = n[p];
    if (mode == "write") {
        ScaFES::DataFile<double, DIM> f(base, {"A", "B"}, nn);
        const std::size_t cnt = f.nNodesTotal();
        std::vector<double> a(cnt), b(cnt);
        for (int r = 0; r < nrec; ++r) {
            for (std::size_t i = 0; i < cnt; ++i) {
                a[i] = (double)i + 1000.0 * r;
                b[i] = -((double)i + r);
            }
            f.write({a.data(), b.data()}, {true, r % 2 == 0});
        }
        return 0;
    }
    ScaFES::DataFile<double, DIM> f(base, vars, nn);
    const std::size_t cnt = f.nNodesTotal();
    std::vector<std::vector<double>> data(vars.size(), std::vector<double>(cnt));
    std::vector<double*> ptrs;
    for (auto& d : data) ptrs.push_back(d.data());
    f.read(ptrs);
    for (const auto& d : data)
        for (double v : d) std::printf("%.17g\n", v);
    return 0;
}

int main(int argc, char* argv[])
{
    try {
        if (argc < 4) throw std::runtime_error("usage: netcdf_tool write|read <base> ...");
        const std::string mode = argv[1], base = argv[2];
        int a = 3, nrec = 0;
        if (mode == "write") nrec = std::atoi(argv[a++]);
        std::vector<int> n;
        std::vector<std::string> vars;
        for (; a < argc && std::strcmp(argv[a], "--") != 0; ++a) n.push_back(std::atoi(argv[a]));
        for (++a; a < argc; ++a) vars.push_back(argv[a]);
        switch (n.size()) {
        case 1: return run<1>(mode, base, nrec, n, vars);
        case 2: return run<2>(mode, base, nrec, n, vars);
        case 3: return run<3>(mode, base, nrec, n, vars);
        default: throw std::runtime_error("1 to 3 grid sizes expected");
        }
    } catch (const std::exception& e) {
        std::fprintf(stderr, "netcdf_tool: %s\n", e.what());
        return 1;
    }
}
