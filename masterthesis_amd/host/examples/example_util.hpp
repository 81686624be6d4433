// example_util.hpp -- result dump shared by the example programs: the node
// values (Fortran order, no ghost layers) of a data field as raw float64,
// written to the file named by the environment variable SCAFES_DUMP.
#ifndef SCAFES_EXAMPLE_UTIL_HPP_
#define SCAFES_EXAMPLE_UTIL_HPP_

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "ScaFES.hpp"

template <typename CT, std::size_t DIM>
inline void dumpNodes(const ScaFES::DataField<CT, DIM>& df, const std::string& suffix = "")
{
    const char* path = std::getenv("SCAFES_DUMP");
    if (!path) return;
    const ScaFES::GridSub<DIM>& part = df.gridGlobal().partition(0);
    std::vector<double> v;
    for (typename ScaFES::GridSub<DIM>::iterator it = part.begin(); it != part.end(); ++it)
        v.push_back((double)df(*it));
    const std::string name = std::string(path) + suffix;
    FILE* f = std::fopen(name.c_str(), "wb");
    if (!f) return;
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

#endif
