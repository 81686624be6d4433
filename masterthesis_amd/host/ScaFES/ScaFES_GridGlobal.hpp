// ScaFES_GridGlobal.hpp -- global discrete domain and node-kind flags.
//
// Role of ScaFES::GridGlobal (src/ScaFES_GridGlobal.hpp): the discrete
// domain built from Parameters (:384-422) and its partitions.  One GPU holds
// the whole domain, so there is exactly one partition covering it; the RCB
// and uniform decompositions (:584-715) belong to the multi-GPU design and
// are not built.  The node-kind flag values are the reference's (:138-151).
#ifndef SCAFES_HOST_GRIDGLOBAL_HPP_
#define SCAFES_HOST_GRIDGLOBAL_HPP_

#include <vector>

#include "ScaFES_GridSub.hpp"
#include "ScaFES_Parameters.hpp"

namespace ScaFES {

template <std::size_t DIM = 3>
class GridGlobal {
public:
    enum scafes_gridnode_type {
        SCAFES_GLOBAL_ALL_NODE = 1,
        SCAFES_GLOBAL_NORMAL_NODE = 2,
        SCAFES_GLOBAL_INNER_NODE = 4,
        SCAFES_GLOBAL_BORDER_NODE = 8,
        SCAFES_GLOBAL_GHOST_NODE = 16,
        SCAFES_GLOBAL_COMM_NODE = 32,
        SCAFES_PARTITITON_ALL_NODE = 64,
        SCAFES_PARTITION_INNER_NODE = 128,
        SCAFES_PARTITION_BORDER_NODE = 256,
        SCAFES_PARTITION_GHOST_NODE = 512,
        SCAFES_PARTITION_COMM_NODE = 1024
    };
    typedef Ntuple<int, DIM> Index;

    SCAFES_HD GridGlobal() {}
    explicit GridGlobal(const Parameters& params)
    {
        if ((int)DIM != params.dim()) {
            std::cerr << "\nERROR: GridGlobal<" << DIM << "> built from parameters with --dim="
                      << params.dim() << ". Exit...\n" << std::endl;
            std::exit(-1);
        }
        Index first(0), last(0);
        Ntuple<double, DIM> x0(0.0), x1(1.0);
        for (std::size_t p = 0; p < DIM; ++p) {
            last[p] = params.nNodes()[p] - 1;
            x0[p] = params.coordNodeFirst()[p];
            x1[p] = params.coordNodeLast()[p];
        }
        mDomain = Grid<DIM>(first, last, x0, x1);
        mPartition = GridSub<DIM>(first, last, mDomain);
    }

    SCAFES_HD const Grid<DIM>& discreteDomain() const { return mDomain; }
    SCAFES_HD const GridSub<DIM>& partition(const int& /*rank*/) const { return mPartition; }
    SCAFES_HD int nPartitions() const { return 1; }
    std::vector<int> neighbourPartitions(const int& /*rank*/) const { return std::vector<int>(); }

private:
    Grid<DIM> mDomain;
    GridSub<DIM> mPartition;
};

}  // namespace ScaFES
#endif
