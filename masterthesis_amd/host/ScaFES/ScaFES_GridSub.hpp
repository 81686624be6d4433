// ScaFES_GridSub.hpp -- sub-rectangle of a grid with a Fortran-order node
// iterator (role of ScaFES::GridSub, src/ScaFES_GridSub.hpp:68-370).
#ifndef SCAFES_HOST_GRIDSUB_HPP_
#define SCAFES_HOST_GRIDSUB_HPP_

#include "ScaFES_Grid.hpp"

namespace ScaFES {

template <std::size_t DIM = 3>
class GridSub {
public:
    typedef Ntuple<int, DIM> Index;

    SCAFES_HD GridSub() : mFirst(0), mLast(-1), mBase() {}
    SCAFES_HD GridSub(const Index& first, const Index& last, const Grid<DIM>& base)
        : mFirst(first), mLast(last), mBase(base)
    {
    }

    SCAFES_HD const Index& idxNodeFirstSub() const { return mFirst; }
    SCAFES_HD const Index& idxNodeLastSub() const { return mLast; }
    SCAFES_HD Index nNodesSub() const
    {
        Index n;
        for (std::size_t p = 0; p < DIM; ++p) n[p] = mLast.elem(p) - mFirst.elem(p) + 1;
        return n;
    }
    SCAFES_HD int nNodesTotal() const
    {
        for (std::size_t p = 0; p < DIM; ++p)
            if (mLast.elem(p) < mFirst.elem(p)) return 0;
        return nNodesSub().size();
    }
    SCAFES_HD const Grid<DIM>& grid() const { return mBase; }
    SCAFES_HD bool inside(const Index& idx) const { return mFirst <= idx && idx <= mLast; }

    // Forward iterator over the nodes in Fortran order (x fastest).
    class iterator {
    public:
        iterator(const GridSub* g, const Index& idx, bool end) : mG(g), mIdx(idx), mEnd(end) {}
        const Index& idxNode() const { return mIdx; }
        const Index& operator*() const { return mIdx; }
        iterator& operator++()
        {
            for (std::size_t p = 0; p < DIM; ++p) {
                if (mIdx[p] < mG->mLast.elem(p)) {
                    ++mIdx[p];
                    return *this;
                }
                mIdx[p] = mG->mFirst.elem(p);
            }
            mEnd = true;
            return *this;
        }
        bool operator!=(const iterator& o) const { return mEnd != o.mEnd || (!mEnd && mIdx != o.mIdx); }
        bool operator==(const iterator& o) const { return !(*this != o); }

    private:
        const GridSub* mG;
        Index mIdx;
        bool mEnd;
    };
    iterator begin() const { return iterator(this, mFirst, nNodesTotal() == 0); }
    iterator end() const { return iterator(this, mFirst, true); }

private:
    Index mFirst, mLast;
    Grid<DIM> mBase;
};

}  // namespace ScaFES
#endif
