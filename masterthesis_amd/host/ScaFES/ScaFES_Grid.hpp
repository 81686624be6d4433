// ScaFES_Grid.hpp -- hyper-rectangular grid of the host API.
//
// Numbering, coordinates and neighbours as ScaFES::Grid (src/ScaFES_Grid.hpp):
//  * gridsize h_p = (x_last - x_first)/(n_p - 1)            (Grid.hpp:547-555)
//  * coordinates x_p = x_first + h_p*(i_p - i_first)         (Grid.hpp:719-733)
//  * Fortran numbering, x fastest (default style)           (Grid.hpp:955-968,1002-1013)
//  * neigh(idx, dir): dir 2p = minus, 2p+1 = plus along p; at the grid edge
//    every component becomes the sentinel -(dir+1)           (Grid.hpp:885-921)
#ifndef SCAFES_HOST_GRID_HPP_
#define SCAFES_HOST_GRID_HPP_

#include "ScaFES_HostDevice.hpp"
#include "ScaFES_Ntuple.hpp"

namespace ScaFES {

template <std::size_t DIM = 3>
class Grid {
public:
    typedef Ntuple<int, DIM> Index;

    SCAFES_HD Grid() : mFirst(0), mLast(1), mN(2), mX0(0.0), mX1(1.0), mH(1.0), mStyleC(false) {}

    SCAFES_HD Grid(const Index& first, const Index& last, Ntuple<double, DIM> x0, Ntuple<double, DIM> x1)
        : mFirst(first), mLast(last), mN(last - first + Index(1)), mX0(x0), mX1(x1), mH(0.0),
          mStyleC(false)
    {
        if (0 < mN.size())
            for (std::size_t p = 0; p < DIM; ++p) mH[p] = (x1.elem(p) - x0.elem(p)) / (mN.elem(p) - 1);
    }

    SCAFES_HD const Index& idxNodeFirst() const { return mFirst; }
    SCAFES_HD const Index& idxNodeLast() const { return mLast; }
    SCAFES_HD const int& idxNodeFirst(const int& p) const { return mFirst.elem(p); }
    SCAFES_HD const int& idxNodeLast(const int& p) const { return mLast.elem(p); }
    SCAFES_HD const Index& nNodes() const { return mN; }
    SCAFES_HD const int& nNodes(const int& p) const { return mN.elem(p); }
    SCAFES_HD const Ntuple<double, DIM>& coordNodeFirst() const { return mX0; }
    SCAFES_HD const Ntuple<double, DIM>& coordNodeLast() const { return mX1; }
    SCAFES_HD const double& gridsize(const int& p) const { return mH.elem(p); }
    SCAFES_HD const Ntuple<double, DIM>& gridsize() const { return mH; }
    SCAFES_HD bool isNumberedStyleC() const { return mStyleC; }
    SCAFES_HD bool valid() const { return mFirst <= mLast; }

    SCAFES_HD Ntuple<double, DIM> coordinates(const Index& idx) const
    {
        Ntuple<double, DIM> c(0.0);
        for (std::size_t p = 0; p < DIM; ++p)
            c[p] = mX0.elem(p) + mH.elem(p) * (idx.elem(p) - mFirst.elem(p));
        return c;
    }

    SCAFES_HD Index neigh(const Index& idx, const int& dir) const
    {
        Index nb(idx);
        const int p = dir / 2;
        const bool minus = (dir % 2) == 0;
        if ((minus && idx.elem(p) == mFirst.elem(p)) || (!minus && idx.elem(p) == mLast.elem(p))) {
            for (std::size_t q = 0; q < DIM; ++q) nb[q] = -(dir + 1);
        } else {
            nb[p] += minus ? -1 : 1;
        }
        return nb;
    }

    SCAFES_HD bool inside(const Index& idx) const { return mFirst <= idx && idx <= mLast; }

    SCAFES_HD void idxNodeTuple2ScalarStyleFortran(unsigned long& s, const Index& idx) const
    {
        s = idx.elem(DIM - 1) - mFirst.elem(DIM - 1);
        for (int p = (int)DIM - 2; p >= 0; --p) s = s * mN.elem(p) + idx.elem(p) - mFirst.elem(p);
    }
    SCAFES_HD unsigned long idxNodeTuple2ScalarStyleFortran(const Index& idx) const
    {
        unsigned long s = 0;
        idxNodeTuple2ScalarStyleFortran(s, idx);
        return s;
    }
    SCAFES_HD void idxNodeScalar2TupleStyleFortran(Index& idx, const unsigned long& s) const
    {
        unsigned long t = s;
        for (std::size_t p = 0; p < DIM; ++p) {
            idx[p] = (int)(t % mN.elem(p)) + mFirst.elem(p);
            t /= mN.elem(p);
        }
    }
    SCAFES_HD Index idxNodeScalar2TupleStyleFortran(const unsigned long& s) const
    {
        Index idx;
        idxNodeScalar2TupleStyleFortran(idx, s);
        return idx;
    }
    SCAFES_HD void idxNodeTuple2ScalarStyleC(unsigned long& s, const Index& idx) const
    {
        s = idx.elem(0) - mFirst.elem(0);
        for (std::size_t p = 1; p < DIM; ++p) s = s * mN.elem(p) + idx.elem(p) - mFirst.elem(p);
    }
    SCAFES_HD void idxNodeScalar2TupleStyleC(Index& idx, const unsigned long& s) const
    {
        unsigned long t = s;
        for (std::size_t p = DIM - 1; p > 0; --p) {
            idx[p] = (int)(t % mN.elem(p)) + mFirst.elem(p);
            t /= mN.elem(p);
        }
        idx[0] = (int)t + mFirst.elem(0);
    }
    // Default numbering (Fortran style unless switched).
    SCAFES_HD void idxNodeTuple2Scalar(unsigned long& s, const Index& idx) const
    {
        if (mStyleC) idxNodeTuple2ScalarStyleC(s, idx);
        else idxNodeTuple2ScalarStyleFortran(s, idx);
    }
    SCAFES_HD void idxNodeScalar2Tuple(Index& idx, const unsigned long& s) const
    {
        if (mStyleC) idxNodeScalar2TupleStyleC(idx, s);
        else idxNodeScalar2TupleStyleFortran(idx, s);
    }

private:
    Index mFirst, mLast, mN;
    Ntuple<double, DIM> mX0, mX1, mH;
    bool mStyleC;
};

}  // namespace ScaFES
#endif
