// ScaFES_DataField.hpp -- non-owning view of one field in a Problem lump.
//
// Same storage contract as ScaFES::DataField (src/ScaFES_DataField.hpp):
//  * memAll = partition grid extended by the stencil width (ghost layers),
//    Fortran numbering, nColumns components per node (:622-642, :1187-1199);
//  * the field's time lives in the last slot of its memory (:947-953);
//  * the view is a pointer into memory owned by the Problem, swapped in
//    O(1) between old and new iterates (:1112-1135).
// setValuesAtMemAll copies nElemsAllocated()*sizeof(CT) bytes (the
// reference's sizeof(CT*) at :1399-1400 copies twice that for CT=float).
// sync() is a no-op: one GPU holds the whole domain (:1920-1923 with one rank).
#ifndef SCAFES_HOST_DATAFIELD_HPP_
#define SCAFES_HOST_DATAFIELD_HPP_

#include <cmath>
#include <cstring>
#include <iostream>
#include <string>
#include <utility>

#include "ScaFES_DataFile.hpp"
#include "ScaFES_GridGlobal.hpp"
#include "ScaFES_Parameters.hpp"

namespace ScaFES {

template <typename CT, std::size_t DIM = 3>
class DataField {
public:
    typedef Ntuple<int, DIM> Index;
    typedef CT value_type;

    DataField() : mParams(0), mGG(0), mStencilWidth(0), mName("EMPTYNAME"), mElemData(0), mNcolumns(1),
                  mNlayers(0), mWriteToFile(WriteHowOften::NEVER) {}

    DataField(const std::string& name, Parameters* params, const GridGlobal<DIM>& gg,
              const Grid<DIM>& memAll, const int& stencilWidth, CT* elemData, const int& nColumns = 1,
              const int& nLayers = 0, const WriteHowOften& writeToFile = WriteHowOften::NEVER)
        : mParams(params), mGG(&gg), mMemAll(memAll), mStencilWidth(stencilWidth), mName(name),
          mElemData(elemData), mNcolumns(nColumns), mNlayers(nLayers), mWriteToFile(writeToFile)
    {
    }

    const std::string& name() const { return mName; }
    const int& stencilWidth() const { return mStencilWidth; }
    const Grid<DIM>& memAll() const { return mMemAll; }
    const GridGlobal<DIM>& gridGlobal() const { return *mGG; }
    Parameters* params() const { return mParams; }
    const int& nColumns() const { return mNcolumns; }
    int nRows() const { return mMemAll.nNodes().size() + 1; }
    int nElemsAllocated() const { return nRows() * mNcolumns; }
    CT* elemData() const { return mElemData; }
    const CT& elemData(const int& i) const { return mElemData[i]; }
    const int& nLayers() const { return mNlayers; }
    const WriteHowOften& writeToFile() const { return mWriteToFile; }

    CT& time() const { return mElemData[nElemsAllocated() - mNcolumns]; }

    CT& operator()(const unsigned long& memoryPos) const { return mElemData[memoryPos]; }
    CT& operator()(const Index& idx, const int& comp = 0) const
    {
        unsigned long s = 0;
        mMemAll.idxNodeTuple2Scalar(s, idx);
        return mElemData[mNcolumns * s + comp];
    }
    CT& operator()(const int& i, const int& j, const int& k, const int& comp = 0) const
    {
        static_assert(DIM == 3, "Method is valid in case DIM=3, only.");
        return (*this)(Index(i, j, k), comp);
    }
    CT& operator()(const int& i, const int& j) const
    {
        static_assert(DIM == 2, "Method is valid in case DIM=2, only.");
        return (*this)(Index(i, j), 0);
    }
    CT& operator[](const Index& idx) const { return (*this)(idx, 0); }

    void idxNode2MemoryPos(unsigned long& pos, const Index& idx, const int& comp = 0) const
    {
        unsigned long s = 0;
        mMemAll.idxNodeTuple2Scalar(s, idx);
        pos = mNcolumns * s + comp;
    }
    void memoryPos2IdxNode(Index& idx, int& comp, const unsigned long& pos) const
    {
        comp = (int)(pos % mNcolumns);
        mMemAll.idxNodeScalar2Tuple(idx, pos / mNcolumns);
    }

    void swapPointerToMemoryWith(DataField& df) { std::swap(mElemData, df.mElemData); }
    void setElemData(CT* p) { mElemData = p; }

    DataField& operator=(const CT& val)
    {
        for (int i = 0; i < nElemsAllocated(); ++i) mElemData[i] = val;
        return *this;
    }
    DataField& operator=(const DataField&) = default;
    DataField(const DataField&) = default;

    bool hasSameSize(const DataField& df) const { return nElemsAllocated() == df.nElemsAllocated(); }
    void assignValues(const DataField& from)
    {
        if (hasSameSize(from))
            for (int i = 0; i < nElemsAllocated(); ++i) mElemData[i] = from.mElemData[i];
    }
    void setValuesAtMemAll(const DataField& df)
    {
        std::memcpy(mElemData, df.elemData(), (std::size_t)nElemsAllocated() * sizeof(CT));
    }

    // Max |this - df| over the partition nodes (host copy of the data;
    // the engine computes the same quantity on the device).
    double compErrLinf(const DataField& df) const
    {
        double e = 0.0;
        const GridSub<DIM>& part = mGG->partition(0);
        for (typename GridSub<DIM>::iterator it = part.begin(); it != part.end(); ++it)
            for (int c = 0; c < mNcolumns; ++c) {
                const double d = std::fabs((double)(*this)(*it, c) - (double)df(*it, c));
                if (d > e) e = d;
            }
        return e;
    }
    bool checkConv(const DataField& dfOld) const
    {
        const GridSub<DIM>& part = mGG->partition(0);
        for (typename GridSub<DIM>::iterator it = part.begin(); it != part.end(); ++it)
            for (int c = 0; c < mNcolumns; ++c)
                if (std::fabs((double)(*this)(*it, c) - (double)dfOld(*it, c)) > mParams->threshold())
                    return true;
        return false;
    }

    void sync(const int& /*timeIter*/) {}

private:
    Parameters* mParams;
    const GridGlobal<DIM>* mGG;
    Grid<DIM> mMemAll;
    int mStencilWidth;
    std::string mName;
    CT* mElemData;
    int mNcolumns;
    int mNlayers;
    WriteHowOften mWriteToFile;
};

}  // namespace ScaFES
#endif
