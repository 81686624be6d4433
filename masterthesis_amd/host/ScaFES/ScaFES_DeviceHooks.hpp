// ScaFES_DeviceHooks.hpp -- the generic device update phase: a problem's
// own updateInner/updateBorder(/2) hooks run as GPU code (hipcc only).
//
// The reference calls the hooks per node in an OpenMP loop
// (src/ScaFES_Problem.hpp:5397-5476): with TT = CT for a plain run and
// TT = adouble when ADOL-C tapes them, so a ScaFES problem writes its hooks
// for a generic active type TT.  Here they are instantiated once more, with
// TT = DevReal<CT>, a value type that is a CT with host+device arithmetic,
// and called from one GPU thread per node:
//   * std::vector<DataField<DevReal<CT>, DIM>> is a fixed-capacity array of
//     device field views (a specialisation for this program-defined element
//     type), so vNew[k](idx) / vOld[k](connect(idx, dir)) / vNew.size() work
//     on the device unchanged;
//   * DataField<DevReal<CT>, DIM> is a device view of one field: its memAll
//     grid, ghost width, time, and its element pointer in HBM;
//   * DevReal<CT> arithmetic is CT arithmetic, operation for operation, so
//     with -ffp-contract=off the device result equals the host hooks' bit
//     for bit.
// The problem object itself is copied byte for byte to device memory; its
// hook-facing accessors (connect, tau, gridsize, nNodes, coordinates, time,
// knownDf, knownDfOld, kind) read only plain members and device views.
// Hooks must write only vNew[k] at their own node (the reference's contract,
// which its OpenMP loop relies on as well).
#ifndef SCAFES_HOST_DEVICEHOOKS_HPP_
#define SCAFES_HOST_DEVICEHOOKS_HPP_

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <type_traits>
#include <vector>

#include "ScaFES_DataField.hpp"
#include "ScaFES_GridGlobal.hpp"
#include "ScaFES_HostDevice.hpp"

namespace ScaFES {

// ------------------------------------------------------------------ DevReal
// The active type of the device update phase: a CT with the arithmetic,
// comparisons and math functions an ADOL-C-ready hook uses.  Mixed
// operations with plain numbers promote as the built-in types do.
template <typename T>
struct DevReal {
    T v;
    DevReal() = default;
    SCAFES_HD constexpr DevReal(T x) : v(x) {}
    template <typename U, typename = typename std::enable_if<std::is_arithmetic<U>::value>::type>
    SCAFES_HD constexpr DevReal(U x) : v(static_cast<T>(x))
    {
    }
    template <typename U>
    SCAFES_HD constexpr DevReal(const DevReal<U>& o) : v(static_cast<T>(o.v))
    {
    }
    SCAFES_HD explicit operator T() const { return v; }
    SCAFES_HD T getValue() const { return v; }  // adouble's accessor
    SCAFES_HD T value() const { return v; }
    SCAFES_HD DevReal& operator+=(const DevReal& o)
    {
        v = v + o.v;
        return *this;
    }
    SCAFES_HD DevReal& operator-=(const DevReal& o)
    {
        v = v - o.v;
        return *this;
    }
    SCAFES_HD DevReal& operator*=(const DevReal& o)
    {
        v = v * o.v;
        return *this;
    }
    SCAFES_HD DevReal& operator/=(const DevReal& o)
    {
        v = v / o.v;
        return *this;
    }
    SCAFES_HD DevReal operator-() const { return DevReal(-v); }
    SCAFES_HD DevReal operator+() const { return *this; }
};

template <typename U>
using EnableArith = typename std::enable_if<std::is_arithmetic<U>::value, int>::type;

#define SCAFES_DEVREAL_BINOP(OP)                                                                     \
    template <typename A, typename B>                                                                \
    SCAFES_HD DevReal<decltype(A() OP B())> operator OP(const DevReal<A>& a, const DevReal<B>& b)    \
    {                                                                                                \
        return DevReal<decltype(A() OP B())>(a.v OP b.v);                                            \
    }                                                                                                \
    template <typename A, typename U, EnableArith<U> = 0>                                            \
    SCAFES_HD DevReal<decltype(A() OP U())> operator OP(const DevReal<A>& a, U b)                    \
    {                                                                                                \
        return DevReal<decltype(A() OP U())>(a.v OP b);                                              \
    }                                                                                                \
    template <typename U, typename B, EnableArith<U> = 0>                                            \
    SCAFES_HD DevReal<decltype(U() OP B())> operator OP(U a, const DevReal<B>& b)                    \
    {                                                                                                \
        return DevReal<decltype(U() OP B())>(a OP b.v);                                              \
    }
SCAFES_DEVREAL_BINOP(+)
SCAFES_DEVREAL_BINOP(-)
SCAFES_DEVREAL_BINOP(*)
SCAFES_DEVREAL_BINOP(/)
#undef SCAFES_DEVREAL_BINOP

#define SCAFES_DEVREAL_CMP(OP)                                                                       \
    template <typename A, typename B>                                                                \
    SCAFES_HD bool operator OP(const DevReal<A>& a, const DevReal<B>& b) { return a.v OP b.v; }      \
    template <typename A, typename U, EnableArith<U> = 0>                                            \
    SCAFES_HD bool operator OP(const DevReal<A>& a, U b) { return a.v OP b; }                        \
    template <typename U, typename B, EnableArith<U> = 0>                                            \
    SCAFES_HD bool operator OP(U a, const DevReal<B>& b) { return a OP b.v; }
SCAFES_DEVREAL_CMP(==)
SCAFES_DEVREAL_CMP(!=)
SCAFES_DEVREAL_CMP(<)
SCAFES_DEVREAL_CMP(<=)
SCAFES_DEVREAL_CMP(>)
SCAFES_DEVREAL_CMP(>=)
#undef SCAFES_DEVREAL_CMP

// Math functions found by argument-dependent lookup, as ADOL-C's are for adouble.
#define SCAFES_DEVREAL_FN1(FN)                                                                       \
    template <typename T>                                                                            \
    SCAFES_HD DevReal<T> FN(const DevReal<T>& a)                                                     \
    {                                                                                                \
        using std::FN;                                                                               \
        return DevReal<T>(FN(a.v));                                                                  \
    }
SCAFES_DEVREAL_FN1(sqrt)
SCAFES_DEVREAL_FN1(exp)
SCAFES_DEVREAL_FN1(log)
SCAFES_DEVREAL_FN1(log10)
SCAFES_DEVREAL_FN1(sin)
SCAFES_DEVREAL_FN1(cos)
SCAFES_DEVREAL_FN1(tan)
SCAFES_DEVREAL_FN1(asin)
SCAFES_DEVREAL_FN1(acos)
SCAFES_DEVREAL_FN1(atan)
SCAFES_DEVREAL_FN1(sinh)
SCAFES_DEVREAL_FN1(cosh)
SCAFES_DEVREAL_FN1(tanh)
SCAFES_DEVREAL_FN1(fabs)
SCAFES_DEVREAL_FN1(floor)
SCAFES_DEVREAL_FN1(ceil)
#undef SCAFES_DEVREAL_FN1
template <typename T>
SCAFES_HD DevReal<T> abs(const DevReal<T>& a)
{
    using std::fabs;
    return DevReal<T>(fabs(a.v));
}

#define SCAFES_DEVREAL_FN2(FN)                                                                       \
    template <typename A, typename B>                                                                \
    SCAFES_HD DevReal<decltype(A() + B())> FN(const DevReal<A>& a, const DevReal<B>& b)              \
    {                                                                                                \
        using std::FN;                                                                               \
        return DevReal<decltype(A() + B())>(FN(a.v, b.v));                                           \
    }                                                                                                \
    template <typename A, typename U, EnableArith<U> = 0>                                            \
    SCAFES_HD DevReal<decltype(A() + U())> FN(const DevReal<A>& a, U b)                              \
    {                                                                                                \
        using std::FN;                                                                               \
        return DevReal<decltype(A() + U())>(FN(a.v, b));                                             \
    }                                                                                                \
    template <typename U, typename B, EnableArith<U> = 0>                                            \
    SCAFES_HD DevReal<decltype(U() + B())> FN(U a, const DevReal<B>& b)                              \
    {                                                                                                \
        using std::FN;                                                                               \
        return DevReal<decltype(U() + B())>(FN(a, b.v));                                             \
    }
SCAFES_DEVREAL_FN2(pow)
SCAFES_DEVREAL_FN2(atan2)
SCAFES_DEVREAL_FN2(fmax)
SCAFES_DEVREAL_FN2(fmin)
#undef SCAFES_DEVREAL_FN2

// -------------------------------------------------------------- field view
// Device view of one data field: the element array in HBM with the field's
// memAll grid (Fortran numbering, ghost layers included, as the host
// DataField), its stencil width, components per node and time.
template <typename T, std::size_t DIM>
class DataField<DevReal<T>, DIM> {
public:
    typedef Ntuple<int, DIM> Index;
    typedef DevReal<T> value_type;

    SCAFES_HD DataField() : mData(nullptr), mGG(nullptr), mNelem(0), mStencilWidth(0), mNcolumns(1), mTime(T(0)) {}
    SCAFES_HD DataField(T* data, const Grid<DIM>& memAll, int stencilWidth, int nColumns, long nElems, T time,
                        const GridGlobal<DIM>* gg)
        : mData(reinterpret_cast<DevReal<T>*>(data)), mMemAll(memAll), mGG(gg), mNelem(nElems),
          mStencilWidth(stencilWidth), mNcolumns(nColumns), mTime(time)
    {
    }

    SCAFES_HD DevReal<T>& operator()(const Index& idx, const int& comp = 0) const
    {
        unsigned long s = 0;
        mMemAll.idxNodeTuple2ScalarStyleFortran(s, idx);
        return mData[mNcolumns * s + comp];
    }
    SCAFES_HD DevReal<T>& operator()(const unsigned long& memoryPos) const { return mData[memoryPos]; }
    SCAFES_HD DevReal<T>& operator()(const int& i, const int& j, const int& k, const int& comp = 0) const
    {
        static_assert(DIM == 3, "Method is valid in case DIM=3, only.");
        return (*this)(Index(i, j, k), comp);
    }
    SCAFES_HD DevReal<T>& operator()(const int& i, const int& j) const
    {
        static_assert(DIM == 2, "Method is valid in case DIM=2, only.");
        return (*this)(Index(i, j), 0);
    }
    SCAFES_HD DevReal<T>& operator[](const Index& idx) const { return (*this)(idx, 0); }
    SCAFES_HD DevReal<T>& time() const { return mTime; }
    SCAFES_HD const int& stencilWidth() const { return mStencilWidth; }
    SCAFES_HD const int& nColumns() const { return mNcolumns; }
    SCAFES_HD const Grid<DIM>& memAll() const { return mMemAll; }
    SCAFES_HD const GridGlobal<DIM>& gridGlobal() const { return *mGG; }
    SCAFES_HD int nElemsAllocated() const { return (int)mNelem; }
    SCAFES_HD DevReal<T>* elemData() const { return mData; }
    SCAFES_HD void idxNode2MemoryPos(unsigned long& pos, const Index& idx, const int& comp = 0) const
    {
        unsigned long s = 0;
        mMemAll.idxNodeTuple2ScalarStyleFortran(s, idx);
        pos = mNcolumns * s + comp;
    }

private:
    DevReal<T>* mData;
    Grid<DIM> mMemAll;
    const GridGlobal<DIM>* mGG;
    long mNelem;
    int mStencilWidth, mNcolumns;
    mutable DevReal<T> mTime;
};

}  // namespace ScaFES

namespace std {
// The vectors of device field views the hooks receive: a fixed-capacity
// array, trivially copyable into device memory.  It provides the part of
// std::vector's interface a hook reads (size, operator[], at, iteration).
template <typename T, std::size_t DIM, typename A>
class vector<ScaFES::DataField<ScaFES::DevReal<T>, DIM>, A> {
public:
    typedef ScaFES::DataField<ScaFES::DevReal<T>, DIM> value_type;
    typedef std::size_t size_type;
    typedef value_type& reference;
    typedef const value_type& const_reference;
    typedef value_type* iterator;
    typedef const value_type* const_iterator;

    SCAFES_HD vector() : mSize(0) {}
    SCAFES_HD size_type size() const { return mSize; }
    SCAFES_HD bool empty() const { return mSize == 0; }
    SCAFES_HD size_type capacity() const { return ScaFES::kMaxDeviceFields; }
    SCAFES_HD reference operator[](size_type i) { return mData[i]; }
    SCAFES_HD const_reference operator[](size_type i) const { return mData[i]; }
    SCAFES_HD reference at(size_type i) { return mData[i]; }
    SCAFES_HD const_reference at(size_type i) const { return mData[i]; }
    SCAFES_HD reference front() { return mData[0]; }
    SCAFES_HD const_reference front() const { return mData[0]; }
    SCAFES_HD reference back() { return mData[mSize - 1]; }
    SCAFES_HD const_reference back() const { return mData[mSize - 1]; }
    SCAFES_HD iterator begin() { return mData; }
    SCAFES_HD iterator end() { return mData + mSize; }
    SCAFES_HD const_iterator begin() const { return mData; }
    SCAFES_HD const_iterator end() const { return mData + mSize; }
    // host side, while the views are assembled
    void clear() { mSize = 0; }
    void push_back(const value_type& v) { mData[mSize++] = v; }

private:
    value_type mData[ScaFES::kMaxDeviceFields];
    size_type mSize;
};
}  // namespace std

namespace ScaFES {

// Everything one launch of the generic update reads besides the problem
// object: the unknown fields' new/old views (the hooks' vNew, vOld), the
// views knownDf/knownDfOld read (one per data field, empty where a field has
// no known memory), the global grid the views point to, and the node kinds.
template <typename T, std::size_t DIM>
struct HookState {
    std::vector<DataField<DevReal<T>, DIM>> vNew, vOld;
    DataField<DevReal<T>, DIM> knownNew[kMaxDeviceFields], knownOld[kMaxDeviceFields];
    GridGlobal<DIM> gg;
    Grid<DIM> domain;
    long nNodes;
    const short* kind;
};

// One GPU thread per node of the domain (grid-stride): updateInner at
// GLOBAL_INNER nodes, updateBorder at GLOBAL_BORDER nodes (phase 1), or the
// leapfrog's updateInner2/updateBorder2 (phase 2), as evalTypeTwoAt
// (src/ScaFES_Problem.hpp:5443-5466).  vNew already holds vOld's values.
template <class P, typename T, std::size_t DIM, int PHASE>
__global__ __launch_bounds__(256) void hooks_kernel(P* prob, HookState<T, DIM>* st, int timestep)
{
    typedef GridGlobal<DIM> GG;
    const long n = st->nNodes;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        Ntuple<int, DIM> idx;
        st->domain.idxNodeScalar2TupleStyleFortran(idx, (unsigned long)t);
        const short kind = st->kind[t];
        if (kind & GG::SCAFES_GLOBAL_INNER_NODE) {
            if (PHASE == 1) prob->template updateInner<DevReal<T>>(st->vNew, st->vOld, idx, timestep);
            else prob->template updateInner2<DevReal<T>>(st->vNew, st->vOld, idx, timestep);
        } else if (kind & GG::SCAFES_GLOBAL_BORDER_NODE) {
            if (PHASE == 1) prob->template updateBorder<DevReal<T>>(st->vNew, st->vOld, idx, timestep);
            else prob->template updateBorder2<DevReal<T>>(st->vNew, st->vOld, idx, timestep);
        }
    }
}

// Launches one phase on the library context's stream.
template <class P, typename T, std::size_t DIM>
inline int launch_hooks(void* stream, P* prob, HookState<T, DIM>* st, long nNodes, int timestep, int phase)
{
    const long blocks = (nNodes + 255) / 256;
    const dim3 grid((unsigned)(blocks < 65536 ? (blocks > 0 ? blocks : 1) : 65536)), block(256);
    if (phase == 1)
        hooks_kernel<P, T, DIM, 1><<<grid, block, 0, (hipStream_t)stream>>>(prob, st, timestep);
    else
        hooks_kernel<P, T, DIM, 2><<<grid, block, 0, (hipStream_t)stream>>>(prob, st, timestep);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace ScaFES
#endif
