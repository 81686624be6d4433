// ScaFES_Ntuple.hpp -- index / coordinate tuple of the host API.
//
// Same role and operations as ScaFES::Ntuple (src/ScaFES_Ntuple.hpp): a
// fixed-size tuple with component access, component-wise arithmetic and the
// component-wise (not lexicographic) comparisons of src/ScaFES_Ntuple.hpp:383-397.
// size() is the product of the components (number of nodes of an extent).
#ifndef SCAFES_HOST_NTUPLE_HPP_
#define SCAFES_HOST_NTUPLE_HPP_

#include <cstddef>
#include <ostream>

namespace ScaFES {

template <typename TT, std::size_t NN>
class Ntuple {
public:
    typedef TT value_type;
    static const std::size_t DIM = NN;

    Ntuple() { for (std::size_t i = 0; i < NN; ++i) mElem[i] = TT(0); }
    explicit Ntuple(const TT& v) { for (std::size_t i = 0; i < NN; ++i) mElem[i] = v; }
    Ntuple(const TT& a, const TT& b)
    {
        static_assert(NN == 2, "two-component constructor needs NN == 2");
        mElem[0] = a;
        mElem[1] = b;
    }
    Ntuple(const TT& a, const TT& b, const TT& c)
    {
        static_assert(NN == 3, "three-component constructor needs NN == 3");
        mElem[0] = a;
        mElem[1] = b;
        mElem[2] = c;
    }
    Ntuple(const TT& a, const TT& b, const TT& c, const TT& d)
    {
        static_assert(NN == 4, "four-component constructor needs NN == 4");
        mElem[0] = a;
        mElem[1] = b;
        mElem[2] = c;
        mElem[3] = d;
    }

    const TT& elem(const std::size_t& i) const { return mElem[i]; }
    TT& operator[](const std::size_t& i) { return mElem[i]; }
    const TT& operator[](const std::size_t& i) const { return mElem[i]; }
    const TT* data() const { return mElem; }

    // Product of all components (e.g. number of nodes of an extent).
    TT size() const
    {
        TT s = TT(1);
        for (std::size_t i = 0; i < NN; ++i) s *= mElem[i];
        return s;
    }

    Ntuple& operator+=(const Ntuple& o)
    {
        for (std::size_t i = 0; i < NN; ++i) mElem[i] += o.mElem[i];
        return *this;
    }
    Ntuple& operator-=(const Ntuple& o)
    {
        for (std::size_t i = 0; i < NN; ++i) mElem[i] -= o.mElem[i];
        return *this;
    }

private:
    TT mElem[NN];
};

template <typename TT, std::size_t NN>
inline Ntuple<TT, NN> operator+(Ntuple<TT, NN> a, const Ntuple<TT, NN>& b) { return a += b; }
template <typename TT, std::size_t NN>
inline Ntuple<TT, NN> operator-(Ntuple<TT, NN> a, const Ntuple<TT, NN>& b) { return a -= b; }

template <typename TT, std::size_t NN>
inline bool operator==(const Ntuple<TT, NN>& a, const Ntuple<TT, NN>& b)
{
    for (std::size_t i = 0; i < NN; ++i)
        if (a.elem(i) != b.elem(i)) return false;
    return true;
}
template <typename TT, std::size_t NN>
inline bool operator!=(const Ntuple<TT, NN>& a, const Ntuple<TT, NN>& b) { return !(a == b); }

// Component-wise: a < b iff a_i < b_i for every i (src/ScaFES_Ntuple.hpp:383-397).
template <typename TT, std::size_t NN>
inline bool operator<(const Ntuple<TT, NN>& a, const Ntuple<TT, NN>& b)
{
    for (std::size_t i = 0; i < NN; ++i)
        if (!(a.elem(i) < b.elem(i))) return false;
    return true;
}
template <typename TT, std::size_t NN>
inline bool operator<=(const Ntuple<TT, NN>& a, const Ntuple<TT, NN>& b)
{
    for (std::size_t i = 0; i < NN; ++i)
        if (!(a.elem(i) <= b.elem(i))) return false;
    return true;
}
template <typename TT, std::size_t NN>
inline bool operator>(const Ntuple<TT, NN>& a, const Ntuple<TT, NN>& b) { return b < a; }
template <typename TT, std::size_t NN>
inline bool operator>=(const Ntuple<TT, NN>& a, const Ntuple<TT, NN>& b) { return b <= a; }

template <typename TT, std::size_t NN>
inline std::ostream& operator<<(std::ostream& os, const Ntuple<TT, NN>& a)
{
    os << "(";
    for (std::size_t i = 0; i < NN; ++i) os << (i ? ", " : "") << a.elem(i);
    return os << ")";
}

}  // namespace ScaFES
#endif
