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

#include "ScaFES_HostDevice.hpp"

namespace ScaFES {

template <typename TT, std::size_t NN>
class Ntuple {
public:
    typedef TT value_type;
    static const std::size_t DIM = NN;

    SCAFES_HD Ntuple() { for (std::size_t i = 0; i < NN; ++i) mElem[i] = TT(0); }
    // Every component set to v; implicit, as in the reference, so that a
    // scalar compares component-wise with a tuple (x < eps).
    SCAFES_HD Ntuple(const TT& v) { for (std::size_t i = 0; i < NN; ++i) mElem[i] = v; }
    SCAFES_HD Ntuple(const TT& a, const TT& b)
    {
        static_assert(NN == 2, "two-component constructor needs NN == 2");
        mElem[0] = a;
        mElem[1] = b;
    }
    SCAFES_HD Ntuple(const TT& a, const TT& b, const TT& c)
    {
        static_assert(NN == 3, "three-component constructor needs NN == 3");
        mElem[0] = a;
        mElem[1] = b;
        mElem[2] = c;
    }
    SCAFES_HD Ntuple(const TT& a, const TT& b, const TT& c, const TT& d)
    {
        static_assert(NN == 4, "four-component constructor needs NN == 4");
        mElem[0] = a;
        mElem[1] = b;
        mElem[2] = c;
        mElem[3] = d;
    }

    SCAFES_HD const TT& elem(const std::size_t& i) const { return mElem[i]; }
    SCAFES_HD TT& operator[](const std::size_t& i) { return mElem[i]; }
    SCAFES_HD const TT& operator[](const std::size_t& i) const { return mElem[i]; }
    SCAFES_HD const TT* data() const { return mElem; }

    // Product of all components (e.g. number of nodes of an extent).
    SCAFES_HD TT size() const
    {
        TT s = TT(1);
        for (std::size_t i = 0; i < NN; ++i) s *= mElem[i];
        return s;
    }

    SCAFES_HD Ntuple& operator+=(const Ntuple& o)
    {
        for (std::size_t i = 0; i < NN; ++i) mElem[i] += o.mElem[i];
        return *this;
    }
    SCAFES_HD Ntuple& operator-=(const Ntuple& o)
    {
        for (std::size_t i = 0; i < NN; ++i) mElem[i] -= o.mElem[i];
        return *this;
    }
    SCAFES_HD Ntuple operator+(const Ntuple& o) const { return Ntuple(*this) += o; }
    SCAFES_HD Ntuple operator-(const Ntuple& o) const { return Ntuple(*this) -= o; }

    SCAFES_HD bool operator==(const Ntuple& o) const
    {
        for (std::size_t i = 0; i < NN; ++i)
            if (mElem[i] != o.mElem[i]) return false;
        return true;
    }
    SCAFES_HD bool operator!=(const Ntuple& o) const { return !(*this == o); }
    // Component-wise: a < b iff a_i < b_i for every i (src/ScaFES_Ntuple.hpp:383-397).
    SCAFES_HD bool operator<(const Ntuple& o) const
    {
        for (std::size_t i = 0; i < NN; ++i)
            if (!(mElem[i] < o.mElem[i])) return false;
        return true;
    }
    SCAFES_HD bool operator<=(const Ntuple& o) const
    {
        for (std::size_t i = 0; i < NN; ++i)
            if (!(mElem[i] <= o.mElem[i])) return false;
        return true;
    }
    SCAFES_HD bool operator>(const Ntuple& o) const { return o < *this; }
    SCAFES_HD bool operator>=(const Ntuple& o) const { return o <= *this; }

private:
    TT mElem[NN];
};

template <typename TT, std::size_t NN>
inline std::ostream& operator<<(std::ostream& os, const Ntuple<TT, NN>& a)
{
    os << "(";
    for (std::size_t i = 0; i < NN; ++i) os << (i ? ", " : "") << a.elem(i);
    return os << ")";
}

}  // namespace ScaFES
#endif
