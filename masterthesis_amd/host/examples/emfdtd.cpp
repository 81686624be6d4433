// emfdtd.cpp -- the reference's EMFDTD example (examples/EMFDTD: EMFDTD.hpp,
// simulate.cpp) as a problem class against the host API: 25 data fields,
// leapfrog (H in updateInner, D/E/I in updateInner2), complete CPML curl
// stencils, a conductive sphere from the initialization parameters, a hard
// Ez source and an Ez sink written to the [sink] data_file.
//
// The per-node hooks below are the CPU form of what the EMFDTD device
// kernel computes; DeviceTraits hands the update phases to the GPU.  The
// configuration comes from --configfile (the reference's ini sections [PML],
// [source], [sink], [initialization]), read by a small ini reader instead of
// boost::property_tree.  The sink file holds one value per step, printed as
// the reference does (scientific, 20 digits).  SCAFES_DUMP=<path> also dumps
// the 24 updated fields (files <path>.<name>).
#include <cmath>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>

#include "ScaFES.hpp"
#include "example_util.hpp"
#include "ini_reader.hpp"

namespace {
const double C_0 = 299792458.0;               // physical_constants.hpp:12
const long double EPSILON_0 = 8.85418781762e-12L;  // physical_constants.hpp:14 (long double)
enum DF {
    SIGMA = 0, Ex, Ey, Ez, Hx, Hy, Hz, Dx, Dy, Dz, Ix, Iy, Iz,
    Psi_ExDy, Psi_ExDz, Psi_EyDx, Psi_EyDz, Psi_EzDx, Psi_EzDy,
    Psi_HxDy, Psi_HxDz, Psi_HyDx, Psi_HyDz, Psi_HzDx, Psi_HzDy, NDF
};
const char* kNames[NDF] = {"SIGMA", "Ex", "Ey", "Ez", "Hx", "Hy", "Hz", "Dx", "Dy", "Dz", "Ix", "Iy", "Iz",
                           "Psi_ExDy", "Psi_ExDz", "Psi_EyDx", "Psi_EyDz", "Psi_EzDx", "Psi_EzDy",
                           "Psi_HxDy", "Psi_HxDz", "Psi_HyDx", "Psi_HyDz", "Psi_HzDx", "Psi_HzDy"};
}  // namespace

// One curl component with its CPML psi pair (CompleteCPMLCurlStencil.hpp:38-80):
// offsets OL / OR of the differenced fields, valid range [start, last + end].
struct CurlCpml {
    int ol[3], orr[3], start[3], end[3];
    double constant;
    const scafes_emfdtd_params* p;
    template <typename TT>
    TT operator()(const ScaFES::Ntuple<int, 3>& idx, const ScaFES::DataField<TT, 3>& dest_old,
                  const ScaFES::DataField<TT, 3>& from_l, const ScaFES::DataField<TT, 3>& psi_l_old,
                  ScaFES::DataField<TT, 3>& psi_l_new, const ScaFES::DataField<TT, 3>& from_r,
                  const ScaFES::DataField<TT, 3>& psi_r_old, ScaFES::DataField<TT, 3>& psi_r_new) const
    {
        const int last = p->n - 1;
        int e[3];
        for (int q = 0; q < 3; ++q) {
            e[q] = last + end[q];
            if (start[q] > idx[q] || idx[q] > e[q]) return 0.;
        }
        int dlo = 0, dro = 0;
        for (int q = 0; q < 3; ++q) {
            const int md = std::min(idx[q], e[q] - idx[q]);
            dlo += md * std::abs(ol[q]);
            dro += md * std::abs(orr[q]);
        }
        const int dl = std::min(dlo, p->d), dr = std::min(dro, p->d);
        TT dfdl = from_l(ScaFES::Ntuple<int, 3>(idx[0] + ol[0], idx[1] + ol[1], idx[2] + ol[2])) - from_l(idx);
        TT psi_l_n = (p->b[dl] * psi_l_old(idx)) + (p->c[dl] * dfdl);
        dfdl = (p->ook[dl] * dfdl) + psi_l_n;
        psi_l_new(idx) = psi_l_n;
        TT dfdr = from_r(ScaFES::Ntuple<int, 3>(idx[0] + orr[0], idx[1] + orr[1], idx[2] + orr[2])) - from_r(idx);
        TT psi_r_n = (p->b[dr] * psi_r_old(idx)) + (p->c[dr] * dfdr);
        dfdr = (p->ook[dr] * dfdr) + psi_r_n;
        psi_r_new(idx) = psi_r_n;
        return dest_old(idx) + (constant * (dfdl - dfdr));
    }
};

template <typename CT>
class EMFDTDng : public ScaFES::Problem<EMFDTDng<CT>, CT, 3> {
public:
    typedef ScaFES::Problem<EMFDTDng<CT>, CT, 3> Super;
    typedef ScaFES::Ntuple<int, 3> Index;
    template <typename TT>
    using VDf = std::vector<ScaFES::DataField<TT, 3>>;

    EMFDTDng(ScaFES::Parameters const& cl, ScaFES::GridGlobal<3> const& gg, const Ini& ini)
        : Super(cl, gg, true, names(), halo(), std::vector<bool>(NDF, false), std::vector<int>(NDF, 0),
                std::vector<CT>(NDF, CT(0.)), writePolicy(),
                std::vector<bool>(NDF, false), ini.list("initialization.parameters", ' ')),
          mIni(ini), cfln_(.5), ds_(1.), dt_((cfln_ * ds_) / C_0)
    {
        // EMFDTD.hpp:184-236: CPML parameters, source and sink, tau = 1
        mP = scafes_emfdtd_params();
        mP.n = cl.nNodes().at(0);
        mP.d = (int)ini.real("PML.d");
        mP.cfln = cfln_;
        mP.dt = dt_;
        const char* pk[5] = {"PML.m_a", "PML.m", "PML.k_max", "PML.s_max", "PML.a_max"};
        for (int q = 0; q < 5; ++q) mPml[q] = ini.real(pk[q]);
        if (scafes_emfdtd_tables(mP.d, mPml[0], mPml[1], mPml[2], mPml[3], mPml[4], dt_, mP.b, mP.c, mP.ook) !=
            SCAFES_OK)
            throw std::runtime_error(scafes_last_error());
        const std::vector<double> src = ini.list("source.location", 'x'), snk = ini.list("sink.location", 'x');
        if (src.size() != 3 || snk.size() != 3) throw std::runtime_error("source / sink location needs AxBxC");
        for (int q = 0; q < 3; ++q) {
            mP.src[q] = (int)src[q];
            mP.sink[q] = (int)snk[q];
            if (mP.src[q] < 0 || mP.src[q] > mP.n - 1) throw std::runtime_error("Source not within simulation area.");
            if (mP.sink[q] < 0 || mP.sink[q] > mP.n - 1) throw std::runtime_error("Sink not within simulation area.");
        }
        if (mP.src[0] == mP.sink[0] && mP.src[1] == mP.sink[1] && mP.src[2] == mP.sink[2])
            throw std::runtime_error("Source and sink cannot be at the same position.");
        mP.f_mod = ini.real("source.f_mod");
        mP.t0 = ini.real("source.t0");
        mP.spread = ini.real("source.spread");
        if (cl.tau() != 1.0)
            throw std::runtime_error("Please choose \"--starttime\" and \"--endtime\" so that tau()=1.0.");
        // stencils (EMFDTD.hpp:167-172, 208-213)
        const int spec[6][13] = {{0, 0, 1, 0, 1, 0, 1, 0, 0, 0, -1, -1, 1},   {1, 0, 0, 0, 0, 1, 0, 1, 0, -1, 0, -1, 1},
                                 {0, 1, 0, 1, 0, 0, 0, 0, 1, -1, -1, 0, 1},  {0, -1, 0, 0, 0, -1, 0, 0, 1, -1, 0, 0, -1},
                                 {0, 0, -1, -1, 0, 0, 1, 0, 1, 0, -1, 0, -1}, {-1, 0, 0, 0, -1, 0, 1, 1, 0, 0, 0, -1, -1}};
        for (int s = 0; s < 6; ++s) {
            for (int q = 0; q < 3; ++q) {
                mCurl[s].ol[q] = spec[s][q];
                mCurl[s].orr[q] = spec[s][3 + q];
                mCurl[s].start[q] = spec[s][6 + q];
                mCurl[s].end[q] = spec[s][9 + q];
            }
            mCurl[s].constant = spec[s][12] * cfln_;
            mCurl[s].p = &mP;
        }
    }

    const scafes_emfdtd_params& params() const { return mP; }
    // CPMLParams arguments m_a, m, k_max, s_max, a_max ([PML] section)
    const double* pml() const { return mPml; }

    void evalInner(VDf<CT>&, Index const&, int const&) {}
    void evalBorder(VDf<CT>&, Index const&, int const&) {}
    // SIGMA from the sphere / ellipsoid indicator at the node indices
    // (EMFDTD.hpp:258-276, geometry/IndicatorEllipsoid.hpp:48-68, h = 1e-8)
    template <typename TT>
    void initInner(VDf<TT>& dfs, std::vector<TT> const& p, Index const& idx, int const&)
    {
        const TT inside = 3.546e7, h = 1.e-8;
        const TT fx = TT(idx[0]) - p.at(0), fy = TT(idx[1]) - p.at(1), fz = TT(idx[2]) - p.at(2);
        const TT tmp = (fx * fx) + (fy * fy);
        const TT theta = std::atan2(std::sqrt(tmp), fz), phi = std::atan2(fy, fx);
        const TT ro = std::sqrt(tmp + (fz * fz));
        const TT x1 = p.at(3) * std::cos(theta) * std::cos(phi), y1 = p.at(4) * std::cos(theta) * std::sin(phi);
        const TT z1 = p.at(5) * std::sin(theta);
        const TT rn = std::sqrt((x1 * x1) + (y1 * y1) + (z1 * z1));
        dfs[SIGMA](idx) = 0.5 * inside *
                          (((ro + rn) / (std::sqrt((ro + rn) * (ro + rn) + h * h))) -
                           ((ro - rn) / (std::sqrt((ro - rn) * (ro - rn) + h * h))));
    }
    template <typename TT>
    void initBorder(VDf<TT>& dfs, std::vector<TT> const& p, Index const& idx, int const& n)
    {
        initInner<TT>(dfs, p, idx, n);
    }
    // H from E (EMFDTD.hpp:295-313)
    template <typename TT>
    void updateInner(VDf<TT>& n, VDf<TT> const& o, Index const& idx, int const&)
    {
        n[Hx](idx) = mCurl[0](idx, o[Hx], o[Ey], o[Psi_EyDz], n[Psi_EyDz], o[Ez], o[Psi_EzDy], n[Psi_EzDy]);
        n[Hy](idx) = mCurl[1](idx, o[Hy], o[Ez], o[Psi_EzDx], n[Psi_EzDx], o[Ex], o[Psi_ExDz], n[Psi_ExDz]);
        n[Hz](idx) = mCurl[2](idx, o[Hz], o[Ex], n[Psi_ExDy], n[Psi_ExDy], o[Ey], n[Psi_EyDx], n[Psi_EyDx]);
    }
    template <typename TT>
    void updateBorder(VDf<TT>& n, VDf<TT> const& o, Index const& idx, int const& t)
    {
        updateInner(n, o, idx, t);
    }
    // D from H, then E, I, source, sink (EMFDTD.hpp:330-386)
    template <typename TT>
    void updateInner2(VDf<TT>& n, VDf<TT> const& o, Index const& idx, int const&)
    {
        TT dx = mCurl[3](idx, o[Dx], o[Hz], o[Psi_HzDy], n[Psi_HzDy], o[Hy], o[Psi_HyDz], n[Psi_HyDz]);
        TT dy = mCurl[4](idx, o[Dy], o[Hx], o[Psi_HxDz], n[Psi_HxDz], o[Hz], o[Psi_HzDx], n[Psi_HzDx]);
        TT dz = mCurl[5](idx, o[Dz], o[Hy], o[Psi_HyDx], n[Psi_HyDx], o[Hx], o[Psi_HxDy], n[Psi_HxDy]);
        TT ga = 1. / (1. + (o[SIGMA](idx)) * dt_ / EPSILON_0);
        TT gb = (o[SIGMA](idx)) * dt_ / EPSILON_0;
        TT ex = ga * (dx - o[Ix](idx)), ey = ga * (dy - o[Iy](idx)), ez = ga * (dz - o[Iz](idx));
        n[Ix](idx) = o[Ix](idx) + (gb * ex);
        n[Iy](idx) = o[Iy](idx) + (gb * ey);
        n[Iz](idx) = o[Iz](idx) + (gb * ez);
        n[Dx](idx) = dx;
        n[Dy](idx) = dy;
        n[Dz](idx) = dz;
        n[Ex](idx) = ex;
        n[Ey](idx) = ey;
        if (idx[0] == mP.src[0] && idx[1] == mP.src[1] && idx[2] == mP.src[2])
            n[Ez](idx) = scafes_emfdtd_source(o[Ez].time(), dt_, mP.f_mod, mP.t0, mP.spread);
        else
            n[Ez](idx) = ez;
    }
    template <typename TT>
    void updateBorder2(VDf<TT>& n, VDf<TT> const& o, Index const& idx, int const& t)
    {
        updateInner2(n, o, idx, t);
    }

    // EMFDTD.hpp:414-448 (single rank)
    void writeSinkData() const
    {
        const std::string fname = mIni.get("sink.data_file");
        if (fname.empty()) throw std::runtime_error("Sink file name empty.");
        std::ofstream of(fname.c_str());
        of.setf(of.scientific);
        of.precision(20);
        for (double v : this->sinkData()) of << v << std::endl;
    }

private:
    static std::vector<std::string> names() { return std::vector<std::string>(kNames, kNames + NDF); }
    // EMFDTD.hpp:105-119: SIGMA at the start, E like --nSnapshots, the rest never
    static std::vector<ScaFES::WriteHowOften> writePolicy()
    {
        std::vector<ScaFES::WriteHowOften> v(NDF, ScaFES::WriteHowOften::NEVER);
        v[SIGMA] = ScaFES::WriteHowOften::AT_START;
        v[Ex] = v[Ey] = v[Ez] = ScaFES::WriteHowOften::LIKE_GIVEN_AT_CL;
        return v;
    }
    static std::vector<int> halo()
    {
        std::vector<int> v(NDF, 0);
        for (int q = Ex; q <= Hz; ++q) v[q] = 1;
        return v;
    }
    const Ini& mIni;
    const double cfln_, ds_, dt_;
    scafes_emfdtd_params mP;
    double mPml[5];
    CurlCpml mCurl[6];
};

namespace ScaFES {
template <typename CT>
struct DeviceTraits<EMFDTDng<CT>> {
    static const bool available = true;
    static void describe(const EMFDTDng<CT>& p, DeviceUpdate& d)
    {
        const scafes_emfdtd_params& q = p.params();
        d.kernel = DeviceKernel::EMFDTD;
        for (int f = 0; f < 25; ++f) d.em[f] = f;  // DF order
        d.pmlD = q.d;
        d.cfln = q.cfln;
        d.pmlMa = p.pml()[0];
        d.pmlM = p.pml()[1];
        d.pmlKmax = p.pml()[2];
        d.pmlSmax = p.pml()[3];
        d.pmlAmax = p.pml()[4];
        for (int c = 0; c < 3; ++c) {
            d.src[c] = q.src[c];
            d.sink[c] = q.sink[c];
        }
        d.fMod = q.f_mod;
        d.t0 = q.t0;
        d.spread = q.spread;
    }
};
}  // namespace ScaFES

int main(int argc, char* argv[])
{
    ScaFES::Parameters cl(argc, argv);
    ScaFES::GridGlobal<3> grid(cl);
    Ini ini(cl.nameConfigFile());
    EMFDTDng<double> EM(cl, grid, ini);
    EM.iterateOverTime();
    EM.writeSinkData();
    for (int q = Ex; q < NDF; ++q) dumpNodes(EM.vectUnknownDfsDomOld()[q], std::string(".") + kNames[q]);
    return 0;
}
