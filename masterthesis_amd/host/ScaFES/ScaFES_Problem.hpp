// ScaFES_Problem.hpp -- the explicit time-stepping engine of the host API.
//
// Same constructor, hook interface and phase order as ScaFES::Problem
// (src/ScaFES_Problem.hpp:264-1316): a CRTP base class whose subclass
// provides evalInner/evalBorder (known fields), initInner/initBorder
// (initial state) and updateInner/updateBorder(/2) (time step).
// iterateOverTime() (reference :2558-2715) runs n = 0..nTimesteps:
//   evalKnownDfs(n) -> [n==0: reset, init hooks, old<-new]
//                      [n>0 : update phase (+ update2 phase with leapfrog)]
//   -> compErrLinf -> checkConv -> swap old/new -> times += tau.
// The init and eval hooks run on the host (once, or every step when the
// known fields depend on t).  The update phase runs on the GPU: through the
// C ABI with the kernel family the problem declares through
// ScaFES::DeviceTraits, or, in a program compiled by hipcc with the problem
// header between SCAFES_DEVICE_HOOKS_BEGIN / _END, as the problem's own
// update hooks compiled for the device (ScaFES_DeviceHooks.hpp).  There is
// no CPU update path.
#ifndef SCAFES_HOST_PROBLEM_HPP_
#define SCAFES_HOST_PROBLEM_HPP_

#include <algorithm>
#include <cmath>
#include <iomanip>
#include <iostream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "ScaFES_DataField.hpp"
#include "ScaFES_Device.hpp"
#include "ScaFES_GridGlobal.hpp"
#include "ScaFES_HostDevice.hpp"
#include "ScaFES_Parameters.hpp"
#include "ScaFES_Timer.hpp"
#if defined(__HIPCC__)
#include "ScaFES_DeviceHooks.hpp"
#endif

namespace ScaFES {

template <class OWNPRBLM, typename CT = double, std::size_t DIM = 3>
class Problem {
public:
    typedef Ntuple<int, DIM> ScaFES_IntNtuple;
    template <typename TT>
    using ScaFES_VectDf = std::vector<DataField<TT, DIM>>;

    Problem(const Parameters& params, const GridGlobal<DIM>& gg, bool useLeapfrog,
            const std::vector<std::string>& nameDatafield, const std::vector<int>& stencilWidth,
            const std::vector<bool>& isKnownDf, const std::vector<int>& nLayers = std::vector<int>(),
            const std::vector<CT>& defaultValue = std::vector<CT>(),
            const std::vector<WriteHowOften>& writeToFile = std::vector<WriteHowOften>(),
            const std::vector<bool>& computeError = std::vector<bool>(),
            const std::vector<CT>& geomparamsInit = std::vector<CT>(),
            const std::vector<bool>& checkConvergence = std::vector<bool>())
        : mParams(params), mGG(gg), mUseLeapfrog(useLeapfrog), mNames(nameDatafield),
          mStencilWidth(stencilWidth), mIsKnownDf(isKnownDf), mNlayers(nLayers),
          mDefaultValue(defaultValue), mWriteToFile(writeToFile), mComputeError(computeError),
          mParamsCurr(geomparamsInit), mCheckConvergence(checkConvergence)
    {
        // plain copies the hook accessors read (also from a device copy of this object)
        mDomainC = mGG.discreteDomain();
        mTauC = mParams.tau();
        mTimeStartC = mParams.timeIntervalStart();
        // The four update hooks are required even if unused (reference :2849-2862, :2895-2908).
        (void)&OWNPRBLM::template updateInner<CT>;
        (void)&OWNPRBLM::template updateBorder<CT>;
        (void)&OWNPRBLM::template updateInner2<CT>;
        (void)&OWNPRBLM::template updateBorder2<CT>;
        const std::size_t nDf = mNames.size();
        if (nDf == 0) throw std::runtime_error("Problem: no data fields given.");
        if (mStencilWidth.size() != nDf || mIsKnownDf.size() != nDf)
            throw std::runtime_error("Problem: sizes of nameDatafield, stencilWidth and isKnownDf differ.");
        if (mNlayers.empty()) mNlayers.assign(nDf, 0);
        if (mDefaultValue.empty()) mDefaultValue.assign(nDf, CT(0));
        if (mWriteToFile.empty()) mWriteToFile.assign(nDf, WriteHowOften::NEVER);
        if (mComputeError.empty()) mComputeError.assign(nDf, false);
        if (mCheckConvergence.empty()) mCheckConvergence.assign(nDf, false);
        if (mNlayers.size() != nDf || mDefaultValue.size() != nDf || mWriteToFile.size() != nDf ||
            mComputeError.size() != nDf || mCheckConvergence.size() != nDf)
            throw std::runtime_error("Problem: size of an optional per-field vector differs from nameDatafield.");
        for (std::size_t ii = 0; ii < nDf; ++ii) {
            if (mStencilWidth[ii] < 0) throw std::runtime_error("Problem: negative stencil width.");
            if (mNlayers[ii] != 0)
                throw std::runtime_error("Problem: border-only data fields (nLayers > 0) are not supported.");
        }
        // memAll of each field = partition +- stencil width (reference :1495-1514)
        const Grid<DIM>& dom = mGG.discreteDomain();
        for (std::size_t ii = 0; ii < nDf; ++ii) {
            const ScaFES_IntNtuple sw(mStencilWidth[ii]);
            Ntuple<double, DIM> x0 = dom.coordNodeFirst(), x1 = dom.coordNodeLast();
            for (std::size_t p = 0; p < DIM; ++p) {
                x0[p] -= mStencilWidth[ii] * dom.gridsize(p);
                x1[p] += mStencilWidth[ii] * dom.gridsize(p);
            }
            mMemAll.push_back(Grid<DIM>(dom.idxNodeFirst() - sw, dom.idxNodeLast() + sw, x0, x1));
        }
        // lumps and views: unknown fields (old/new), known fields and the
        // reference solutions of unknown fields (old/new; reference :2198-2392)
        for (std::size_t ii = 0; ii < nDf; ++ii) {
            const std::size_t ne = (std::size_t)mMemAll[ii].nNodes().size() + 1;
            if (!mIsKnownDf[ii]) {
                mUnknownIndex.push_back((int)ii);
                mLumpUnknownOld.push_back(std::vector<CT>(ne, CT(0)));
                mLumpUnknownNew.push_back(std::vector<CT>(ne, CT(0)));
            }
            const bool hasMem = mIsKnownDf[ii] || mComputeError[ii];
            mLumpKnownOld.push_back(std::vector<CT>(hasMem ? ne : 0, CT(0)));
            mLumpKnownNew.push_back(std::vector<CT>(hasMem ? ne : 0, CT(0)));
        }
        for (std::size_t u = 0; u < mUnknownIndex.size(); ++u) {
            const int ii = mUnknownIndex[u];
            mVectUnknownDfsDomOld.push_back(DataField<CT, DIM>(mNames[ii] + "OldDom", &mParams, mGG, mMemAll[ii],
                                                               mStencilWidth[ii], mLumpUnknownOld[u].data(), 1,
                                                               0, mWriteToFile[ii]));
            mVectUnknownDfsDomNew.push_back(DataField<CT, DIM>(mNames[ii] + "NewDom", &mParams, mGG, mMemAll[ii],
                                                               mStencilWidth[ii], mLumpUnknownNew[u].data(), 1,
                                                               0, mWriteToFile[ii]));
        }
        for (std::size_t ii = 0; ii < nDf; ++ii) {
            const int sw = mIsKnownDf[ii] ? mStencilWidth[ii] : 0;
            CT* po = mLumpKnownOld[ii].empty() ? 0 : mLumpKnownOld[ii].data();
            CT* pn = mLumpKnownNew[ii].empty() ? 0 : mLumpKnownNew[ii].data();
            mVectKnownDfsDomOld.push_back(DataField<CT, DIM>(mNames[ii] + "OldSolDom", &mParams, mGG,
                                                             mMemAll[ii], sw, po, 1, 0, mWriteToFile[ii]));
            mVectKnownDfsDomNew.push_back(DataField<CT, DIM>(mNames[ii] + "NewSolDom", &mParams, mGG,
                                                             mMemAll[ii], sw, pn, 1, 0, mWriteToFile[ii]));
        }
        // node-kind flags (reference setGridGlobalFlags :3902-3980)
        const Grid<DIM> kindGrid(dom.idxNodeFirst(), dom.idxNodeLast(), dom.coordNodeFirst(), dom.coordNodeLast());
        mLumpKind.assign((std::size_t)kindGrid.nNodes().size() + 1, 0);
        mKind = DataField<short, DIM>("kind", &mParams, mGG, kindGrid, 0, mLumpKind.data());
        const int nl = mParams.nLayersAtBorder();
        const GridSub<DIM>& part = mGG.partition(0);
        for (typename GridSub<DIM>::iterator it = part.begin(); it != part.end(); ++it) {
            short k = GridGlobal<DIM>::SCAFES_GLOBAL_INNER_NODE;
            for (std::size_t p = 0; p < DIM; ++p)
                if (dom.idxNodeFirst(p) + nl > (*it).elem(p) || dom.idxNodeLast(p) - nl < (*it).elem(p)) {
                    k = GridGlobal<DIM>::SCAFES_GLOBAL_BORDER_NODE;
                    break;
                }
            mKind(*it) = k;
        }
        const char* names[] = {"trace", "write", "evalKnownDf", "sync", "update", "init", "compErr", "checkConv"};
        for (int i = 0; i < 8; ++i) mWallClockTimes[names[i]] = 0.0;
        setupDataFiles();
    }

    virtual ~Problem() { releaseDevice(); }

    // -------------------------------------------------------------- accessors
    const Parameters& params() const { return mParams; }
    const GridGlobal<DIM>& globalGrid() const { return mGG; }
    const int& nTimesteps() const { return mParams.nTimesteps(); }
    // The accessors update hooks call are host+device functions reading plain
    // members (SCAFES_HD; plain functions under g++).
    SCAFES_HD const double& tau() const { return mTauC; }
    SCAFES_HD const double& gridsize(const int& p) const { return mDomainC.gridsize(p); }
    SCAFES_HD const ScaFES_IntNtuple& nNodes() const { return mDomainC.nNodes(); }
    SCAFES_HD const int& nNodes(const int& p) const { return mDomainC.nNodes(p); }
    SCAFES_HD Ntuple<double, DIM> coordinates(const ScaFES_IntNtuple& idx) const { return mDomainC.coordinates(idx); }
    SCAFES_HD double time(const int& timestep) const { return mTimeStartC + timestep * mTauC; }
    SCAFES_HD ScaFES_IntNtuple connect(const ScaFES_IntNtuple& idx, const int& dir) { return mDomainC.neigh(idx, dir); }
    SCAFES_HD const short& kind(const ScaFES_IntNtuple& idx) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return hookState()->kind[mDomainC.idxNodeTuple2ScalarStyleFortran(idx)];
#else
        return mKind[idx];
#endif
    }
    const DataField<short, DIM>& kind() const { return mKind; }
    const std::vector<CT>& vectParamsCurr() const { return mParamsCurr; }
    SCAFES_HD const CT& knownDf(const int& idx, const ScaFES_IntNtuple& node) const { return knownDfNew(idx, node); }
    SCAFES_HD const CT& knownDfNew(const int& idx, const ScaFES_IntNtuple& node) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return hookState()->knownNew[idx](node).v;
#else
        return mVectKnownDfsDomNew[idx](node);
#endif
    }
    SCAFES_HD const CT& knownDfOld(const int& idx, const ScaFES_IntNtuple& node) const
    {
#if defined(__HIP_DEVICE_COMPILE__)
        return hookState()->knownOld[idx](node).v;
#else
        return mVectKnownDfsDomOld[idx](node);
#endif
    }
    const ScaFES_VectDf<CT>& vectSol() const { return mVectKnownDfsDomNew; }
    const ScaFES_VectDf<CT>& vectKnownDfsDomNew() const { return mVectKnownDfsDomNew; }
    const ScaFES_VectDf<CT>& vectKnownDfsDomOld() const { return mVectKnownDfsDomOld; }
    const ScaFES_VectDf<CT>& vectUnknownDfsDomNew() const { return mVectUnknownDfsDomNew; }
    const ScaFES_VectDf<CT>& vectUnknownDfsDomOld() const { return mVectUnknownDfsDomOld; }
    bool useLeapfrogIntegration() const { return mUseLeapfrog; }
    bool useAsynchronMode() const { return false; }
    int myRank() const { return 0; }
    const std::vector<bool>& isKnownDf() const { return mIsKnownDf; }
    const std::vector<int>& nLayers() const { return mNlayers; }
    const std::vector<CT>& defaultValue() const { return mDefaultValue; }
    const std::map<std::string, double>& wallClockTimes() const { return mWallClockTimes; }
    // "Local Linf" errors of every compErrLinf call, in call order.
    const std::vector<double>& errorHistory() const { return mErrors; }
    // Discrete energies recorded by the Yee kernel (every energyEvery steps).
    const std::vector<double>& energyHistory() const { return mEnergies; }
    // EMFDTD: Ez at the sink node after each step (the reference's sink_,
    // EMFDTD.hpp:376-380), available after iterateOverTime().
    const std::vector<double>& sinkData() const { return mSink; }
    // Time iterate at which the loop stopped (nTimesteps if it ran to the end).
    int lastTimeIter() const { return mLastIter; }

    // Whether gradients were requested (reference :3870; --computeGradients).
    bool computeGradients() const { return mParams.computeGradients(); }

    // -------------------------------------------------------------- driver
    void iterateOverTime()
    {
#if !defined(__HIPCC__)
        if (!DeviceTraits<OWNPRBLM>::available)
            throw std::runtime_error(
                "ScaFES: this problem declares no device update kernel (specialise ScaFES::DeviceTraits for "
                "it) and this program was not compiled by hipcc, so its hooks cannot run on the GPU (compile "
                "it with hipcc, the problem header between SCAFES_DEVICE_HOOKS_BEGIN and SCAFES_DEVICE_HOOKS_END); "
                "the engine has no CPU update path.");
#endif
        // The reference records an ADOL-C tape of the hooks and propagates
        // the derivatives w.r.t. geomparamsInit alongside the iterate
        // (Problem.hpp:1391, :4598-4660).  This engine has no tape; its
        // gradients are the field adjoints of the C ABI.
        if (mParams.computeGradients())
            throw std::runtime_error(
                "ScaFES: --computeGradients (ADOL-C tapes of the hooks) is not available in this engine; "
                "use the field gradients of the C ABI (scafes_pennes_gradient, scafes_heat7var_gradient, "
                "scafes_yee_eps_gradient).");
        mDev = DeviceUpdate();
        if (DeviceTraits<OWNPRBLM>::available)
            DeviceTraits<OWNPRBLM>::describe(static_cast<const OWNPRBLM&>(*this), mDev);
        else
            mDev.kernel = DeviceKernel::HOOKS;  // the problem's own hooks on the device
        validateDevice();
        Timer timerIterate;
        if (out()) std::cout << mParams.getPrefix() << " * Iterate over time..." << std::endl;
        mParams.decreaseLevel();
        bool cont = true;
        mLastIter = mParams.nTimesteps();
        for (int n = 0; n <= mParams.nTimesteps(); ++n) {
            if (out()) std::cout << mParams.getPrefix() << " * Process TIME ITERATE = " << n << "..." << std::endl;
            evalKnownDfs(n);
            if (0 < n) {
                evalUpdatePhase(n);
                if (n == 1) checkTraitsAgainstHooks(n);
            } else {
                Timer t;
                resetValsAtNodesOfUnknownDfs();
                resetTimesOfUnknownDfs();
                evalInitPbPhase();
                deviceSetup();
                mWallClockTimes["init"] += t.elapsed();
            }
            writeAllDfs(n);
            compErrOfUnknownDfs(n);
            cont = checkConvOfUnknownDfs(n);
            swapPointersOfKnownDfs();
            swapPointersOfUnknownDfs();
            updateTimesOfKnownDfs();
            updateTimesOfUnknownDfs();
            if (!cont) {
                if (out())
                    std::cout << mParams.getPrefix() << " * Convergence after " << n << " of "
                              << mParams.nTimesteps() << " timesteps." << std::endl;
                mLastIter = n;
                writeAllDfs(mParams.nTimesteps());  // reference :2650-2660, after the swaps
                break;
            }
        }
        deviceDownload();
        const double tIterate = timerIterate.elapsed();
        mParams.increaseLevel();
        if (out()) printTimes(tIterate);
        releaseDevice();
    }

    // Device buffers hold the context's handle: free them before the context.
    void releaseDevice()
    {
        mUOld = DeviceBuffer();
        mUNew = DeviceBuffer();
        mCOld = DeviceBuffer();
        mCNew = DeviceBuffer();
        mKnownA = DeviceBuffer();
        mKnownB = DeviceBuffer();
        mKnownC = DeviceBuffer();
        mYee = DeviceBuffer();
        mYee2 = DeviceBuffer();
        mEm = DeviceBuffer();
        mEmSink = DeviceBuffer();
        mMat.clear();
        mRef.clear();
        mHookOld.clear();
        mHookNew.clear();
        mHookKnownOld.clear();
        mHookKnownNew.clear();
        mHookKind = DeviceBuffer();
        mHookStateBuf = DeviceBuffer();
        mHookSelf = DeviceBuffer();
        mDevHookState = 0;
        mCtx.reset();
    }

private:
    bool out() const { return mParams.rankOutput() == 0 && 0 < mParams.indentDepth(); }
    std::size_t nUnknown() const { return mUnknownIndex.size(); }
    int unknownSlot(int df) const
    {
        for (std::size_t u = 0; u < mUnknownIndex.size(); ++u)
            if (mUnknownIndex[u] == df) return (int)u;
        return -1;
    }
    bool evalDf() const
    {
        for (std::size_t ii = 0; ii < mNames.size(); ++ii)
            if (mComputeError[ii] || mIsKnownDf[ii]) return true;
        return false;
    }

    // ---------------------------------------------------- known data fields
    // evalKnownDfs (reference :3433-3485): evalInner/evalBorder at every node.
    void evalKnownDfs(const int& n)
    {
        if (!evalDf()) return;
        if (0 < n && !mDev.knownTimeDependent && !anyComputeError()) return;
        Timer t;
        if (0 < n && mDev.knownGen != DeviceKnownGen::NONE) {
            evalKnownOnDevice(n);
            mWallClockTimes["evalKnownDf"] += t.elapsed();
            return;
        }
        const GridSub<DIM>& part = mGG.partition(0);
        OWNPRBLM* self = static_cast<OWNPRBLM*>(this);
        for (typename GridSub<DIM>::iterator it = part.begin(); it != part.end(); ++it) {
            if (mKind[*it] & GridGlobal<DIM>::SCAFES_GLOBAL_INNER_NODE) self->evalInner(mVectKnownDfsDomNew, *it, n);
            else self->evalBorder(mVectKnownDfsDomNew, *it, n);
        }
        if (0 < n) uploadKnown(n);
        mWallClockTimes["evalKnownDf"] += t.elapsed();
    }
    // The kernel reads F from the old known buffer and G from the new one
    // (flux borders also G_old): the values the hooks produced at steps n-1
    // and n, i.e. at time(n-1) and time(n).  The exact solution for compErr
    // is the one at time(n).
    void evalKnownOnDevice(const int& n)
    {
        scafes_timelin_params p;
        p.form = (int)mDev.knownGen;
        p.border_mode = (int)mDev.border;
        const Ntuple<double, DIM> x0 = coordinates(ScaFES_IntNtuple(0));
        for (int q = 0; q < 3; ++q) p.x0[q] = q < (int)DIM ? x0.elem(q) : 0.0;
        p.rho_c = mDev.rhoC;
        p.rhob_cb_w = mDev.rhobCbW;
        p.lambda = mDev.lambda;
        p.alpha = mDev.alpha;
        p.epsilon = mDev.epsilon;
        p.sigma = mDev.sigma;
        const scafes_geom g = geom();
        scafes_ctx* c = mCtx->get();
        double* exact = 0;
        const int u = unknownSlot(mDev.unknown);
        if (u >= 0 && (std::size_t)u < mRef.size()) exact = (double*)mRef[u].get();
        deviceCheck(scafes_known_timelin_f64(c, &g, &p, time(n - 1), (double*)mKnownA.get(),
                                             fluxBorder() ? (double*)mKnownC.get() : 0, 0, 0),
                    "known fields (old)");
        // the reference solution has the unknown field's memAll layout
        deviceCheck(scafes_known_timelin_f64(c, &g, &p, time(n), 0, (double*)mKnownB.get(), exact,
                                             mStencilWidth[mDev.unknown]),
                    "known fields (new)");
    }
    bool anyComputeError() const
    {
        for (std::size_t ii = 0; ii < mNames.size(); ++ii)
            if (mComputeError[ii] && !mIsKnownDf[ii]) return true;
        return false;
    }

    // ---------------------------------------------------- init phase
    void resetValsAtNodesOfUnknownDfs()
    {
        for (std::size_t u = 0; u < nUnknown(); ++u) {
            mVectUnknownDfsDomOld[u] = mDefaultValue[mUnknownIndex[u]];
            mVectUnknownDfsDomNew[u] = mDefaultValue[mUnknownIndex[u]];
        }
    }
    void resetTimesOfUnknownDfs()
    {
        for (std::size_t u = 0; u < nUnknown(); ++u) mVectUnknownDfsDomNew[u].time() = mParams.timeIntervalStart();
    }
    // evalInitPbPhase (reference :2754-2826): init hooks, then old <- new.
    void evalInitPbPhase()
    {
        OWNPRBLM* self = static_cast<OWNPRBLM*>(this);
        for (std::size_t u = 0; u < nUnknown(); ++u) mVectUnknownDfsDomNew[u].time() = mParams.timeIntervalStart();
        const GridSub<DIM>& part = mGG.partition(0);
        if (mParams.readInitFile()) {
            readDfsFromFile();  // instead of the init hooks (reference :2763-2795)
        } else {
            for (typename GridSub<DIM>::iterator it = part.begin(); it != part.end(); ++it) {
                if (mKind[*it] & GridGlobal<DIM>::SCAFES_GLOBAL_INNER_NODE)
                    self->template initInner<CT>(mVectUnknownDfsDomNew, mParamsCurr, *it, 0);
                else
                    self->template initBorder<CT>(mVectUnknownDfsDomNew, mParamsCurr, *it, 0);
            }
        }
        for (std::size_t u = 0; u < nUnknown(); ++u) mVectUnknownDfsDomOld[u].setValuesAtMemAll(mVectUnknownDfsDomNew[u]);
        for (std::size_t ii = 0; ii < mNames.size(); ++ii)
            if (!mLumpKnownNew[ii].empty()) mVectKnownDfsDomOld[ii].setValuesAtMemAll(mVectKnownDfsDomNew[ii]);
    }

    // ---------------------------------------------------- snapshot files
    WriteHowOften policy(std::size_t ii) const
    {
        return ii < mWriteToFile.size() ? mWriteToFile[ii] : WriteHowOften::NEVER;
    }
    // Variables of <outputfile>.nc and <initfile>.nc (reference :1760-1995).
    // Known fields are looked up in the all-fields vector of known views by
    // a counter of known fields, as the reference does.
    void setupDataFiles()
    {
        std::vector<std::string> names;
        std::size_t ik = 0, iu = 0;
        for (std::size_t ii = 0; ii < mNames.size(); ++ii) {
            if (mIsKnownDf[ii]) {
                if (policy(ii) != WriteHowOften::NEVER) names.push_back(mVectKnownDfsDomNew[ik].name());
                ++ik;
            } else {
                if (policy(ii) != WriteHowOften::NEVER) names.push_back(mVectUnknownDfsDomNew[iu].name());
                ++iu;
            }
        }
        mCommonFile = DataFile<CT, DIM>(mParams.nameDataFile(), names, nNodes());
        if (mParams.readInitFile()) {
            std::vector<std::string> init;
            for (std::size_t ii = 0; ii < mNames.size(); ++ii)
                if (!mIsKnownDf[ii] || mParams.initKnownDfs()) init.push_back(mNames[ii]);
            mInitFile = DataFile<CT, DIM>(mParams.nameInitFile(), init, nNodes());
        }
    }
    // node block (x fastest) of a view / back into a view
    std::vector<CT> nodeBlock(const DataField<CT, DIM>& df) const
    {
        std::vector<CT> v;
        v.reserve((std::size_t)nNodes().size());
        const GridSub<DIM>& part = mGG.partition(0);
        for (typename GridSub<DIM>::iterator it = part.begin(); it != part.end(); ++it) v.push_back(df(*it));
        return v;
    }
    // writeDfsToFile (reference :4256-4420): one record when some field's
    // policy selects this time iterate; intervals from --nSnapshots.  The
    // data of the k-th written known field is the k-th entry of the known
    // views (the reference's counter advances on written known fields only).
    void writeAllDfs(const int& n)
    {
        Timer t;
        int interval = mParams.nTimesteps();
        if (0 < mParams.nSnapshots() && mParams.nSnapshots() <= mParams.nTimesteps())
            interval = mParams.nTimesteps() / mParams.nSnapshots();
        if (interval < 1) interval = 1;
        std::vector<bool> writeData;
        std::vector<std::size_t> fieldOf;
        for (std::size_t ii = 0; ii < mNames.size(); ++ii) {
            const WriteHowOften w = policy(ii);
            if (w == WriteHowOften::NEVER) continue;
            bool on = false;
            switch (w) {
            case WriteHowOften::AT_START: on = n == 0; break;
            case WriteHowOften::AT_END: on = n == mParams.nTimesteps(); break;
            case WriteHowOften::AT_START_AND_END: on = n == 0 || n == mParams.nTimesteps(); break;
            case WriteHowOften::LIKE_GIVEN_AT_CL: on = n % interval == 0; break;
            default: on = true; break;
            }
            writeData.push_back(on);
            fieldOf.push_back(ii);
        }
        bool any = false;
        for (bool b : writeData) any = any || b;
        if (!any) return;
        std::vector<std::vector<CT>> blocks(writeData.size());
        std::vector<const CT*> ptrs(writeData.size(), 0);
        std::size_t ik = 0;
        for (std::size_t k = 0; k < fieldOf.size(); ++k) {
            const std::size_t ii = fieldOf[k];
            if (mIsKnownDf[ii]) {
                const std::size_t src = ik++;
                if (!writeData[k]) continue;
                if (!mVectKnownDfsDomNew[src].elemData())
                    throw std::runtime_error("writeDfsToFile: known field " + mNames[ii] +
                                             " maps to an unallocated reference solution (reference indexing)");
                blocks[k] = nodeBlock(mVectKnownDfsDomNew[src]);
            } else {
                if (!writeData[k]) continue;
                const std::size_t u = (std::size_t)unknownSlot((int)ii);
                if (mCtx) downloadUnknownTo(u, mVectUnknownDfsDomNew[u]);
                blocks[k] = nodeBlock(mVectUnknownDfsDomNew[u]);
            }
            ptrs[k] = blocks[k].data();
        }
        mCommonFile.write(ptrs, writeData);
        mWallClockTimes["write"] += t.elapsed();
    }
    // readDfsFromFile (reference :4427-4520): record 0 of each unknown field
    // (and known field with --initKnownDfs=1) by its plain name.
    void readDfsFromFile()
    {
        const std::size_t nn = (std::size_t)nNodes().size();
        std::vector<std::vector<CT>> blocks;
        std::vector<CT*> ptrs;
        std::vector<DataField<CT, DIM>*> dst;
        std::size_t ik = 0, iu = 0;
        for (std::size_t ii = 0; ii < mNames.size(); ++ii) {
            if (!mIsKnownDf[ii]) {
                dst.push_back(&mVectUnknownDfsDomNew[iu++]);
            } else if (mParams.initKnownDfs()) {
                DataField<CT, DIM>* d = &mVectKnownDfsDomNew[ik++];
                if (!d->elemData())
                    throw std::runtime_error("readDfsFromFile: known field " + mNames[ii] +
                                             " maps to an unallocated reference solution (reference indexing)");
                dst.push_back(d);
            }
        }
        blocks.assign(dst.size(), std::vector<CT>(nn));
        for (std::size_t k = 0; k < dst.size(); ++k) ptrs.push_back(blocks[k].data());
        mInitFile.read(ptrs);
        const GridSub<DIM>& part = mGG.partition(0);
        for (std::size_t k = 0; k < dst.size(); ++k) {
            std::size_t q = 0;
            for (typename GridSub<DIM>::iterator it = part.begin(); it != part.end(); ++it) (*dst[k])(*it) = blocks[k][q++];
        }
    }

    // ---------------------------------------------------- device set-up
    void validateDevice()
    {
        const int nDf = (int)mNames.size();
        auto known = [&](int ii, const char* what) {
            if (ii < 0 || ii >= nDf || !mIsKnownDf[ii])
                throw std::runtime_error(std::string("DeviceUpdate: ") + what + " must name a known data field");
        };
        auto unknown = [&](int ii) {
            if (ii < 0 || ii >= nDf || mIsKnownDf[ii])
                throw std::runtime_error("DeviceUpdate: the updated field must be an unknown data field");
            if (mStencilWidth[ii] < 1) throw std::runtime_error("DeviceUpdate: updated field needs stencil width >= 1");
        };
        switch (mDev.kernel) {
        case DeviceKernel::HEAT7:
            unknown(mDev.unknown);
            if (mDev.precision != DevicePrecision::F64) throw std::runtime_error("DeviceUpdate: HEAT7 runs in FP64");
            if (mDev.sourceKnown >= 0) known(mDev.sourceKnown, "sourceKnown");
            if (mDev.border == DeviceBorder::KNOWN_FIELD || fluxBorder()) known(mDev.borderKnown, "borderKnown");
            if (fluxBorder() && mDev.lambda == 0.0) throw std::runtime_error("DeviceUpdate: flux borders need lambda != 0");
            if (mDev.knownGen != DeviceKnownGen::NONE) {
                if (mDev.sourceKnown < 0 || !(mDev.border == DeviceBorder::KNOWN_FIELD || fluxBorder()))
                    throw std::runtime_error("DeviceUpdate: device known fields need a source F and a G border field");
                for (std::size_t ii = 0; ii < mNames.size(); ++ii)
                    if (mIsKnownDf[ii] && mStencilWidth[ii] != 0)
                        throw std::runtime_error("DeviceUpdate: device known fields have stencil width 0");
            }
            break;
        case DeviceKernel::HEAT7VAR:
            unknown(mDev.unknown);
            if (fluxBorder()) throw std::runtime_error("DeviceUpdate: flux borders are HEAT7 only");
            known(mDev.kappaKnown, "kappaKnown");
            if (mStencilWidth[mDev.kappaKnown] != mStencilWidth[mDev.unknown])
                throw std::runtime_error("DeviceUpdate: kappa needs the stencil width of the updated field");
            if (mDev.border != DeviceBorder::KEEP && mDev.border != DeviceBorder::VALUE)
                throw std::runtime_error("DeviceUpdate: HEAT7VAR supports KEEP or VALUE borders");
            break;
        case DeviceKernel::PENNES:
            unknown(mDev.unknown);
            if (!mDev.tumorFields) {
                known(mDev.perfusionKnown, "perfusionKnown");
                known(mDev.sourceFieldKnown, "sourceFieldKnown");
                if (mStencilWidth[mDev.perfusionKnown] != mStencilWidth[mDev.unknown] ||
                    mStencilWidth[mDev.sourceFieldKnown] != mStencilWidth[mDev.unknown])
                    throw std::runtime_error(
                        "DeviceUpdate: perfusion and source need the stencil width of the updated field");
            }
            if (mDev.border != DeviceBorder::KEEP && mDev.border != DeviceBorder::TUMOR)
                throw std::runtime_error("DeviceUpdate: PENNES supports KEEP or TUMOR borders");
            if (mDev.border == DeviceBorder::TUMOR && mDev.kCond == 0.0)
                throw std::runtime_error("DeviceUpdate: the TUMOR border needs kCond != 0");
            if (mDev.precision == DevicePrecision::F32_KAHAN)
                throw std::runtime_error("DeviceUpdate: PENNES runs in FP64 or FP32");
            break;
        case DeviceKernel::YEE:
            if (!mUseLeapfrog) throw std::runtime_error("DeviceUpdate: YEE needs the leapfrog mode");
            for (int c = 0; c < 6; ++c) unknown(mDev.yee[c]);
            if (DIM != 3) throw std::runtime_error("DeviceUpdate: YEE is three-dimensional");
            if (mDev.precision == DevicePrecision::F32_KAHAN)
                throw std::runtime_error("DeviceUpdate: YEE runs in FP64 or FP32");
            for (std::size_t p = 1; p < DIM; ++p)
                if (nNodes(p) != nNodes(0)) throw std::runtime_error("DeviceUpdate: YEE needs a cubic grid");
            break;
        case DeviceKernel::EMFDTD:
            if (!mUseLeapfrog) throw std::runtime_error("DeviceUpdate: EMFDTD needs the leapfrog mode");
            if (DIM != 3) throw std::runtime_error("DeviceUpdate: EMFDTD is three-dimensional");
            if (mDev.precision != DevicePrecision::F64) throw std::runtime_error("DeviceUpdate: EMFDTD runs in FP64");
            for (std::size_t p = 1; p < DIM; ++p)
                if (nNodes(p) != nNodes(0)) throw std::runtime_error("DeviceUpdate: EMFDTD needs a cubic grid");
            for (int q = 0; q < 25; ++q)
                if (mDev.em[q] < 0 || mDev.em[q] >= nDf || mIsKnownDf[mDev.em[q]])
                    throw std::runtime_error("DeviceUpdate: EMFDTD fields must be unknown data fields");
            if (mDev.pmlD < 2 || mDev.pmlD > SCAFES_EMFDTD_MAX_D)
                throw std::runtime_error("DeviceUpdate: EMFDTD needs 2 <= PML depth <= 64");
            for (int q = 0; q < 3; ++q)
                if (mDev.src[q] < 0 || mDev.src[q] >= nNodes(0) || mDev.sink[q] < 0 || mDev.sink[q] >= nNodes(0))
                    throw std::runtime_error("DeviceUpdate: EMFDTD source or sink outside the grid");
            break;
        case DeviceKernel::MRI:
            unknown(mDev.unknown);
            if (mDev.precision != DevicePrecision::F64) throw std::runtime_error("DeviceUpdate: MRI runs in FP64");
            for (int q = 0; q < SCAFES_MRI_NFIELDS; ++q) {
                known(mDev.mri[q], "an MRI material field");
                if (mStencilWidth[mDev.mri[q]] != 0)
                    throw std::runtime_error("DeviceUpdate: MRI material fields have stencil width 0");
            }
            break;
        case DeviceKernel::HOOKS:
            if (mNames.size() > kMaxDeviceFields)
                throw std::runtime_error("device hooks: at most 32 data fields");
            if (mDev.precision != DevicePrecision::F64)
                throw std::runtime_error("device hooks: the problem's own hooks run in its CT");
            break;
        default:
            throw std::runtime_error("DeviceUpdate: no kernel selected");
        }
        if (mDev.knownGen != DeviceKnownGen::NONE && mDev.kernel != DeviceKernel::HEAT7)
            throw std::runtime_error("DeviceUpdate: device known fields are a HEAT7 option");
        if (mDev.knownGen != DeviceKnownGen::NONE)
            for (std::size_t ii = 0; ii < mNames.size(); ++ii)
                if (mIsKnownDf[ii] && policy(ii) != WriteHowOften::NEVER)
                    throw std::runtime_error("DeviceUpdate: known fields evaluated on the device are not written to "
                                             "file (their host copies keep the n = 0 values)");
    }

    scafes_geom geom() const
    {
        scafes_geom g;
        g.dim = (int)DIM;
        for (int p = 0; p < 3; ++p) {
            g.n[p] = p < (int)DIM ? nNodes(p) : 1;
            g.h[p] = p < (int)DIM ? gridsize(p) : 1.0;
        }
        g.nlayers = mParams.nLayersAtBorder();
        return g;
    }
    std::size_t nodesOf(int ii) const { return (std::size_t)mMemAll[ii].nNodes().size(); }
    bool f32() const { return mDev.precision != DevicePrecision::F64; }
    bool fluxBorder() const
    {
        return mDev.border == DeviceBorder::NEUMANN || mDev.border == DeviceBorder::CAUCHY ||
               mDev.border == DeviceBorder::RADIATION;
    }
    double shift() const { return (mDev.kernel == DeviceKernel::PENNES && f32()) ? mDev.tA : 0.0; }

    // host lump (memAll + time slot) -> device array (memAll) of the field's dtype
    void put(DeviceBuffer& d, const std::vector<CT>& lump, std::size_t n, bool asFloat, double sub = 0.0)
    {
        if (asFloat) {
            std::vector<float> tmp(n);
            for (std::size_t i = 0; i < n; ++i) tmp[i] = (float)((double)lump[i] - sub);
            d.upload(tmp.data(), n * sizeof(float));
        } else {
            std::vector<double> tmp(n);
            for (std::size_t i = 0; i < n; ++i) tmp[i] = (double)lump[i] - sub;
            d.upload(tmp.data(), n * sizeof(double));
        }
    }
    DeviceBuffer alloc(std::size_t n, bool asFloat) { return DeviceBuffer(*mCtx, n * (asFloat ? 4 : 8)); }

    void deviceSetup()
    {
        mCtx.reset(new DeviceContext(mParams.device()));
        if (mDev.kernel == DeviceKernel::HOOKS) {
            hooksSetup();
            return;
        }
        const bool fl = f32();
        if (mDev.kernel == DeviceKernel::EMFDTD) {
            emSetup();
        } else if (mDev.kernel == DeviceKernel::YEE) {
            const std::size_t nn = (std::size_t)nNodes().size();
            const std::size_t es = fl ? sizeof(float) : sizeof(double);
            mYee = DeviceBuffer(*mCtx, 6 * nn * es);
            mYee2 = DeviceBuffer(*mCtx, 6 * nn * es);  // out-of-place fused step
            if (fl) {
                std::vector<float> tmp(6 * nn);
                for (int c = 0; c < 6; ++c)
                    packNodes(mLumpUnknownNew[unknownSlot(mDev.yee[c])], mDev.yee[c], &tmp[c * nn]);
                mYee.upload(tmp.data(), tmp.size() * es);
            } else {
                std::vector<double> tmp(6 * nn);
                for (int c = 0; c < 6; ++c)
                    packNodes(mLumpUnknownNew[unknownSlot(mDev.yee[c])], mDev.yee[c], &tmp[c * nn]);
                mYee.upload(tmp.data(), tmp.size() * es);
            }
        } else {
            const int ii = mDev.unknown;
            const int u = unknownSlot(ii);
            const std::size_t n = nodesOf(ii);
            mUOld = alloc(n, fl);
            mUNew = alloc(n, fl);
            put(mUOld, mLumpUnknownOld[u], n, fl, shift());
            put(mUNew, mLumpUnknownNew[u], n, fl, shift());
            if (mDev.precision == DevicePrecision::F32_KAHAN) {
                mCOld = alloc(n, true);
                mCNew = alloc(n, true);
                deviceCheck(scafes_fill(mCtx->get(), mCOld.get(), SCAFES_F32, n, 0.0), "fill");
                deviceCheck(scafes_fill(mCtx->get(), mCNew.get(), SCAFES_F32, n, 0.0), "fill");
            }
        }
        uploadKnown(0);
    }

    // EMFDTD state: the 24 updated fields' node blocks, then ga and gb from
    // SIGMA (long double, as the reference's hook evaluates them); CPML
    // tables from the CPMLParams arguments.
    void emSetup()
    {
        const std::size_t nn = (std::size_t)nNodes().size();
        std::vector<double> tmp(26 * nn), sigma(nn);
        for (int q = 1; q < 25; ++q) packNodes(mLumpUnknownNew[unknownSlot(mDev.em[q])], mDev.em[q], &tmp[(q - 1) * nn]);
        packNodes(mLumpUnknownNew[unknownSlot(mDev.em[0])], mDev.em[0], sigma.data());
        mEmP = scafes_emfdtd_params();
        mEmP.n = nNodes(0);
        mEmP.d = mDev.pmlD;
        mEmP.cfln = mDev.cfln;
        mEmP.dt = scafes_emfdtd_dt(mDev.cfln);
        deviceCheck(scafes_emfdtd_tables(mDev.pmlD, mDev.pmlMa, mDev.pmlM, mDev.pmlKmax, mDev.pmlSmax, mDev.pmlAmax,
                                         mEmP.dt, mEmP.b, mEmP.c, mEmP.ook),
                    "CPML tables");
        deviceCheck(scafes_emfdtd_ga_gb(nn, sigma.data(), mEmP.dt, &tmp[24 * nn], &tmp[25 * nn]), "ga/gb");
        for (int q = 0; q < 3; ++q) {
            mEmP.src[q] = mDev.src[q];
            mEmP.sink[q] = mDev.sink[q];
        }
        mEmP.f_mod = mDev.fMod;
        mEmP.t0 = mDev.t0;
        mEmP.spread = mDev.spread;
        mEmP.flags = 0;
        mEm = DeviceBuffer(*mCtx, tmp.size() * sizeof(double));
        mEm.upload(tmp.data(), tmp.size() * sizeof(double));
        mEmSink = DeviceBuffer(*mCtx, (std::size_t)std::max(1, mParams.nTimesteps()) * sizeof(double));
        mEmSteps = 0;
        int ok = 0;
        deviceCheck(scafes_emfdtd_vacuum_ok(mCtx->get(), &mEmP, (const double*)mEm.get(), &ok), "vacuum check");
        if (ok) mEmP.flags |= SCAFES_EMFDTD_VACUUM_OK;
    }

    // Known fields the kernel reads (and reference solutions for compErr).
    void uploadKnown(const int& n)
    {
        if (!mCtx) return;
        if (mDev.kernel == DeviceKernel::HOOKS) {
            // every lump with memory, as evalKnownDfs just rewrote them on the host
            for (std::size_t ii = 0; ii < mNames.size(); ++ii)
                if (mHookKnownNew[ii].get())
                    mHookKnownNew[ii].upload(mVectKnownDfsDomNew[ii].elemData(), nodesOf((int)ii) * sizeof(CT));
            return;
        }
        const bool fl = f32();
        auto up = [&](int ii, DeviceBuffer& d, bool old, bool asFloat) {
            const std::size_t nE = nodesOf(ii);
            if (!d.get()) d = alloc(nE, asFloat);
            put(d, old ? mLumpKnownOldView(ii) : mLumpKnownNewView(ii), nE, asFloat);
        };
        const bool firstOrTime = (n == 0) || mDev.knownTimeDependent;
        if (firstOrTime) {
            if (mDev.kernel == DeviceKernel::HEAT7) {
                // F from the old known buffer (knownDfOld), G from the new one (knownDf)
                if (mDev.sourceKnown >= 0) up(mDev.sourceKnown, mKnownA, n > 0, false);
                if (mDev.border == DeviceBorder::KNOWN_FIELD || fluxBorder())
                    up(mDev.borderKnown, mKnownB, false, false);
                if (fluxBorder()) up(mDev.borderKnown, mKnownC, n > 0, false);
            } else if (mDev.kernel == DeviceKernel::HEAT7VAR) {
                up(mDev.kappaKnown, mKnownA, false, fl);
            } else if (mDev.kernel == DeviceKernel::MRI) {
                mriUpload();
            } else if (mDev.kernel == DeviceKernel::PENNES) {
                if (mDev.tumorFields) {
                    if (n == 0) tumorFieldsUpload();
                } else {
                    up(mDev.perfusionKnown, mKnownA, false, fl);
                    up(mDev.sourceFieldKnown, mKnownB, false, fl);
                }
            }
        }
        // reference solutions of unknown fields, for compErrLinf on the device
        for (std::size_t u = 0; u < nUnknown(); ++u) {
            const int ii = mUnknownIndex[u];
            if (mComputeError[ii] && !fl && mDev.kernel != DeviceKernel::YEE && mDev.kernel != DeviceKernel::EMFDTD) {
                if (mRef.size() < nUnknown()) mRef.resize(nUnknown());
                up(ii, mRef[u], false, false);
            }
        }
    }
    // The nine material fields (node blocks, stencil width 0) and their
    // per-node coefficients for this tau, laid out like the updated field.
    void mriUpload()
    {
        if (mMat.size() != SCAFES_MRI_NFIELDS) mMat.resize(SCAFES_MRI_NFIELDS);
        const double* ptr[SCAFES_MRI_NFIELDS];
        for (int q = 0; q < SCAFES_MRI_NFIELDS; ++q) {
            const int ii = mDev.mri[q];
            const std::size_t nE = nodesOf(ii);
            if (!mMat[q].get()) mMat[q] = alloc(nE, false);
            put(mMat[q], mLumpKnownNewView(ii), nE, false);
            ptr[q] = (const double*)mMat[q].get();
        }
        const std::size_t ne = nodesOf(mDev.unknown);
        if (!mKnownA.get()) mKnownA = alloc(SCAFES_MRI_NCOEF * ne, false);
        const scafes_geom g = geom();
        const scafes_mri_params p = mriParams();
        deviceCheck(scafes_mri_coef_f64(mCtx->get(), &g, mStencilWidth[mDev.unknown], ptr, &p,
                                        (double*)mKnownA.get()),
                    "mri coefficients");
    }
    scafes_mri_params mriParams() const
    {
        scafes_mri_params p;
        p.tau = mParams.tau();
        p.h_conv = mDev.hConv;
        p.t_amb = mDev.tInf;
        p.q_bc = mDev.qBc;
        p.q_skull = mDev.qSkull;
        p.epsilon = mDev.epsilon;
        p.sigma = mDev.sigma;
        return p;
    }
    // omega and Q per node of the updated field's memAll, selected by the
    // tumor sphere like TumorHeatEqnFDM::updateInner (:239-254): squared
    // distance summed over the axes in order, compared with radius^2.
    void tumorFieldsUpload()
    {
        const int ii = mDev.unknown;
        const std::size_t ne = nodesOf(ii);
        std::vector<CT> w(ne, CT(0)), q(ne, CT(0));
        const GridSub<DIM>& part = mGG.partition(0);
        for (typename GridSub<DIM>::iterator it = part.begin(); it != part.end(); ++it) {
            const Ntuple<double, DIM> x = coordinates(*it);
            CT distance = 0.0;
            for (std::size_t pp = 0; pp < DIM; ++pp)
                distance += (x[pp] - mDev.tumorCenter[pp]) * (x[pp] - mDev.tumorCenter[pp]);
            const bool in = distance <= (mDev.tumorRadius * mDev.tumorRadius);
            unsigned long s = 0;
            mMemAll[ii].idxNodeTuple2Scalar(s, *it);
            w[s] = in ? mDev.omegaIn : mDev.omegaOut;
            q[s] = in ? mDev.qIn : mDev.qOut;
        }
        const bool fl = f32();
        mKnownA = alloc(ne, fl);
        mKnownB = alloc(ne, fl);
        put(mKnownA, w, ne, fl);
        put(mKnownB, q, ne, fl);
    }
    // Known lumps as seen through the (swapped) views.
    const std::vector<CT>& mLumpKnownOldView(int ii) const { return lumpOf(mVectKnownDfsDomOld[ii], ii); }
    const std::vector<CT>& mLumpKnownNewView(int ii) const { return lumpOf(mVectKnownDfsDomNew[ii], ii); }
    const std::vector<CT>& lumpOf(const DataField<CT, DIM>& df, int ii) const
    {
        return df.elemData() == mLumpKnownOld[ii].data() ? mLumpKnownOld[ii] : mLumpKnownNew[ii];
    }

    // node block (without ghosts) of an unknown lump, as FP32 or FP64
    template <typename T>
    void packNodes(const std::vector<CT>& lump, int ii, T* dst) const
    {
        const GridSub<DIM>& part = mGG.partition(0);
        const Grid<DIM>& m = mMemAll[ii];
        std::size_t k = 0;
        for (typename GridSub<DIM>::iterator it = part.begin(); it != part.end(); ++it) {
            unsigned long s = 0;
            m.idxNodeTuple2Scalar(s, *it);
            dst[k++] = (T)lump[s];
        }
    }
    template <typename T>
    void unpackNodes(std::vector<CT>& lump, int ii, const T* src) const
    {
        const GridSub<DIM>& part = mGG.partition(0);
        const Grid<DIM>& m = mMemAll[ii];
        std::size_t k = 0;
        for (typename GridSub<DIM>::iterator it = part.begin(); it != part.end(); ++it) {
            unsigned long s = 0;
            m.idxNodeTuple2Scalar(s, *it);
            lump[s] = (CT)src[k++];
        }
    }

    // ---------------------------------------------------- update phase
    void evalUpdatePhase(const int& n)
    {
        Timer t;
        // time of the new iterate = old + tau before the update (reference :5419-5423)
        for (std::size_t u = 0; u < nUnknown(); ++u)
            mVectUnknownDfsDomNew[u].time() = mVectUnknownDfsDomOld[u].time() + mParams.tau();
        scafes_ctx* c = mCtx->get();
        const scafes_geom g = geom();
        const int hw = mStencilWidth[mDev.unknown];
        switch (mDev.kernel) {
        case DeviceKernel::HEAT7: {
            scafes_heat7_params p;
            p.tau = mParams.tau();
            p.coeff_a = mDev.coeffA;
            p.rhoc_inv = mDev.rhoCInv;
            p.perf_c = mDev.perfC;
            p.perf_w = mDev.perfW;
            p.has_perf = mDev.perfusion ? 1 : 0;
            p.has_source = mDev.sourceKnown >= 0 ? 1 : 0;
            p.border_mode = (int)mDev.border;
            p.border_value = mDev.borderValue;
            p.lambda = mDev.lambda;
            p.alpha = mDev.alpha;
            p.epsilon = mDev.epsilon;
            p.sigma = mDev.sigma;
            deviceCheck(scafes_heat7_step_f64(c, &g, hw, (const double*)mUOld.get(), (double*)mUNew.get(),
                                              (const double*)mKnownA.get(), (const double*)mKnownC.get(),
                                              (const double*)mKnownB.get(), &p),
                        "heat7 step");
            break;
        }
        case DeviceKernel::HEAT7VAR: {
            int which = 0;
            deviceCheck(scafes_heat7var_run(c, &g, hw, (int)mDev.precision, mUOld.get(), mUNew.get(), mCOld.get(),
                                            mCNew.get(), mKnownA.get(), mParams.tau(), 1, (int)mDev.border,
                                            mDev.borderValue, -1, &which),
                        "heat7var step");
            break;
        }
        case DeviceKernel::PENNES: {
            scafes_pennes_params p = scafes_pennes_params();
            p.tau = mParams.tau();
            p.cdiff = mDev.cdiff;
            p.cperf = mDev.cperf;
            p.csrc = mDev.csrc;
            p.t_a = mDev.tA - shift();
            p.border_mode = (int)mDev.border;
            p.k_cond = mDev.kCond;
            p.h_conv = mDev.hConv;
            p.t_inf = mDev.tInf - shift();
            p.q_bc = mDev.qBc;
            deviceCheck(scafes_pennes_step(c, &g, hw, f32() ? SCAFES_F32 : SCAFES_F64, mUOld.get(), mUNew.get(),
                                           mKnownA.get(), mKnownB.get(), &p),
                        "pennes step");
            break;
        }
        case DeviceKernel::MRI: {
            const double* mat[SCAFES_MRI_NFIELDS];
            for (int q = 0; q < SCAFES_MRI_NFIELDS; ++q) mat[q] = (const double*)mMat[q].get();
            const scafes_mri_params p = mriParams();
            deviceCheck(scafes_mri_step_f64(c, &g, hw, (const double*)mUOld.get(), (double*)mUNew.get(), mat,
                                            (const double*)mKnownA.get(), &p),
                        "mri step");
            break;
        }
        case DeviceKernel::EMFDTD: {
            // updateInner (H) and updateInner2 (D, E, I, source, sink); the
            // source reads the old Ez's time (EMFDTD.hpp:370-371)
            const double tOld = mVectUnknownDfsDomOld[unknownSlot(mDev.em[3])].time();
            deviceCheck(scafes_emfdtd_run(c, &mEmP, (double*)mEm.get(), 1, tOld, (double*)mEmSink.get() + mEmSteps),
                        "emfdtd step");
            mEmP.flags |= SCAFES_EMFDTD_PSI_SETTLED;
            ++mEmSteps;
            break;
        }
        case DeviceKernel::HOOKS: {
            hooksPhase(n, 1);
            if (mUseLeapfrog) {
                // reference :2603-2619: swap old/new known and unknown fields
                // but not their times, then the update2 phase
                swapKeepingTimes(mVectKnownDfsDomNew, mVectKnownDfsDomOld);
                std::swap(mHookKnownNew, mHookKnownOld);
                swapKeepingTimes(mVectUnknownDfsDomNew, mVectUnknownDfsDomOld);
                std::swap(mHookNew, mHookOld);
                for (std::size_t u = 0; u < nUnknown(); ++u)
                    mVectUnknownDfsDomNew[u].time() = mVectUnknownDfsDomOld[u].time() + mParams.tau();
                hooksPhase(n, 2);
            }
            break;
        }
        case DeviceKernel::YEE: {
            // updateInner (H) and, after the leapfrog swap, updateInner2 (E):
            // one fused leapfrog step on the device (reference :2603-2619)
            double e = 0.0;
            const bool rec = mDev.energyEvery > 0 && (n % mDev.energyEvery) == 0;
            int second = 0;
            deviceCheck(scafes_yee_run(c, nNodes(0) - 1, f32() ? SCAFES_F32 : SCAFES_F64, mDev.courant, mYee.get(),
                                       mYee2.get(), 1, rec ? 1 : 0, rec ? &e : 0, &second),
                        "yee step");
            if (second) std::swap(mYee, mYee2);
            if (rec) mEnergies.push_back(e);
            // the reference swaps unknown pointers and times twice inside a
            // leapfrog step; the net effect on the time slots is nil
            break;
        }
        default:
            break;
        }
        mCtx->sync();
        mWallClockTimes["update"] += t.elapsed();
    }

    // ---------------------------------------------------- DeviceTraits check
    // After the first step: the declared kernel family's new state of the
    // updated field against the problem's own hooks run on the host (TT = CT)
    // from the same old state, at sampled nodes (every combination of the
    // per-axis indices 0, 1, 2, n/4, n/2, 3n/4, n-3, n-2, n-1, and about
    // 2048 nodes spread over the grid).  A DeviceTraits that names the wrong
    // family, coefficient, field or border rule throws here instead of
    // silently running another update.  Tolerance relative to the largest
    // sampled value: 1e-12 for FP64 kernels (only the operation order may
    // differ), 1e-5 for the FP32 ones.  Not for leapfrog problems (the
    // second phase needs the first one's neighbourhood) or device-evaluated
    // known fields (their host copies are not current).
    void checkTraitsAgainstHooks(const int& n)
    {
        if (!mParams.checkDeviceTraits()) return;
        if (mDev.kernel == DeviceKernel::HOOKS || mUseLeapfrog || mDev.knownGen != DeviceKnownGen::NONE) return;
        const int u = unknownSlot(mDev.unknown);
        if (u < 0) return;
        downloadUnknownTo((std::size_t)u, mVectUnknownDfsDomNew[u]);
        // the hooks write into copies of the old lumps (new <- old first, as
        // the reference's update phase does)
        std::vector<std::vector<CT>> scratch(nUnknown());
        ScaFES_VectDf<CT> vNew(mVectUnknownDfsDomNew);
        for (std::size_t q = 0; q < nUnknown(); ++q) {
            const DataField<CT, DIM>& o = mVectUnknownDfsDomOld[q];
            scratch[q].assign(o.elemData(), o.elemData() + o.nElemsAllocated());
            vNew[q].setElemData(scratch[q].data());
        }
        std::vector<ScaFES_IntNtuple> nodes;
        int axis[DIM][9];
        int nax[DIM];
        for (std::size_t p = 0; p < DIM; ++p) {
            const int m = nNodes((int)p), f = mDomainC.idxNodeFirst((int)p);
            const int cand[9] = {0, 1, 2, m / 4, m / 2, (3 * m) / 4, m - 3, m - 2, m - 1};
            nax[p] = 0;
            for (int c = 0; c < 9; ++c) {
                if (cand[c] < 0 || cand[c] >= m) continue;
                bool dup = false;
                for (int e = 0; e < nax[p]; ++e) dup = dup || axis[p][e] == f + cand[c];
                if (!dup) axis[p][nax[p]++] = f + cand[c];
            }
        }
        int pick[DIM] = {};
        for (;;) {
            ScaFES_IntNtuple x;
            for (std::size_t p = 0; p < DIM; ++p) x[p] = axis[p][pick[p]];
            nodes.push_back(x);
            std::size_t p = 0;
            while (p < DIM && ++pick[p] == nax[p]) pick[p++] = 0;
            if (p == DIM) break;
        }
        const unsigned long total = (unsigned long)nNodes().size();
        const unsigned long stride = total > 2048 ? total / 2048 + 1 : 1;
        for (unsigned long t = 0; t < total; t += stride) nodes.push_back(mDomainC.idxNodeScalar2TupleStyleFortran(t));
        OWNPRBLM* self = static_cast<OWNPRBLM*>(this);
        const DataField<CT, DIM>& dev = mVectUnknownDfsDomNew[u];
        double scale = 0.0, worst = -1.0, hw = 0.0, dw = 0.0;
        ScaFES_IntNtuple at;
        for (std::size_t k = 0; k < nodes.size(); ++k) {
            const ScaFES_IntNtuple& x = nodes[k];
            if (mKind[x] & GridGlobal<DIM>::SCAFES_GLOBAL_INNER_NODE)
                self->template updateInner<CT>(vNew, mVectUnknownDfsDomOld, x, n);
            else
                self->template updateBorder<CT>(vNew, mVectUnknownDfsDomOld, x, n);
            const double h = (double)vNew[u](x), d = (double)dev(x);
            scale = std::max(scale, std::fabs(h));
            const double e = std::fabs(h - d);
            if (!(e <= worst)) {  // NaN counts as the worst
                worst = e;
                hw = h;
                dw = d;
                at = x;
            }
        }
        const double tol = (f32() ? 1e-5 : 1e-12) * std::max(scale, 1e-300);
        if (!(worst <= tol)) {
            std::ostringstream msg;
            msg << std::setprecision(17) << "ScaFES: the DeviceTraits kernel family disagrees with the problem's own "
                << "update hooks after the first step at node " << at << ": device " << dw << ", hooks " << hw
                << " (tolerance " << tol << "); fix the DeviceTraits specialisation, or pass "
                << "--checkDeviceTraits=0 to skip this check.";
            throw std::runtime_error(msg.str());
        }
    }

    // ---------------------------------------------------- generic device update
    // The problem's own updateInner/updateBorder(/2), instantiated with
    // TT = DevReal<CT> and run by one GPU thread per node
    // (ScaFES_DeviceHooks.hpp).  Device state: one memAll array per unknown
    // field (old and new), per known lump (known fields and the reference
    // solutions of compErr fields; old and new), the node kinds, the views
    // of one launch (HookState) and a byte copy of this problem object.
#if defined(__HIPCC__)
    SCAFES_HD const HookState<CT, DIM>* hookState() const
    {
        return static_cast<const HookState<CT, DIM>*>(mDevHookState);
    }
    void hooksSetup()
    {
        const std::size_t es = sizeof(CT);
        mHookOld.clear();
        mHookNew.clear();
        for (std::size_t u = 0; u < nUnknown(); ++u) {
            const std::size_t n = nodesOf(mUnknownIndex[u]);
            mHookOld.push_back(DeviceBuffer(*mCtx, n * es));
            mHookNew.push_back(DeviceBuffer(*mCtx, n * es));
            mHookOld.back().upload(mVectUnknownDfsDomOld[u].elemData(), n * es);
            mHookNew.back().upload(mVectUnknownDfsDomNew[u].elemData(), n * es);
        }
        mHookKnownOld.clear();
        mHookKnownNew.clear();
        for (std::size_t ii = 0; ii < mNames.size(); ++ii) {
            mHookKnownOld.push_back(DeviceBuffer());
            mHookKnownNew.push_back(DeviceBuffer());
            if (!mVectKnownDfsDomNew[ii].elemData()) continue;
            const std::size_t n = nodesOf((int)ii);
            mHookKnownOld.back() = DeviceBuffer(*mCtx, n * es);
            mHookKnownNew.back() = DeviceBuffer(*mCtx, n * es);
            mHookKnownOld.back().upload(mVectKnownDfsDomOld[ii].elemData(), n * es);
            mHookKnownNew.back().upload(mVectKnownDfsDomNew[ii].elemData(), n * es);
        }
        mHookKind = DeviceBuffer(*mCtx, mLumpKind.size() * sizeof(short));
        mHookKind.upload(mLumpKind.data(), mLumpKind.size() * sizeof(short));
        mHookStateBuf = DeviceBuffer(*mCtx, sizeof(HookState<CT, DIM>));
        mDevHookState = mHookStateBuf.get();
        // the device copy of this object: its hook accessors read plain
        // members and, on the device, the HookState above
        mHookSelf = DeviceBuffer(*mCtx, sizeof(OWNPRBLM));
        mHookSelf.upload(static_cast<OWNPRBLM*>(this), sizeof(OWNPRBLM));
    }
    // The views of the current old/new iterates (pointers, layouts, times).
    void hooksUploadState()
    {
        typedef DataField<DevReal<CT>, DIM> View;
        std::unique_ptr<HookState<CT, DIM>> st(new HookState<CT, DIM>());
        st->gg = mGG;
        st->domain = mDomainC;
        st->nNodes = (long)nNodes().size();
        st->kind = static_cast<const short*>(mHookKind.get());
        const GridGlobal<DIM>* ggDev = reinterpret_cast<const GridGlobal<DIM>*>(
            static_cast<char*>(mHookStateBuf.get()) + (reinterpret_cast<char*>(&st->gg) - reinterpret_cast<char*>(st.get())));
        for (std::size_t u = 0; u < nUnknown(); ++u) {
            const int ii = mUnknownIndex[u];
            const long n = (long)nodesOf(ii);
            st->vNew.push_back(View(static_cast<CT*>(mHookNew[u].get()), mMemAll[ii], mStencilWidth[ii], 1, n,
                                    mVectUnknownDfsDomNew[u].time(), ggDev));
            st->vOld.push_back(View(static_cast<CT*>(mHookOld[u].get()), mMemAll[ii], mStencilWidth[ii], 1, n,
                                    mVectUnknownDfsDomOld[u].time(), ggDev));
        }
        for (std::size_t ii = 0; ii < mNames.size(); ++ii) {
            if (!mHookKnownNew[ii].get()) continue;
            const int sw = mIsKnownDf[ii] ? mStencilWidth[ii] : 0;
            const long n = (long)nodesOf((int)ii);
            st->knownNew[ii] = View(static_cast<CT*>(mHookKnownNew[ii].get()), mMemAll[ii], sw, 1, n,
                                    mVectKnownDfsDomNew[ii].time(), ggDev);
            st->knownOld[ii] = View(static_cast<CT*>(mHookKnownOld[ii].get()), mMemAll[ii], sw, 1, n,
                                    mVectKnownDfsDomOld[ii].time(), ggDev);
        }
        mHookStateBuf.upload(st.get(), sizeof(HookState<CT, DIM>));
    }
    // One update phase (1: updateInner/updateBorder, 2: the leapfrog's
    // updateInner2/updateBorder2): new <- old for every unknown field
    // (setVectDfDepToDfIndep, reference :6226-6256), then the hooks.
    void hooksPhase(int n, int phase)
    {
        scafes_ctx* c = mCtx->get();
        for (std::size_t u = 0; u < nUnknown(); ++u)
            deviceCheck(scafes_memcpy_d2d(c, mHookNew[u].get(), mHookOld[u].get(), mHookOld[u].bytes()),
                        "old to new copy");
        hooksUploadState();
        if (launch_hooks<OWNPRBLM, CT, DIM>(mCtx->stream(), static_cast<OWNPRBLM*>(mHookSelf.get()),
                                            static_cast<HookState<CT, DIM>*>(mHookStateBuf.get()),
                                            (long)nNodes().size(), n, phase))
            throw std::runtime_error("ScaFES device: launch of the problem's update hooks failed");
    }
#else
    void hooksSetup() { throw std::logic_error("device hooks need a hipcc-compiled program"); }
    void hooksPhase(int, int) { throw std::logic_error("device hooks need a hipcc-compiled program"); }
#endif

    // ---------------------------------------------------- error / convergence
    void compErrOfUnknownDfs(const int& /*n*/)
    {
        Timer t;
        for (std::size_t u = 0; u < nUnknown(); ++u) {
            const int ii = mUnknownIndex[u];
            if (!mComputeError[ii]) continue;
            double e = 0.0;
            const scafes_geom g = geom();
            if (mDev.kernel == DeviceKernel::HOOKS) {
                deviceCheck(scafes_reduce(mCtx->get(), &g, sizeof(CT) == 8 ? SCAFES_F64 : SCAFES_F32,
                                          mHookNew[u].get(), mStencilWidth[ii], mHookKnownNew[ii].get(),
                                          mStencilWidth[ii], 0, 0, &e),
                            "compErrLinf");
            } else if (!f32() && mDev.kernel != DeviceKernel::YEE && ii == mDev.unknown) {
                deviceCheck(scafes_reduce(mCtx->get(), &g, SCAFES_F64, mUNew.get(), mStencilWidth[ii], mRef[u].get(),
                                          mStencilWidth[ii], 0, 0, &e),
                            "compErrLinf");
            } else {
                downloadUnknownTo(u, mVectUnknownDfsDomNew[u]);
                e = mVectUnknownDfsDomNew[u].compErrLinf(mVectKnownDfsDomNew[ii]);
            }
            mErrors.push_back(e);
            if (out())
                std::cout << mParams.getPrefix() << " * Local Linf error(" << mVectUnknownDfsDomNew[u].name() << ", "
                          << mVectKnownDfsDomNew[ii].name() << "): " << std::scientific << e
                          << std::resetiosflags(std::ios::scientific) << std::endl;
        }
        mWallClockTimes["compErr"] += t.elapsed();
    }

    bool checkConvOfUnknownDfs(const int& n)
    {
        Timer t;
        std::vector<bool> cont;
        for (std::size_t u = 0; u < nUnknown(); ++u) {
            const int ii = mUnknownIndex[u];
            if (!(mCheckConvergence[ii] && n >= mParams.checkConvFirstAtIter() &&
                  (n - mParams.checkConvFirstAtIter()) % mParams.checkConvAtEveryNIter() == 0))
                continue;
            double d = 0.0;
            const scafes_geom g = geom();
            if (mDev.kernel == DeviceKernel::HOOKS) {
                deviceCheck(scafes_reduce(mCtx->get(), &g, sizeof(CT) == 8 ? SCAFES_F64 : SCAFES_F32, mHookNew[u].get(),
                                          mStencilWidth[ii], mHookOld[u].get(), mStencilWidth[ii], 0, 0, &d),
                            "checkConv");
            } else {
                if (mDev.kernel == DeviceKernel::YEE || ii != mDev.unknown)
                    throw std::runtime_error(
                        "checkConvergence is supported for the updated field of a heat/Pennes kernel");
                deviceCheck(scafes_reduce(mCtx->get(), &g, f32() ? SCAFES_F32 : SCAFES_F64, mUNew.get(),
                                          mStencilWidth[ii], mUOld.get(), mStencilWidth[ii], 0, 0, &d),
                            "checkConv");
            }
            cont.push_back(d > mParams.threshold());
            if (out())
                std::cout << mParams.getPrefix() << " * Is (" << mVectUnknownDfsDomNew[u].name() << ", "
                          << mVectUnknownDfsDomOld[u].name() << ") converged?: " << !cont.back() << std::endl;
        }
        mWallClockTimes["checkConv"] += t.elapsed();
        if (cont.empty()) return true;
        for (std::size_t i = 0; i < cont.size(); ++i)
            if (cont[i]) return true;
        return false;
    }

    // ---------------------------------------------------- swaps and times
    void swapPointersOfKnownDfs()
    {
        for (std::size_t ii = 0; ii < mNames.size(); ++ii)
            mVectKnownDfsDomNew[ii].swapPointerToMemoryWith(mVectKnownDfsDomOld[ii]);
        std::swap(mHookKnownNew, mHookKnownOld);
    }
    void swapPointersOfUnknownDfs()
    {
        for (std::size_t u = 0; u < nUnknown(); ++u)
            mVectUnknownDfsDomNew[u].swapPointerToMemoryWith(mVectUnknownDfsDomOld[u]);
        mUOld.swap(mUNew);
        mCOld.swap(mCNew);
        std::swap(mHookNew, mHookOld);
    }
    // Pointer swap that leaves each view's time where it was (the
    // reference's swapPointers + swapTimes pair, :2604-2616).
    static void swapKeepingTimes(ScaFES_VectDf<CT>& a, ScaFES_VectDf<CT>& b)
    {
        for (std::size_t k = 0; k < a.size(); ++k) {
            if (!a[k].elemData() || !b[k].elemData()) continue;
            const CT ta = a[k].time(), tb = b[k].time();
            a[k].swapPointerToMemoryWith(b[k]);
            a[k].time() = ta;
            b[k].time() = tb;
        }
    }
    void updateTimesOfKnownDfs()
    {
        if (!evalDf()) return;
        for (std::size_t ii = 0; ii < mNames.size(); ++ii)
            if (mVectKnownDfsDomNew[ii].elemData())
                mVectKnownDfsDomNew[ii].time() = mVectKnownDfsDomOld[ii].time() + mParams.tau();
    }
    void updateTimesOfUnknownDfs()
    {
        for (std::size_t u = 0; u < nUnknown(); ++u)
            mVectUnknownDfsDomNew[u].time() = mVectUnknownDfsDomOld[u].time() + mParams.tau();
    }

    // ---------------------------------------------------- download
    // Copies device state of unknown slot u into the lump the view df points to.
    void downloadUnknownTo(std::size_t u, DataField<CT, DIM>& df)
    {
        const int ii = mUnknownIndex[u];
        std::vector<CT>& lump = df.elemData() == mLumpUnknownOld[u].data() ? mLumpUnknownOld[u] : mLumpUnknownNew[u];
        const bool isNew = (&df == &mVectUnknownDfsDomNew[u]);
        if (mDev.kernel == DeviceKernel::HOOKS) {
            (isNew ? mHookNew[u] : mHookOld[u]).download(lump.data(), nodesOf(ii) * sizeof(CT));
            return;
        }
        if (mDev.kernel == DeviceKernel::EMFDTD) {
            const std::size_t nn = (std::size_t)nNodes().size();
            int q = -1;
            for (int r = 1; r < 25; ++r)
                if (mDev.em[r] == ii) q = r;
            if (q < 0) return;  // SIGMA: never updated, both lumps hold it
            std::vector<double> tmp(nn);
            deviceCheck(scafes_memcpy_d2h(mCtx->get(), tmp.data(), (const double*)mEm.get() + (q - 1) * nn,
                                          nn * sizeof(double)),
                        "download");
            unpackNodes(lump, ii, tmp.data());
            return;
        }
        if (mDev.kernel == DeviceKernel::YEE) {
            const std::size_t nn = (std::size_t)nNodes().size();
            int c = -1;
            for (int q = 0; q < 6; ++q)
                if (mDev.yee[q] == ii) c = q;
            if (c < 0) return;
            if (f32()) {
                std::vector<float> tmp(nn);
                deviceCheck(scafes_memcpy_d2h(mCtx->get(), tmp.data(), (const float*)mYee.get() + c * nn,
                                              nn * sizeof(float)),
                            "download");
                unpackNodes(lump, ii, tmp.data());
            } else {
                std::vector<double> tmp(nn);
                deviceCheck(scafes_memcpy_d2h(mCtx->get(), tmp.data(), (const double*)mYee.get() + c * nn,
                                              nn * sizeof(double)),
                            "download");
                unpackNodes(lump, ii, tmp.data());
            }
            return;
        }
        if (ii != mDev.unknown) return;
        const std::size_t n = nodesOf(ii);
        const DeviceBuffer& d = isNew ? mUNew : mUOld;
        if (!f32()) {
            std::vector<double> tmp(n);
            d.download(tmp.data(), n * sizeof(double));
            for (std::size_t i = 0; i < n; ++i) lump[i] = (CT)tmp[i];
        } else {
            std::vector<float> tmp(n), cmp(n, 0.0f);
            d.download(tmp.data(), n * sizeof(float));
            if (mDev.precision == DevicePrecision::F32_KAHAN)
                (isNew ? mCNew : mCOld).download(cmp.data(), n * sizeof(float));
            for (std::size_t i = 0; i < n; ++i) lump[i] = (CT)((double)tmp[i] + (double)cmp[i] + shift());
        }
    }
    // After the loop the latest iterate is behind the "old" views (swapped).
    void deviceDownload()
    {
        if (!mCtx) return;
        mCtx->sync();
        if (mDev.kernel == DeviceKernel::EMFDTD && mEmSteps > 0) {
            mSink.assign((std::size_t)mEmSteps, 0.0);
            deviceCheck(scafes_memcpy_d2h(mCtx->get(), mSink.data(), mEmSink.get(), mSink.size() * sizeof(double)),
                        "sink download");
        }
        for (std::size_t u = 0; u < nUnknown(); ++u) downloadUnknownTo(u, mVectUnknownDfsDomOld[u]);
    }

    void printTimes(double tIterate)
    {
        double tSum = 0.0;
        for (std::map<std::string, double>::const_iterator it = mWallClockTimes.begin(); it != mWallClockTimes.end(); ++it)
            tSum += it->second;
        std::cout << mParams.getPrefix() << " * Summary of Wall clock times(iterateOverTime):" << std::endl;
        for (std::map<std::string, double>::const_iterator it = mWallClockTimes.begin(); it != mWallClockTimes.end(); ++it)
            std::cout << mParams.getPrefix() << "       t(" << std::setw(11) << it->first << ") = " << std::fixed
                      << std::setw(8) << std::setprecision(5) << it->second << " s (" << std::setw(5)
                      << std::setprecision(2) << (tSum > 0 ? it->second / tSum * 100.0 : 0.0) << " %)" << std::endl;
        std::cout << mParams.getPrefix() << "       t(    iterate) = " << std::setw(8) << std::setprecision(5) << tSum
                  << " s." << std::endl
                  << mParams.getPrefix() << " t(iterate) w.barriers= " << std::setw(8) << std::setprecision(5)
                  << tIterate << " s." << std::resetiosflags(std::ios::fixed) << std::setprecision(6) << std::endl;
    }

protected:
    Parameters mParams;
    const GridGlobal<DIM>& mGG;

private:
    bool mUseLeapfrog;
    std::vector<std::string> mNames;
    std::vector<int> mStencilWidth;
    std::vector<bool> mIsKnownDf;
    std::vector<int> mNlayers;
    std::vector<CT> mDefaultValue;
    std::vector<WriteHowOften> mWriteToFile;
    std::vector<bool> mComputeError;
    std::vector<CT> mParamsCurr;
    std::vector<bool> mCheckConvergence;
    std::vector<Grid<DIM>> mMemAll;
    std::vector<int> mUnknownIndex;
    std::vector<std::vector<CT>> mLumpUnknownOld, mLumpUnknownNew, mLumpKnownOld, mLumpKnownNew;
    std::vector<short> mLumpKind;
    ScaFES_VectDf<CT> mVectUnknownDfsDomOld, mVectUnknownDfsDomNew, mVectKnownDfsDomOld, mVectKnownDfsDomNew;
    DataField<short, DIM> mKind;
    std::map<std::string, double> mWallClockTimes;
    std::vector<double> mErrors, mEnergies, mSink;
    int mLastIter = 0;
    // device state
    DeviceUpdate mDev;
    std::unique_ptr<DeviceContext> mCtx;
    DeviceBuffer mUOld, mUNew, mCOld, mCNew, mKnownA, mKnownB, mKnownC, mYee, mYee2, mEm, mEmSink;
    scafes_emfdtd_params mEmP;
    int mEmSteps = 0;
    std::vector<DeviceBuffer> mRef, mMat;
    // generic device update (HOOKS)
    std::vector<DeviceBuffer> mHookOld, mHookNew, mHookKnownOld, mHookKnownNew;
    DeviceBuffer mHookKind, mHookStateBuf, mHookSelf;
    const void* mDevHookState = 0;  // device address of the HookState the device copy reads
    // plain copies for the hook accessors (readable from the device copy)
    Grid<DIM> mDomainC;
    double mTauC = 0.0, mTimeStartC = 0.0;
    DataFile<CT, DIM> mCommonFile, mInitFile;
};

}  // namespace ScaFES
#endif
