// ScaFES_Parameters.hpp -- command-line parameters of the host API.
//
// Accepts the reference's option names (src/ScaFES_Parameters.hpp:473-540)
// in the form --name=value or --name value, with the reference's defaults
// (:424-463) and time-step rule (:769-812): --tau if given, else
// (endtime - starttime)/nTimesteps if both are given, else tau = 1.
// Bad values abort with exit(-1) after a message, as Environment::abort
// does without MPI (src/ScaFES_Environment.hpp:167-179).
// Options for partitioning and ADOL-C that this single-GPU engine does not
// implement are parsed; asking for ADOL-C or more than one partition is an
// error.  --device selects the HIP device (new).
#ifndef SCAFES_HOST_PARAMETERS_HPP_
#define SCAFES_HOST_PARAMETERS_HPP_

#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

namespace ScaFES {

class Parameters {
public:
    Parameters(int argc, char* argv[])
        : mDim(3), mNnodes(3, 32), mCoordNodeFirst(3, 0.0), mCoordNodeLast(3, 1.0),
          mNpartitions(3, 1), mTimeIntervalStart(0.0), mTimeIntervalEnd(1.0), mNtimesteps(1),
          mTau(1.0), mThreshold(0.00001), mCheckConvFirstAtIter(1), mCheckConvAtEveryNIter(1),
          mNsnapshots(1), mNlayersAtBorder(1), mEnabledAdolc(false), mAsynchronMode(true),
          mComputeGradients(false), mRankOutput(0), mOutputLevelMax(2), mOutputLevelCurr(2),
          mDevice(0)
    {
        bool haveTau = false, haveEnd = false, haveSteps = false;
        for (int a = 1; a < argc; ++a) {
            std::string arg(argv[a]);
            if (arg.size() < 3 || arg[0] != '-' || arg[1] != '-') abortWith("unknown argument " + arg);
            std::string key = arg.substr(2), val;
            const std::size_t eq = key.find('=');
            if (eq != std::string::npos) {
                val = key.substr(eq + 1);
                key = key.substr(0, eq);
            } else if (a + 1 < argc && std::string(argv[a + 1]).compare(0, 2, "--") != 0) {
                val = argv[++a];
            } else {
                val = "1";
            }
            if (key == "dim") mDim = toInt(key, val);
            else if (key == "nNodes") mNnodes = toIntList(key, val);
            else if (key == "coordNodeFirst") mCoordNodeFirst = toDoubleList(key, val);
            else if (key == "coordNodeLast") mCoordNodeLast = toDoubleList(key, val);
            else if (key == "nPartitions" || key == "divideGrid") mNpartitions = toIntList(key, val);
            else if (key == "starttime") mTimeIntervalStart = toDouble(key, val);
            else if (key == "endtime") { mTimeIntervalEnd = toDouble(key, val); haveEnd = true; }
            else if (key == "nTimesteps") { mNtimesteps = toInt(key, val); haveSteps = true; }
            else if (key == "tau") { mTau = toDouble(key, val); haveTau = true; }
            else if (key == "threshold") mThreshold = toDouble(key, val);
            else if (key == "checkConvFirstAtIter") mCheckConvFirstAtIter = toInt(key, val);
            else if (key == "checkConvAtEveryNIter") mCheckConvAtEveryNIter = toInt(key, val);
            else if (key == "nSnapshots") mNsnapshots = toInt(key, val);
            else if (key == "nLayersAtBorder") mNlayersAtBorder = toInt(key, val);
            else if (key == "outputlevel") mOutputLevelMax = mOutputLevelCurr = toInt(key, val);
            else if (key == "rankOutput") mRankOutput = toInt(key, val);
            else if (key == "enabledAdolc") mEnabledAdolc = toBool(val);
            else if (key == "asynchronMode") mAsynchronMode = toBool(val);
            else if (key == "computeGradients") mComputeGradients = toBool(val);
            else if (key == "device") mDevice = toInt(key, val);
            else if (key == "checkDeviceTraits") mCheckDeviceTraits = toBool(val);
            else if (key == "configfile") mNameConfigFile = val;
            else if (key == "outputfile") { mNameDataFile = val; mWriteDataFile = true; }
            else if (key == "initfile") { mNameInitFile = val; mReadInitFile = true; }
            else if (key == "initKnownDfs") mInitKnownDfs = toInt(key, val) == 1;
            else if (key == "typeDomainDecomposition" || key == "useBoostMpiSkeletonConcept" ||
                     key == "checkGradients" || key == "kindfile" || key == "writeKindFile") {
                // accepted for command-line compatibility; no effect on one GPU
            } else {
                abortWith("unknown option --" + key);
            }
        }
        if (mDim < 1 || mDim > 3) abortWith("--dim must be 1, 2 or 3");
        resizeTo(mNnodes, 32);
        resizeTo(mCoordNodeFirst, 0.0);
        resizeTo(mCoordNodeLast, 1.0);
        resizeTo(mNpartitions, 1);
        for (int p = 0; p < mDim; ++p) {
            if (mNnodes[p] < 2) abortWith("--nNodes must be >= 2 along every axis");
            if (mNpartitions[p] != 1) abortWith("only one partition is supported (one GPU)");
        }
        if (mEnabledAdolc) abortWith("--enabledAdolc: ADOL-C taping is replaced by the adjoint kernels");
        if (mNtimesteps < 0) abortWith("--nTimesteps must be >= 0");
        if (haveTau) {
            if (0.0 >= mTau) abortWith("Time step width tau <= 0.0.");
        } else if (haveEnd && haveSteps) {
            mTau = (0 == mNtimesteps) ? 0.0 : (mTimeIntervalEnd - mTimeIntervalStart) / mNtimesteps;
        } else {
            mTau = 1.0;
        }
        if (mNlayersAtBorder < 1) abortWith("--nLayersAtBorder must be >= 1");
        if (mCheckConvAtEveryNIter < 1) abortWith("--checkConvAtEveryNIter must be >= 1");
    }

    int dim() const { return mDim; }
    const std::vector<int>& nNodes() const { return mNnodes; }
    const std::vector<double>& coordNodeFirst() const { return mCoordNodeFirst; }
    const std::vector<double>& coordNodeLast() const { return mCoordNodeLast; }
    const std::vector<int>& nPartitions() const { return mNpartitions; }
    const double& timeIntervalStart() const { return mTimeIntervalStart; }
    const double& timeIntervalEnd() const { return mTimeIntervalEnd; }
    const int& nTimesteps() const { return mNtimesteps; }
    const double& tau() const { return mTau; }
    const double& threshold() const { return mThreshold; }
    const int& checkConvFirstAtIter() const { return mCheckConvFirstAtIter; }
    const int& checkConvAtEveryNIter() const { return mCheckConvAtEveryNIter; }
    const int& nSnapshots() const { return mNsnapshots; }
    const int& nLayersAtBorder() const { return mNlayersAtBorder; }
    bool enabledAdolc() const { return mEnabledAdolc; }
    bool asynchronMode() const { return mAsynchronMode; }
    bool computeGradients() const { return mComputeGradients; }
    const int& rankOutput() const { return mRankOutput; }
    int rank() const { return 0; }
    int nProcesses() const { return 1; }
    int device() const { return mDevice; }
    // --checkDeviceTraits (default on): after the first time step, compare the
    // declared kernel family's result with the problem's own hooks evaluated
    // on the host at sampled nodes (no counterpart in the reference, which
    // runs only the hooks).
    bool checkDeviceTraits() const { return mCheckDeviceTraits; }
    // --configfile (Parameters.hpp:516, 928-935): read by problem classes.
    const std::string& nameConfigFile() const { return mNameConfigFile; }
    // --outputfile / --initfile / --initKnownDfs (Parameters.hpp:514-520,
    // 917-963): snapshot file <outputfile>.nc (default EMPTYDATAFILE), initial
    // state from <initfile>.nc instead of the init hooks, known fields too.
    const std::string& nameDataFile() const { return mNameDataFile; }
    const bool& writeDataFile() const { return mWriteDataFile; }
    const std::string& nameInitFile() const { return mNameInitFile; }
    const bool& readInitFile() const { return mReadInitFile; }
    const bool& initKnownDfs() const { return mInitKnownDfs; }

    // Indented, level-gated output (Parameters.hpp:1726-1737).
    int indentDepth() const { return mOutputLevelCurr; }
    void decreaseLevel() { --mOutputLevelCurr; }
    void increaseLevel() { ++mOutputLevelCurr; }
    std::string getPrefix() const
    {
        const int depth = mOutputLevelMax - mOutputLevelCurr;
        return std::string("[ScaFES]") + std::string(depth > 0 ? 2 * depth : 0, ' ');
    }

private:
    static void abortWith(const std::string& msg)
    {
        std::cerr << "\nERROR: " << msg << " Exit...\n" << std::endl;
        std::exit(-1);
    }
    static int toInt(const std::string& k, const std::string& v)
    {
        std::istringstream is(v);
        int x = 0;
        if (!(is >> x)) abortWith("--" + k + " expects an integer, got " + v);
        return x;
    }
    static double toDouble(const std::string& k, const std::string& v)
    {
        std::istringstream is(v);
        double x = 0;
        if (!(is >> x)) abortWith("--" + k + " expects a real value, got " + v);
        return x;
    }
    static bool toBool(const std::string& v)
    {
        return v == "1" || v == "true" || v == "TRUE" || v == "yes" || v == "YES" || v == "on";
    }
    static std::vector<std::string> split(const std::string& v)
    {
        std::vector<std::string> out;
        std::string cur;
        for (std::size_t i = 0; i < v.size(); ++i) {
            if (v[i] == 'x' || v[i] == 'X') {
                out.push_back(cur);
                cur.clear();
            } else {
                cur += v[i];
            }
        }
        out.push_back(cur);
        return out;
    }
    static std::vector<int> toIntList(const std::string& k, const std::string& v)
    {
        std::vector<int> r;
        std::vector<std::string> parts = split(v);
        for (std::size_t i = 0; i < parts.size(); ++i) r.push_back(toInt(k, parts[i]));
        return r;
    }
    static std::vector<double> toDoubleList(const std::string& k, const std::string& v)
    {
        std::vector<double> r;
        std::vector<std::string> parts = split(v);
        for (std::size_t i = 0; i < parts.size(); ++i) r.push_back(toDouble(k, parts[i]));
        return r;
    }
    template <typename T>
    void resizeTo(std::vector<T>& v, T fill)
    {
        if ((int)v.size() > mDim) v.resize(mDim);
        while ((int)v.size() < mDim) v.push_back(v.empty() ? fill : v.back());
    }

    int mDim;
    std::vector<int> mNnodes;
    std::vector<double> mCoordNodeFirst, mCoordNodeLast;
    std::vector<int> mNpartitions;
    double mTimeIntervalStart, mTimeIntervalEnd;
    int mNtimesteps;
    double mTau, mThreshold;
    int mCheckConvFirstAtIter, mCheckConvAtEveryNIter, mNsnapshots, mNlayersAtBorder;
    bool mEnabledAdolc, mAsynchronMode, mComputeGradients;
    int mRankOutput, mOutputLevelMax, mOutputLevelCurr, mDevice;
    bool mCheckDeviceTraits = true;
    std::string mNameConfigFile;
    std::string mNameDataFile = "EMPTYDATAFILE", mNameInitFile = "EMPTYINITFILE";
    bool mWriteDataFile = false, mReadInitFile = false, mInitKnownDfs = false;
};

}  // namespace ScaFES
#endif
