// ScaFES_Device.hpp -- C++ side of the C ABI (include/scafes_hip.h).
//
// RAII wrappers over the HIP library and the kernel descriptor a problem
// class provides so that the update phase runs on the GPU.  The reference
// calls the update hooks per node on the CPU (src/ScaFES_Problem.hpp:5443-5466);
// arbitrary C++ hooks cannot run on a GPU, so a problem names its update
// family and coefficients once, by specialising ScaFES::DeviceTraits for the
// class it passes to Problem as OWNPRBLM, next to the (unchanged) class:
//
//   template <typename CT, std::size_t DIM, typename Class>
//   struct ScaFES::DeviceTraits<HeatEqnFDM<CT, DIM, Class>> {
//       static const bool available = true;
//       static void describe(const HeatEqnFDM<CT, DIM, Class>& p, ScaFES::DeviceUpdate& d) {...}
//   };
//
// A problem without a specialisation has no device path, and
// Problem::iterateOverTime() reports that instead of running anything on
// the CPU.
#ifndef SCAFES_HOST_DEVICE_HPP_
#define SCAFES_HOST_DEVICE_HPP_

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "scafes_hip.h"

namespace ScaFES {

enum class DeviceKernel {
    NONE = 0,
    HEAT7,     // HeatEqnFDM / PennesBioheatEqnFDM system form, FP64
    HEAT7VAR,  // conservative variable-kappa heat, FP64 / FP32 / FP32+Kahan
    PENNES,    // TumorHeatEqnFDM form with perfusion and source fields
    YEE,       // leapfrog Yee FDTD, PEC box, vacuum (updateInner = H, updateInner2 = E)
    EMFDTD,    // the EMFDTD example: Yee leapfrog + complete CPML + conductive media, FP64
    MRI,       // the MRIDataPBHEqnFDM example: Pennes with nine per-node material fields, FP64
    HOOKS      // the problem's own update hooks compiled as device code (hipcc; no DeviceTraits)
};
enum class DevicePrecision { F64 = 0, F32 = 1, F32_KAHAN = 2 };
// KEEP: the copied old value; KNOWN_FIELD: Dirichlet from G_new; VALUE:
// Dirichlet from a constant; NEUMANN / CAUCHY / RADIATION: the flux border
// classes of the reference's system tests (tests/system/HeatEqnFDM/Neumann,
// .../Cauchy, tests/system/PennesBioheatEqnFDM/Radiation): G_new on faces
// with an index 0, the one-sided flux update with G_old on the far faces.
enum class DeviceBorder {
    KEEP = SCAFES_BORDER_KEEP,
    KNOWN_FIELD = SCAFES_BORDER_FIELD,
    VALUE = SCAFES_BORDER_VALUE,
    NEUMANN = SCAFES_BORDER_NEUMANN,
    CAUCHY = SCAFES_BORDER_CAUCHY,
    RADIATION = SCAFES_BORDER_RADIATION,
    TUMOR = SCAFES_BORDER_TUMOR  // PENNES: the Tumor example's Neumann / Cauchy borders
};

// Known fields computed on the device instead of by the host hooks: the
// TimeLinSpaceLin system problems' evalInner/evalBorder are closed forms of
// (x, t), so from step 1 on the engine writes F_old, G (and the exact
// solution for compErr) with scafes_known_timelin_f64 and runs no host loop
// and no upload.  The host copies of the known fields keep their n = 0 values.
enum class DeviceKnownGen {
    NONE = 0,
    HEAT_TIMELIN = SCAFES_KNOWN_HEAT_TIMELIN,
    PENNES_TIMELIN = SCAFES_KNOWN_PENNES_TIMELIN
};

// Everything the engine needs to run a problem's update phase on the GPU.
// Known-field indices refer to the problem's data-field list (the order of
// the constructor's nameDatafield vector).
struct DeviceUpdate {
    DeviceKernel kernel = DeviceKernel::NONE;
    DevicePrecision precision = DevicePrecision::F64;
    int unknown = 0;  // data-field index of the updated field
    // HEAT7: u' = u + sum_p tau*a*(W+E-2u)/h_p^2 [- tau*perfC*perfW*u] [+ tau*rhoCInv*F_old]
    double coeffA = 1.0, rhoCInv = 1.0, perfC = 0.0, perfW = 0.0;
    bool perfusion = false;
    int sourceKnown = -1;  // F, read from the old known buffer (knownDfOld)
    DeviceBorder border = DeviceBorder::KEEP;
    int borderKnown = -1;  // G, read from the new known buffer (knownDf); flux borders also read knownDfOld
    double borderValue = 0.0;
    double lambda = 1.0, alpha = 1.0, epsilon = 1.0, sigma = 1.0;  // flux borders (HeatEqnFDM.hpp:98-110)
    // HEAT7VAR: diffusivity field
    int kappaKnown = -1;
    // PENNES: T' = T + sum_p tau*cdiff*(W+E-2T)/h_p^2 + tau*cperf*w*(T_a-T) + tau*csrc*Q
    int perfusionKnown = -1, sourceFieldKnown = -1;
    double cdiff = 1.0, cperf = 1.0, csrc = 1.0, tA = 0.0;
    // PENNES with border TUMOR (TumorHeatEqnFDM.hpp:272-349): K, H, T_INF,
    // Q_BC.  tumorFields: perfusion w and source Q chosen per node by the
    // tumor sphere as the example's hooks do (omegaIn/qIn inside, where the
    // squared distance of the node's coordinates to tumorCenter is <=
    // tumorRadius^2), instead of known data fields.
    double kCond = 1.0, hConv = 0.0, tInf = 0.0, qBc = 0.0;
    bool tumorFields = false;
    double tumorCenter[3] = {0.0, 0.0, 0.0};
    double tumorRadius = 0.0, omegaIn = 0.0, omegaOut = 0.0, qIn = 0.0, qOut = 0.0;
    // YEE: fields Ex,Ey,Ez,Hx,Hy,Hz are data fields yee[0..5]; Courant number
    // tau/h; discrete energy recorded every energyEvery steps (0 = never)
    int yee[6] = {0, 1, 2, 3, 4, 5};
    double courant = 0.5;
    int energyEvery = 0;
    // EMFDTD: data-field indices of SIGMA and the 24 updated fields in the
    // reference's DFs order (EMFDTD.hpp:91-101: SIGMA, Ex..Ez, Hx..Hz, Dx..Dz,
    // Ix..Iz, Psi_ExDy, Psi_ExDz, Psi_EyDx, Psi_EyDz, Psi_EzDx, Psi_EzDy,
    // Psi_HxDy, Psi_HxDz, Psi_HyDx, Psi_HyDz, Psi_HzDx, Psi_HzDy); CPMLParams
    // arguments; Courant factor cfln; source / sink node indices and the
    // ModulatedGaussianSource parameters.  The sink series is recorded on the
    // device and read with Problem::sinkData().
    int em[25] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24};
    int pmlD = 10;
    double pmlMa = 1.0, pmlM = 3.0, pmlKmax = 1.0, pmlSmax = 0.01222, pmlAmax = 1e-3;
    double cfln = 0.5;
    int src[3] = {0, 0, 0}, sink[3] = {0, 0, 0};
    double fMod = 0.0, t0 = 40.0, spread = 8.0;
    // MRI: data-field indices of the material fields rho, c, lambda,
    // rho_blood, c_blood, omega, T_blood, q, surface (main3D.cpp:17-26; known,
    // stencil width 0) and the border constants H (hConv), T_AMB (tInf),
    // Q_BC (qBc), Q_SKULL, EPSILON (epsilon), SIGMA (sigma); see
    // scafes_mri_step_f64 for the update.
    int mri[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    double qSkull = 0.0;
    // Re-evaluate and upload the known fields every step (evalKnownDfs each
    // time iterate, Problem.hpp:3433-3485); false when they do not depend on t.
    bool knownTimeDependent = true;
    // Device evaluation of the known fields (HEAT7 only); F = s*rhoC
    // [+ rhobCbW*s*(1+t)] with s = 1 + sum x, see scafes_known_timelin_f64.
    DeviceKnownGen knownGen = DeviceKnownGen::NONE;
    double rhoC = 1.0, rhobCbW = 0.0;
};

template <class OWNPRBLM>
struct DeviceTraits {
    static const bool available = false;
    static void describe(const OWNPRBLM&, DeviceUpdate&) {}
};

inline void deviceCheck(int rc, const char* what)
{
    if (rc != SCAFES_OK)
        throw std::runtime_error(std::string("ScaFES device: ") + what + ": " + scafes_last_error());
}

class DeviceContext {
public:
    explicit DeviceContext(int device = 0) : mCtx(0)
    {
        deviceCheck(scafes_ctx_create(device, &mCtx), "context creation");
    }
    ~DeviceContext()
    {
        if (mCtx) scafes_ctx_destroy(mCtx);
    }
    DeviceContext(const DeviceContext&) = delete;
    DeviceContext& operator=(const DeviceContext&) = delete;
    scafes_ctx* get() const { return mCtx; }
    void sync() const { deviceCheck(scafes_ctx_sync(mCtx), "sync"); }
    // The context's HIP stream (hipStream_t), for kernels a hipcc-compiled
    // problem launches itself (the generic device update phase).
    void* stream() const
    {
        void* s = 0;
        deviceCheck(scafes_ctx_stream(mCtx, &s), "stream");
        return s;
    }

private:
    scafes_ctx* mCtx;
};

class DeviceBuffer {
public:
    DeviceBuffer() : mCtx(0), mPtr(0), mBytes(0) {}
    DeviceBuffer(const DeviceContext& ctx, std::size_t bytes) : mCtx(ctx.get()), mPtr(0), mBytes(bytes)
    {
        deviceCheck(scafes_malloc(mCtx, bytes, &mPtr), "allocation");
    }
    ~DeviceBuffer() { release(); }
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    DeviceBuffer(DeviceBuffer&& o) noexcept : mCtx(o.mCtx), mPtr(o.mPtr), mBytes(o.mBytes)
    {
        o.mPtr = 0;
    }
    DeviceBuffer& operator=(DeviceBuffer&& o) noexcept
    {
        if (this != &o) {
            release();
            mCtx = o.mCtx;
            mPtr = o.mPtr;
            mBytes = o.mBytes;
            o.mPtr = 0;
        }
        return *this;
    }
    void* get() const { return mPtr; }
    std::size_t bytes() const { return mBytes; }
    void upload(const void* src, std::size_t bytes)
    {
        deviceCheck(scafes_memcpy_h2d(mCtx, mPtr, src, bytes), "upload");
    }
    void download(void* dst, std::size_t bytes) const
    {
        deviceCheck(scafes_memcpy_d2h(mCtx, dst, mPtr, bytes), "download");
    }
    void swap(DeviceBuffer& o)
    {
        std::swap(mCtx, o.mCtx);
        std::swap(mPtr, o.mPtr);
        std::swap(mBytes, o.mBytes);
    }

private:
    void release()
    {
        if (mPtr) scafes_free(mCtx, mPtr);
        mPtr = 0;
    }
    scafes_ctx* mCtx;
    void* mPtr;
    std::size_t mBytes;
};

}  // namespace ScaFES
#endif
