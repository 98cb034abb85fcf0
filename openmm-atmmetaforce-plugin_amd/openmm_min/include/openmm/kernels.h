// Minimal OpenMM-compatible core: kernel interfaces of the built-in forces.
#pragma once
#include <string>
#include "openmm/HarmonicAngleForce.h"
#include "openmm/HarmonicBondForce.h"
#include "openmm/KernelImpl.h"
#include "openmm/NonbondedForce.h"
#include "openmm/PeriodicTorsionForce.h"
#include "openmm/Platform.h"
#include "openmm/System.h"

namespace OpenMM {

class ContextImpl;

// Energy and forces of a NonbondedForce.
class CalcNonbondedForceKernel : public KernelImpl {
public:
    static std::string Name() { return "CalcNonbondedForce"; }
    CalcNonbondedForceKernel(std::string name, const Platform& platform) : KernelImpl(name, platform) {}
    virtual void initialize(const System& system, const NonbondedForce& force) = 0;
    virtual double execute(ContextImpl& context, bool includeForces, bool includeEnergy) = 0;
    virtual void copyParametersToContext(ContextImpl& context, const NonbondedForce& force) = 0;
};

// Energy and forces of a bonded force F (HarmonicBondForce, HarmonicAngleForce,
// PeriodicTorsionForce).
template <class F>
class CalcBondedForceKernel : public KernelImpl {
public:
    CalcBondedForceKernel(std::string name, const Platform& platform) : KernelImpl(name, platform) {}
    virtual void initialize(const System& system, const F& force) = 0;
    virtual double execute(ContextImpl& context, bool includeForces, bool includeEnergy) = 0;
    virtual void copyParametersToContext(ContextImpl& context, const F& force) = 0;
};

class CalcHarmonicBondForceKernel : public CalcBondedForceKernel<HarmonicBondForce> {
public:
    static std::string Name() { return "CalcHarmonicBondForce"; }
    using CalcBondedForceKernel<HarmonicBondForce>::CalcBondedForceKernel;
};

class CalcHarmonicAngleForceKernel : public CalcBondedForceKernel<HarmonicAngleForce> {
public:
    static std::string Name() { return "CalcHarmonicAngleForce"; }
    using CalcBondedForceKernel<HarmonicAngleForce>::CalcBondedForceKernel;
};

class CalcPeriodicTorsionForceKernel : public CalcBondedForceKernel<PeriodicTorsionForce> {
public:
    static std::string Name() { return "CalcPeriodicTorsionForce"; }
    using CalcBondedForceKernel<PeriodicTorsionForce>::CalcBondedForceKernel;
};

}  // namespace OpenMM
