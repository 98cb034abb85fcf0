// Minimal OpenMM-compatible core: kernel interfaces of the built-in forces.
#pragma once
#include <string>
#include "openmm/KernelImpl.h"
#include "openmm/NonbondedForce.h"
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

}  // namespace OpenMM
