#ifndef ATM_RESTRAINT_KERNELS_H_
#define ATM_RESTRAINT_KERNELS_H_
// Kernel interfaces of the flat-bottom restraint forces.
#include <string>
#include "ATMRestraintForces.h"
#include "openmm/KernelImpl.h"
#include "openmm/System.h"
#include "openmm/internal/ContextImpl.h"

namespace ATMMetaForcePlugin {

class CalcCentroidRestraintForceKernel : public OpenMM::KernelImpl {
public:
    static std::string Name() { return "CalcCentroidRestraintForce"; }
    CalcCentroidRestraintForceKernel(std::string name, const OpenMM::Platform& platform)
        : OpenMM::KernelImpl(name, platform) {}
    virtual void initialize(const OpenMM::System& system, const CentroidRestraintForce& force) = 0;
    virtual double execute(OpenMM::ContextImpl& context, bool includeForces, bool includeEnergy) = 0;
    virtual void copyParametersToContext(OpenMM::ContextImpl& context, const CentroidRestraintForce& force) = 0;
};

class CalcPositionRestraintForceKernel : public OpenMM::KernelImpl {
public:
    static std::string Name() { return "CalcPositionRestraintForce"; }
    CalcPositionRestraintForceKernel(std::string name, const OpenMM::Platform& platform)
        : OpenMM::KernelImpl(name, platform) {}
    virtual void initialize(const OpenMM::System& system, const PositionRestraintForce& force) = 0;
    virtual double execute(OpenMM::ContextImpl& context, bool includeForces, bool includeEnergy) = 0;
    virtual void copyParametersToContext(OpenMM::ContextImpl& context, const PositionRestraintForce& force) = 0;
};

}  // namespace ATMMetaForcePlugin

#endif
