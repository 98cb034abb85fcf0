#ifndef ATMMETAFORCE_KERNELS_H_
#define ATMMETAFORCE_KERNELS_H_
/*
 * Kernel interface of ATMMetaForce (reference: openmmapi/include/ATMMetaForceKernels.h:15-59).
 *
 * The reference's interface receives the energies of two inner contexts that
 * OpenMM evaluated for it (execute(..., State1Energy, State2Energy, ...)) and
 * fills those contexts with x and x + d (copyState).  Those pure virtuals are
 * kept with the same signatures.  A kernel that evaluates both coordinate sets
 * itself -- the MI355X kernel, which runs the displacement, the pair kernel on
 * x and x + d and the combine on the GPU with no inner contexts -- returns true
 * from computesInnerStates() and implements executeFused(); the ForceImpl then
 * calls it instead of creating inner contexts.
 */
#include <string>
#include "ATMMetaForce.h"
#include "openmm/KernelImpl.h"
#include "openmm/OpenMMException.h"
#include "openmm/Platform.h"
#include "openmm/System.h"
#include "openmm/internal/ContextImpl.h"

namespace ATMMetaForcePlugin {

class CalcATMMetaForceKernel : public OpenMM::KernelImpl {
public:
    static std::string Name() { return "CalcATMMetaForce"; }
    CalcATMMetaForceKernel(std::string name, const OpenMM::Platform& platform) : OpenMM::KernelImpl(name, platform) {}
    /** Set up for `system` and `force` (displacements, variable-group forces). */
    virtual void initialize(const OpenMM::System& system, const ATMMetaForce& force) = 0;
    /** Combine inner-context energies U1 (State1Energy) and U2 and add the hybrid force;
     *  returns W if includeEnergy, else 0. */
    virtual double execute(OpenMM::ContextImpl& context, OpenMM::ContextImpl& innerContext1,
                           OpenMM::ContextImpl& innerContext2, double State1Energy, double State2Energy,
                           bool includeForces, bool includeEnergy) = 0;
    /** Copy changed displacements (the particle count and indices cannot change). */
    virtual void copyParametersToContext(OpenMM::ContextImpl& context, const ATMMetaForce& force) = 0;
    /** Copy x, x + d, box, time and global parameters into the two inner contexts. */
    virtual void copyState(OpenMM::ContextImpl& context, OpenMM::ContextImpl& innerContext1,
                           OpenMM::ContextImpl& innerContext2) = 0;
    /** Soft-core perturbation energy of the last evaluation. */
    virtual double getPerturbationEnergy(void) = 0;

    /** True when the kernel evaluates the variable-group energy at x and x + d itself. */
    virtual bool computesInnerStates() const { return false; }
    /** Whole ATM evaluation without inner contexts; returns W if includeEnergy, else 0. */
    virtual double executeFused(OpenMM::ContextImpl& context, bool includeForces, bool includeEnergy) {
        (void)context; (void)includeForces; (void)includeEnergy;
        throw OpenMM::OpenMMException("This CalcATMMetaForce kernel needs inner contexts");
    }
};

}  // namespace ATMMetaForcePlugin

#endif /*ATMMETAFORCE_KERNELS_H_*/
