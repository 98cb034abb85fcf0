// ATMMetaForceImpl: the per-step orchestration of the ATM evaluation
// (reference: openmmapi/src/ATMMetaForceImpl.cpp:69-153).
//
// The reference clones the System into two inner Systems, evaluates the
// variable force groups in two linked inner Contexts at x and x + d, and hands
// their energies to the platform kernel.  Here the platform kernel evaluates
// both coordinate sets itself (CalcATMMetaForceKernel::computesInnerStates),
// so no inner Context exists; the group gate, the variable-group mask, the
// validation and the cached perturbation energy behave as in the reference.
#include "internal/ATMMetaForceImpl.h"
#include "ATMMetaForceKernels.h"
#include "openmm/OpenMMException.h"
#include "openmm/Platform.h"
#include "openmm/internal/ContextImpl.h"

using namespace ATMMetaForcePlugin;
using namespace OpenMM;

ATMMetaForceImpl::ATMMetaForceImpl(const ATMMetaForce& owner)
    : owner(owner), PerturbationEnergy(0.0), variable_force_groups_mask(0) {}

ATMMetaForceImpl::~ATMMetaForceImpl() {}

void ATMMetaForceImpl::initialize(ContextImpl& context) {
    const int atmGroup = owner.getForceGroup();
    variable_force_groups_mask = 0;
    for (int g : owner.getVariableForceGroups()) {
        if (g == atmGroup) throw OpenMMException("The ATM Meta Force group cannot be one of the variable force groups.");
        if (g < 0 || g > 31) throw OpenMMException("ATMMetaForce: variable force groups must be between 0 and 31");
        variable_force_groups_mask |= 1 << g;
    }
    kernel = context.getPlatform().createKernel(CalcATMMetaForceKernel::Name(), context);
    kernel.getAs<CalcATMMetaForceKernel>().initialize(context.getSystem(), owner);
}

double ATMMetaForceImpl::calcForcesAndEnergy(ContextImpl& context, bool includeForces, bool includeEnergy, int groups) {
    if ((groups & (1 << owner.getForceGroup())) == 0) return 0.0;
    CalcATMMetaForceKernel& k = kernel.getAs<CalcATMMetaForceKernel>();
    if (!k.computesInnerStates())
        throw OpenMMException("ATMMetaForce: the platform's CalcATMMetaForce kernel needs inner contexts, "
                              "which this build does not create");
    const double energy = k.executeFused(context, includeForces, includeEnergy);
    PerturbationEnergy = k.getPerturbationEnergy();
    return includeEnergy ? energy : 0.0;
}

std::map<std::string, double> ATMMetaForceImpl::getDefaultParameters() {
    return {{ATMMetaForce::Lambda1(), owner.getDefaultLambda1()}, {ATMMetaForce::Lambda2(), owner.getDefaultLambda2()},
            {ATMMetaForce::Alpha(), owner.getDefaultAlpha()},     {ATMMetaForce::U0(), owner.getDefaultU0()},
            {ATMMetaForce::W0(), owner.getDefaultW0()},           {ATMMetaForce::Umax(), owner.getDefaultUmax()},
            {ATMMetaForce::Ubcore(), owner.getDefaultUbcore()},   {ATMMetaForce::Acore(), owner.getDefaultAcore()},
            {ATMMetaForce::Direction(), owner.getDefaultDirection()}};
}

std::vector<std::string> ATMMetaForceImpl::getKernelNames() { return {CalcATMMetaForceKernel::Name()}; }

void ATMMetaForceImpl::updateParametersInContext(ContextImpl& context) {
    kernel.getAs<CalcATMMetaForceKernel>().copyParametersToContext(context, owner);
}
