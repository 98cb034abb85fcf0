#ifndef OPENMM_ATMMETAFORCEFORCEIMPL_H_
#define OPENMM_ATMMETAFORCEFORCEIMPL_H_
// Per-context implementation of ATMMetaForce
// (reference: openmmapi/include/internal/ATMMetaForceImpl.h:22-57).
#include <map>
#include <string>
#include <vector>
#include "ATMMetaForce.h"
#include "internal/windowsExportATMMetaForce.h"
#include "openmm/Kernel.h"
#include "openmm/internal/ForceImpl.h"

namespace ATMMetaForcePlugin {

class OPENMM_EXPORT_ATMMETAFORCE ATMMetaForceImpl : public OpenMM::ForceImpl {
public:
    explicit ATMMetaForceImpl(const ATMMetaForce& owner);
    ~ATMMetaForceImpl() override;
    void initialize(OpenMM::ContextImpl& context) override;
    const ATMMetaForce& getOwner() const override { return owner; }
    void updateContextState(OpenMM::ContextImpl& context, bool& forcesInvalid) override {}
    double calcForcesAndEnergy(OpenMM::ContextImpl& context, bool includeForces, bool includeEnergy, int groups) override;
    std::map<std::string, double> getDefaultParameters() override;
    std::vector<std::string> getKernelNames() override;
    void updateParametersInContext(OpenMM::ContextImpl& context);
    double getPerturbationEnergy() const { return PerturbationEnergy; }
    int getVariableForceGroupsMask() const { return variable_force_groups_mask; }

private:
    const ATMMetaForce& owner;
    OpenMM::Kernel kernel;
    double PerturbationEnergy;
    int variable_force_groups_mask;
};

}  // namespace ATMMetaForcePlugin

#endif /*OPENMM_ATMMETAFORCEIMPL_H_*/
