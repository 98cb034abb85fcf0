// Minimal OpenMM-compatible core: per-context implementation of a Force.
#pragma once
#include <map>
#include <string>
#include <vector>
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class ContextImpl;
class Force;

class OPENMM_EXPORT ForceImpl {
public:
    virtual ~ForceImpl() {}
    virtual void initialize(ContextImpl& context) = 0;
    virtual const Force& getOwner() const = 0;
    virtual void updateContextState(ContextImpl& context, bool& forcesInvalid) = 0;
    // Adds this force's contribution to the context's forces and returns its
    // energy, if its group is in `groups` (a bit mask of force groups).
    virtual double calcForcesAndEnergy(ContextImpl& context, bool includeForces, bool includeEnergy, int groups) = 0;
    virtual std::map<std::string, double> getDefaultParameters() = 0;
    virtual std::vector<std::string> getKernelNames() = 0;
};

}  // namespace OpenMM
