// Minimal OpenMM-compatible core: integrators advance a Context in time.
#pragma once
#include <set>
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class Context;
class ContextImpl;

class OPENMM_EXPORT Integrator {
public:
    explicit Integrator(double stepSize) : stepSize(stepSize), groups(0xFFFFFFFF), context(nullptr) {}
    virtual ~Integrator() {}
    double getStepSize() const { return stepSize; }
    void setStepSize(double s) { stepSize = s; }
    // Force groups whose forces drive the integration (all by default).
    int getIntegrationForceGroups() const { return groups; }
    void setIntegrationForceGroups(int mask) { groups = mask; }
    void setIntegrationForceGroups(const std::set<int>& s);
    virtual void step(int steps) = 0;

protected:
    friend class ContextImpl;
    virtual void initialize(ContextImpl& ctx) { context = &ctx; }
    double stepSize;
    int groups;
    ContextImpl* context;
};

// Velocity Verlet (leapfrog form) evaluated on the host around the platform's
// force evaluation.  Constraints are not applied.
class OPENMM_EXPORT VerletIntegrator : public Integrator {
public:
    explicit VerletIntegrator(double stepSize) : Integrator(stepSize) {}
    void step(int steps) override;
};

}  // namespace OpenMM
