// Minimal OpenMM-compatible core: integrators advance a Context in time.
#pragma once
#include <set>
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class Context;
class ContextImpl;

class OPENMM_EXPORT Integrator {
public:
    explicit Integrator(double stepSize)
        : stepSize(stepSize), constraintTolerance(1e-5), groups(0xFFFFFFFF), context(nullptr) {}
    virtual ~Integrator() {}
    double getStepSize() const { return stepSize; }
    void setStepSize(double s) { stepSize = s; }
    // Relative tolerance of the distance constraints (OpenMM's default 1e-5).
    double getConstraintTolerance() const { return constraintTolerance; }
    void setConstraintTolerance(double tol) { constraintTolerance = tol; }
    // Force groups whose forces drive the integration (all by default).
    int getIntegrationForceGroups() const { return groups; }
    void setIntegrationForceGroups(int mask) { groups = mask; }
    void setIntegrationForceGroups(const std::set<int>& s);
    virtual void step(int steps) = 0;

protected:
    friend class ContextImpl;
    virtual void initialize(ContextImpl& ctx) { context = &ctx; }
    double stepSize, constraintTolerance;
    int groups;
    ContextImpl* context;
};

// Velocity Verlet (leapfrog form) evaluated on the host around the platform's
// force evaluation; distance constraints by SHAKE on the new positions, the
// velocities then (x' - x) / dt.
class OPENMM_EXPORT VerletIntegrator : public Integrator {
public:
    explicit VerletIntegrator(double stepSize) : Integrator(stepSize) {}
    void step(int steps) override;
};

// Langevin dynamics, OpenMM's leapfrog form (the integrator of the ATM example
// scripts, example/abfe/abfe.py:100-101): per step
//   v = a v + (1 - a)/friction F/m + sqrt(kT (1 - a^2)/m) R,  x += dt v,
// a = exp(-friction dt) ((1 - a)/friction -> dt without friction).
// temperature K, friction 1/ps, step size ps.  Host update around the
// platform's forces; Gaussian noise from a seeded std::mt19937_64.  Distance
// constraints (createSystem's HBonds) by SHAKE on the new positions, the
// velocities then (x' - x) / dt.
class OPENMM_EXPORT LangevinIntegrator : public Integrator {
public:
    LangevinIntegrator(double temperature, double frictionCoeff, double stepSize)
        : Integrator(stepSize), temperature(temperature), friction(frictionCoeff), seed(0) {}
    double getTemperature() const { return temperature; }
    void setTemperature(double t) { temperature = t; }
    double getFriction() const { return friction; }
    void setFriction(double f) { friction = f; }
    int getRandomNumberSeed() const { return seed; }
    void setRandomNumberSeed(int s) { seed = s; }
    void step(int steps) override;

protected:
    double temperature, friction;
    int seed;
    unsigned long long drawn = 0;  // deviates drawn so far (the generator restarts from seed + drawn)
};

// LangevinMiddle / BAOAB splitting: v += dt F/m; x += dt/2 v;
// v = a v + sqrt(kT (1 - a^2)/m) R; x += dt/2 v (the scheme of the batched
// device MD, atm_md_step).  Constraints are not applied (a System with any is
// refused).
class OPENMM_EXPORT LangevinMiddleIntegrator : public LangevinIntegrator {
public:
    LangevinMiddleIntegrator(double temperature, double frictionCoeff, double stepSize)
        : LangevinIntegrator(temperature, frictionCoeff, stepSize) {}
    void step(int steps) override;
};

}  // namespace OpenMM
