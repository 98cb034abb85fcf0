// Minimal OpenMM-compatible core: a System bound to a Platform and an
// Integrator, holding the simulation state.
#pragma once
#include <map>
#include <set>
#include <string>
#include <vector>
#include "openmm/State.h"
#include "openmm/Vec3.h"
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class ContextImpl;
class Integrator;
class Platform;
class System;

class OPENMM_EXPORT Context {
public:
    Context(const System& system, Integrator& integrator);
    Context(const System& system, Integrator& integrator, Platform& platform);
    Context(const System& system, Integrator& integrator, Platform& platform,
            const std::map<std::string, std::string>& properties);
    ~Context();
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    const System& getSystem() const;
    Integrator& getIntegrator();
    Platform& getPlatform();
    // `groups` is a bit mask of force groups (default: all).
    State getState(int types, bool enforcePeriodicBox = false, int groups = 0xFFFFFFFF) const;
    State getState(int types, bool enforcePeriodicBox, const std::set<int>& groups) const;
    void setTime(double t);
    void setPositions(const std::vector<Vec3>& positions);
    void setVelocities(const std::vector<Vec3>& velocities);
    double getParameter(const std::string& name) const;
    void setParameter(const std::string& name, double value);
    const std::map<std::string, double>& getParameters() const;
    void setPeriodicBoxVectors(const Vec3& a, const Vec3& b, const Vec3& c);
    // Restores positions, velocities, time, box and parameters from a State.
    void setState(const State& state);

private:
    friend class Force;
    friend class ContextImpl;
    ContextImpl& getImpl() { return *impl; }
    const ContextImpl& getImpl() const { return *impl; }
    ContextImpl* impl;
};

}  // namespace OpenMM
