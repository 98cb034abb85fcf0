// Minimal OpenMM-compatible core: the implementation behind a Context.
// Positions, velocities and forces live on the host in FP64; a platform
// stores its own data (device buffers) in platformData.
#pragma once
#include <map>
#include <string>
#include <vector>
#include "openmm/Vec3.h"
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class Context;
class ForceImpl;
class Integrator;
class Platform;
class System;

class OPENMM_EXPORT ContextImpl {
public:
    ContextImpl(Context& owner, const System& system, Integrator& integrator, Platform* platform,
                const std::map<std::string, std::string>& properties);
    ~ContextImpl();
    Context& getOwner() { return owner; }
    const System& getSystem() const { return system; }
    Integrator& getIntegrator() { return integrator; }
    Platform& getPlatform() { return *platform; }
    const std::vector<ForceImpl*>& getForceImpls() const { return forceImpls; }
    double getTime() const { return time; }
    void setTime(double t) { time = t; }
    void getPositions(std::vector<Vec3>& p) const { p = positions; }
    void setPositions(const std::vector<Vec3>& p);
    std::vector<Vec3>& positionsRef() { return positions; }
    void getVelocities(std::vector<Vec3>& v) const { v = velocities; }
    void setVelocities(const std::vector<Vec3>& v);
    std::vector<Vec3>& velocitiesRef() { return velocities; }
    // Forces of the last calcForcesAndEnergy (sum over the evaluated groups).
    std::vector<Vec3>& forcesRef() { return forces; }
    void getForces(std::vector<Vec3>& f) const { f = forces; }
    double getParameter(const std::string& name) const;
    void setParameter(const std::string& name, double value);
    const std::map<std::string, double>& getParameters() const { return parameters; }
    void getPeriodicBoxVectors(Vec3& a, Vec3& b, Vec3& c) const { a = box[0]; b = box[1]; c = box[2]; }
    void setPeriodicBoxVectors(const Vec3& a, const Vec3& b, const Vec3& c);
    // Zero the forces, then let every ForceImpl add its contribution for the
    // requested groups; returns the potential energy of those groups.
    double calcForcesAndEnergy(bool includeForces, bool includeEnergy, int groups = 0xFFFFFFFF);
    // Lets every force update the state before a step (barostats, thermostats).
    void updateContextState();
    double calcKineticEnergy() const;
    void* getPlatformData() { return platformData; }
    const void* getPlatformData() const { return platformData; }
    void setPlatformData(void* data) { platformData = data; }
    const std::map<std::string, std::string>& getProperties() const { return properties; }

private:
    Context& owner;
    const System& system;
    Integrator& integrator;
    Platform* platform;
    std::vector<ForceImpl*> forceImpls;
    std::vector<Vec3> positions, velocities, forces;
    std::map<std::string, double> parameters;
    std::map<std::string, std::string> properties;
    Vec3 box[3];
    double time;
    void* platformData;
};

}  // namespace OpenMM
