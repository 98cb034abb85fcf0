// Minimal OpenMM-compatible core: a snapshot of a Context.
#pragma once
#include <map>
#include <string>
#include <vector>
#include "openmm/OpenMMException.h"
#include "openmm/Vec3.h"
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class OPENMM_EXPORT State {
public:
    enum DataType { Positions = 1, Velocities = 2, Forces = 4, Energy = 8, Parameters = 16 };
    State() : types(0), time(0.0), potentialEnergy(0.0), kineticEnergy(0.0) {}
    int getDataTypes() const { return types; }
    double getTime() const { return time; }
    const std::vector<Vec3>& getPositions() const;
    const std::vector<Vec3>& getVelocities() const;
    const std::vector<Vec3>& getForces() const;
    double getPotentialEnergy() const;
    double getKineticEnergy() const;
    const std::map<std::string, double>& getParameters() const;
    void getPeriodicBoxVectors(Vec3& a, Vec3& b, Vec3& c) const { a = box[0]; b = box[1]; c = box[2]; }

private:
    friend class Context;
    friend class StateBuilder;
    friend class StateProxy;
    int types;
    double time, potentialEnergy, kineticEnergy;
    std::vector<Vec3> positions, velocities, forces;
    std::map<std::string, double> parameters;
    Vec3 box[3];
};

}  // namespace OpenMM
