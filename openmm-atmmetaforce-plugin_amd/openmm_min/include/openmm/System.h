// Minimal OpenMM-compatible core: particles, constraints, forces and the
// default periodic box of a molecular system.  The System owns its forces.
#pragma once
#include <vector>
#include "openmm/Force.h"
#include "openmm/OpenMMException.h"
#include "openmm/Vec3.h"
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class OPENMM_EXPORT System {
public:
    System();
    ~System();
    System(const System&) = delete;
    System& operator=(const System&) = delete;
    int getNumParticles() const { return (int)masses.size(); }
    int addParticle(double mass) { masses.push_back(mass); return (int)masses.size() - 1; }
    double getParticleMass(int index) const;
    void setParticleMass(int index, double mass);
    int getNumConstraints() const { return (int)constraints.size(); }
    int addConstraint(int particle1, int particle2, double distance);
    void getConstraintParameters(int index, int& particle1, int& particle2, double& distance) const;
    int getNumForces() const { return (int)forces.size(); }
    // Takes ownership of `force`.
    int addForce(Force* force);
    Force& getForce(int index);
    const Force& getForce(int index) const;
    void removeForce(int index);
    void getDefaultPeriodicBoxVectors(Vec3& a, Vec3& b, Vec3& c) const;
    void setDefaultPeriodicBoxVectors(const Vec3& a, const Vec3& b, const Vec3& c);
    bool usesPeriodicBoundaryConditions() const;

private:
    struct Constraint {
        int p1, p2;
        double d;
    };
    std::vector<double> masses;
    std::vector<Constraint> constraints;
    std::vector<Force*> forces;
    Vec3 box[3];
};

}  // namespace OpenMM
