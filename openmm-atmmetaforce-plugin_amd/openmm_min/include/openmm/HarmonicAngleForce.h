// Minimal OpenMM-compatible core: harmonic angles, E = k/2 (theta - theta0)^2
// (OpenMM's HarmonicAngleForce API: particle2 is the vertex; angle in radians,
// k in kJ/mol/rad^2).
#pragma once
#include <vector>
#include "openmm/Force.h"
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class OPENMM_EXPORT HarmonicAngleForce : public Force {
public:
    HarmonicAngleForce() = default;
    int getNumAngles() const { return (int)angles.size(); }
    int addAngle(int particle1, int particle2, int particle3, double angle, double k);
    void getAngleParameters(int index, int& particle1, int& particle2, int& particle3, double& angle, double& k) const;
    void setAngleParameters(int index, int particle1, int particle2, int particle3, double angle, double k);
    void updateParametersInContext(Context& context);
    void setUsesPeriodicBoundaryConditions(bool periodic);
    bool usesPeriodicBoundaryConditions() const override { return false; }

protected:
    ForceImpl* createImpl() const override;
    const char* defaultName() const override { return "HarmonicAngleForce"; }

private:
    struct Angle {
        int p1, p2, p3;
        double angle, k;
    };
    std::vector<Angle> angles;
};

}  // namespace OpenMM
