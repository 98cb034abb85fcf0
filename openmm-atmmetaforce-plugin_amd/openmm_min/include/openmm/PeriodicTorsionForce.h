// Minimal OpenMM-compatible core: periodic torsions, E = k (1 + cos(n phi -
// phase)) with the IUPAC dihedral angle phi (OpenMM's PeriodicTorsionForce
// API: phase in radians, k in kJ/mol).
#pragma once
#include <vector>
#include "openmm/Force.h"
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class OPENMM_EXPORT PeriodicTorsionForce : public Force {
public:
    PeriodicTorsionForce() = default;
    int getNumTorsions() const { return (int)torsions.size(); }
    int addTorsion(int particle1, int particle2, int particle3, int particle4, int periodicity, double phase, double k);
    void getTorsionParameters(int index, int& particle1, int& particle2, int& particle3, int& particle4,
                              int& periodicity, double& phase, double& k) const;
    void setTorsionParameters(int index, int particle1, int particle2, int particle3, int particle4, int periodicity,
                              double phase, double k);
    void updateParametersInContext(Context& context);
    void setUsesPeriodicBoundaryConditions(bool periodic);
    bool usesPeriodicBoundaryConditions() const override { return false; }

protected:
    ForceImpl* createImpl() const override;
    const char* defaultName() const override { return "PeriodicTorsionForce"; }

private:
    struct Torsion {
        int p1, p2, p3, p4, periodicity;
        double phase, k;
    };
    std::vector<Torsion> torsions;
};

}  // namespace OpenMM
