// Minimal OpenMM-compatible core: harmonic bonds, E = k/2 (r - r0)^2 (OpenMM's
// HarmonicBondForce API: length in nm, k in kJ/mol/nm^2).
#pragma once
#include <vector>
#include "openmm/Force.h"
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class OPENMM_EXPORT HarmonicBondForce : public Force {
public:
    HarmonicBondForce() = default;
    int getNumBonds() const { return (int)bonds.size(); }
    int addBond(int particle1, int particle2, double length, double k);
    void getBondParameters(int index, int& particle1, int& particle2, double& length, double& k) const;
    void setBondParameters(int index, int particle1, int particle2, double length, double k);
    void updateParametersInContext(Context& context);
    void setUsesPeriodicBoundaryConditions(bool periodic);
    bool usesPeriodicBoundaryConditions() const override { return false; }

protected:
    ForceImpl* createImpl() const override;
    const char* defaultName() const override { return "HarmonicBondForce"; }

private:
    struct Bond {
        int p1, p2;
        double length, k;
    };
    std::vector<Bond> bonds;
};

}  // namespace OpenMM
