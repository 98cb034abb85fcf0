#ifndef OPENMM_ATMRESTRAINTFORCES_H_
#define OPENMM_ATMRESTRAINTFORCES_H_
// Flat-bottom restraints of ATM binding free energy workflows, evaluated
// beside the ATM force every step (force group 0 in
// python/tests/test_abfe.py:65-79).  The reference builds them in
// python/ATMMetaForceUtils.py as OpenMM custom forces:
//   CentroidRestraintForce  <- addVsiteRestraintForceCMCM (:79-138), a
//     CustomCentroidBondForce with per-bond (kfcm, tolcm, offx, offy, offz);
//   PositionRestraintForce  <- addPosRestraints (:707-753), a
//     CustomExternalForce with per-particle (x0, y0, z0, fc, tol).
// Energies: kf/2 (d - tol)^2 for d > tol, else 0; units nm, kJ/mol.
#include <vector>
#include "internal/windowsExportATMMetaForce.h"
#include "openmm/Context.h"
#include "openmm/Force.h"
#include "openmm/Vec3.h"

namespace ATMMetaForcePlugin {

class OPENMM_EXPORT_ATMMETAFORCE CentroidRestraintForce : public OpenMM::Force {
public:
    CentroidRestraintForce() {}
    int getNumGroups() const { return (int)groups.size(); }
    int getNumBonds() const { return (int)bonds.size(); }
    // An empty weight list weights the members by their masses (as
    // CustomCentroidBondForce does).
    int addGroup(const std::vector<int>& particles, const std::vector<double>& weights = std::vector<double>());
    void getGroupParameters(int index, std::vector<int>& particles, std::vector<double>& weights) const;
    // E = kf/2 (d - tol)^2 for d = |c1 - offset - c2| > tol.
    int addBond(int group1, int group2, double kf, double tol, const OpenMM::Vec3& offset = OpenMM::Vec3());
    void getBondParameters(int index, int& group1, int& group2, double& kf, double& tol, OpenMM::Vec3& offset) const;
    void setBondParameters(int index, int group1, int group2, double kf, double tol, const OpenMM::Vec3& offset);
    void updateParametersInContext(OpenMM::Context& context);
    bool usesPeriodicBoundaryConditions() const override { return false; }

protected:
    OpenMM::ForceImpl* createImpl() const override;
    const char* defaultName() const override { return "CentroidRestraintForce"; }

private:
    struct Group {
        std::vector<int> particles;
        std::vector<double> weights;
    };
    struct Bond {
        int g1, g2;
        double kf, tol;
        OpenMM::Vec3 offset;
    };
    std::vector<Group> groups;
    std::vector<Bond> bonds;
};

class OPENMM_EXPORT_ATMMETAFORCE PositionRestraintForce : public OpenMM::Force {
public:
    // periodic: distances to the reference positions use the minimum image.
    explicit PositionRestraintForce(bool periodic = true) : periodic(periodic) {}
    bool getPeriodic() const { return periodic; }
    int getNumParticles() const { return (int)entries.size(); }
    // E = kf/2 (d - tol)^2 for d = |x - ref| > tol.
    int addParticle(int particle, const OpenMM::Vec3& ref, double kf, double tol);
    void getParticleParameters(int index, int& particle, OpenMM::Vec3& ref, double& kf, double& tol) const;
    void setParticleParameters(int index, int particle, const OpenMM::Vec3& ref, double kf, double tol);
    void updateParametersInContext(OpenMM::Context& context);
    bool usesPeriodicBoundaryConditions() const override { return periodic; }

protected:
    OpenMM::ForceImpl* createImpl() const override;
    const char* defaultName() const override { return "PositionRestraintForce"; }

private:
    struct Entry {
        int particle;
        OpenMM::Vec3 ref;
        double kf, tol;
    };
    bool periodic;
    std::vector<Entry> entries;
};

}  // namespace ATMMetaForcePlugin

#endif
