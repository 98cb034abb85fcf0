// Minimal OpenMM-compatible core: the subset of NonbondedForce the ATM
// evaluation supports on the MI355X platform -- per-particle charge, sigma and
// epsilon with CutoffPeriodic (reaction-field) or PME electrostatics, and
// exceptions (scaled 1-4 pairs, exclusions) as AMBER systems use them, at the
// plain (non-periodic) distance; the dispersion correction is included in
// either method.
#pragma once
#include <unordered_map>
#include <vector>
#include "openmm/Force.h"
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class OPENMM_EXPORT NonbondedForce : public Force {
public:
    enum NonbondedMethod { NoCutoff = 0, CutoffNonPeriodic = 1, CutoffPeriodic = 2, Ewald = 3, PME = 4, LJPME = 5 };
    NonbondedForce();
    int getNumParticles() const { return (int)particles.size(); }
    int getNumExceptions() const { return (int)exceptions.size(); }
    NonbondedMethod getNonbondedMethod() const { return method; }
    void setNonbondedMethod(NonbondedMethod m);
    double getCutoffDistance() const { return cutoff; }
    void setCutoffDistance(double distance);
    double getReactionFieldDielectric() const { return rfDielectric; }
    void setReactionFieldDielectric(double d);
    bool getUseDispersionCorrection() const { return useDispersionCorrection; }
    void setUseDispersionCorrection(bool use) { useDispersionCorrection = use; }
    double getEwaldErrorTolerance() const { return ewaldTolerance; }
    void setEwaldErrorTolerance(double tol);
    // alpha = 0 (the default): chosen from the Ewald error tolerance, as is the grid
    void getPMEParameters(double& alpha, int& nx, int& ny, int& nz) const;
    void setPMEParameters(double alpha, int nx, int ny, int nz);
    int getReciprocalSpaceForceGroup() const { return recipGroup; }
    void setReciprocalSpaceForceGroup(int group);
    int addParticle(double charge, double sigma, double epsilon);
    void getParticleParameters(int index, double& charge, double& sigma, double& epsilon) const;
    void setParticleParameters(int index, double charge, double sigma, double epsilon);
    int addException(int particle1, int particle2, double chargeProd, double sigma, double epsilon, bool replace = false);
    void getExceptionParameters(int index, int& particle1, int& particle2, double& chargeProd, double& sigma,
                                double& epsilon) const;
    void setExceptionParameters(int index, int particle1, int particle2, double chargeProd, double sigma,
                                double epsilon);
    void updateParametersInContext(Context& context);
    bool usesPeriodicBoundaryConditions() const override {
        return method == CutoffPeriodic || method == Ewald || method == PME || method == LJPME;
    }

protected:
    ForceImpl* createImpl() const override;
    const char* defaultName() const override { return "NonbondedForce"; }

private:
    struct Particle {
        double q, sigma, eps;
    };
    struct Exception {
        int p1, p2;
        double qq, sigma, eps;
    };
    NonbondedMethod method;
    double cutoff, rfDielectric, ewaldTolerance, pmeAlpha;
    int pmeGrid[3];
    bool useDispersionCorrection;
    int recipGroup;
    std::vector<Particle> particles;
    std::vector<Exception> exceptions;
    std::unordered_map<unsigned long long, int> exceptionIndex;  // (min, max) particle pair -> exception
};

}  // namespace OpenMM
