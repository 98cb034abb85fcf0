// Minimal OpenMM-compatible core: the subset of NonbondedForce the ATM
// evaluation supports on the MI355X platform -- per-particle charge, sigma and
// epsilon, CutoffPeriodic with reaction-field electrostatics.  Exceptions are
// stored (so a System round-trips) but the MI355X kernels reject nonzero ones;
// the dispersion correction is evaluated on the host.
#pragma once
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
    int getReciprocalSpaceForceGroup() const { return recipGroup; }
    void setReciprocalSpaceForceGroup(int group);
    int addParticle(double charge, double sigma, double epsilon);
    void getParticleParameters(int index, double& charge, double& sigma, double& epsilon) const;
    void setParticleParameters(int index, double charge, double sigma, double epsilon);
    int addException(int particle1, int particle2, double chargeProd, double sigma, double epsilon, bool replace = false);
    void getExceptionParameters(int index, int& particle1, int& particle2, double& chargeProd, double& sigma,
                                double& epsilon) const;
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
    double cutoff, rfDielectric;
    bool useDispersionCorrection;
    int recipGroup;
    std::vector<Particle> particles;
    std::vector<Exception> exceptions;
};

}  // namespace OpenMM
