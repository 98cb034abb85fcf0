#ifndef OPENMM_ATMMETAFORCE_H_
#define OPENMM_ATMMETAFORCE_H_
/*
 * ATMMetaForce: the Alchemical Transfer Method potential as an OpenMM Force.
 *
 * Public interface identical to the reference plugin's
 * (openmmapi/include/ATMMetaForce.h:54-335): the constructor's nine parameters
 * and variable force groups, per-particle displacement vectors, the nine
 * global-parameter names, the default-value getters and
 * getPerturbationEnergy.  ATM and its parameters are described in
 * Azimi et al., J. Chem. Inf. Model. 62, 309 (2022).
 *
 * For each evaluation the force computes U1 = U(x) and U2 = U(x + d) of the
 * forces in the variable force groups, the perturbation energy
 * u = U2 - U1 (direction > 0) or U1 - U2 (otherwise), its soft-core map
 * u_sc(u; umax, ubcore, acore), and the alchemical potential
 *   W = e0 + (lambda2 - lambda1)/alpha ln(1 + exp(-alpha (u_sc - u0))) + lambda2 u_sc + w0
 * (the logarithmic term only for alpha > 0) with e0 = U1 (direction > 0) or U2,
 * and its forces, a blend of the two states' forces.
 */
#include <string>
#include <vector>
#include "ATMMetaForceVersion.h"
#include "internal/windowsExportATMMetaForce.h"
#include "openmm/Context.h"
#include "openmm/Force.h"

namespace ATMMetaForcePlugin {

class OPENMM_EXPORT_ATMMETAFORCE ATMMetaForce : public OpenMM::Force {
public:
    /**
     * lambda1, lambda2 (dimensionless), alpha ((kJ/mol)^-1), u0, w0 (kJ/mol):
     * softplus parameters; umax, ubcore (kJ/mol) and acore (dimensionless):
     * soft-core parameters; direction: > 0 transfers from the reference state
     * to the displaced one.  The values become the Context's global parameters
     * named by Lambda1() ... Direction().  VariableForceGroups: the groups whose
     * forces are re-evaluated at the displaced coordinates.
     */
    ATMMetaForce(double lambda1, double lambda2, double alpha, double u0, double w0, double umax, double ubcore,
                 double acore, double direction, const std::vector<int>& VariableForceGroups)
        : defaultLambda1(lambda1), defaultLambda2(lambda2), defaultAlpha(alpha), defaultU0(u0), defaultW0(w0),
          defaultUmax(umax), defaultUbcore(ubcore), defaultAcore(acore), defaultDirection(direction),
          VariableForceGroups(VariableForceGroups) {}

    /** Number of particles added (normally every particle of the System, in order). */
    int getNumParticles() const { return (int)particles.size(); }
    /** Add particle `particle` with displacement (dx, dy, dz) nm; returns its index in this force. */
    int addParticle(int particle, double dx, double dy, double dz);
    void getParticleParameters(int index, int& particle, double& dx, double& dy, double& dz) const;
    void setParticleParameters(int index, int particle, double dx, double dy, double dz);
    /** Push changed displacements into a Context (the particle count cannot change). */
    void updateParametersInContext(OpenMM::Context& context);
    /** Always false, as in the reference: the periodic forces are in the System itself. */
    bool usesPeriodicBoundaryConditions() const override { return false; }
    /** The soft-core perturbation energy u_sc of the last evaluation in `context`. */
    double getPerturbationEnergy(const OpenMM::Context& context) const;

    static const std::string& Lambda1() { static const std::string k = "ATMLambda1"; return k; }
    static const std::string& Lambda2() { static const std::string k = "ATMLambda2"; return k; }
    static const std::string& Alpha() { static const std::string k = "ATMAlpha"; return k; }
    static const std::string& U0() { static const std::string k = "ATMU0"; return k; }
    static const std::string& W0() { static const std::string k = "ATMW0"; return k; }
    static const std::string& Umax() { static const std::string k = "ATMUmax"; return k; }
    static const std::string& Ubcore() { static const std::string k = "ATMUbcore"; return k; }
    static const std::string& Acore() { static const std::string k = "ATMAcore"; return k; }
    static const std::string& Direction() { static const std::string k = "ATMDirection"; return k; }
    static const std::string& Version() { static const std::string v = ATMMETAFORCE_VERSION; return v; }

    double getDefaultLambda1() const { return defaultLambda1; }
    double getDefaultLambda2() const { return defaultLambda2; }
    double getDefaultAlpha() const { return defaultAlpha; }
    double getDefaultU0() const { return defaultU0; }
    double getDefaultW0() const { return defaultW0; }
    double getDefaultUmax() const { return defaultUmax; }
    double getDefaultUbcore() const { return defaultUbcore; }
    double getDefaultAcore() const { return defaultAcore; }
    double getDefaultDirection() const { return defaultDirection; }
    const std::vector<int>& getVariableForceGroups() const { return VariableForceGroups; }

protected:
    OpenMM::ForceImpl* createImpl() const override;
    const char* defaultName() const override { return "ATMMetaForce"; }

private:
    class ParticleInfo;
    std::vector<ParticleInfo> particles;
    double defaultLambda1, defaultLambda2, defaultAlpha, defaultU0, defaultW0;
    double defaultUmax, defaultUbcore, defaultAcore;
    double defaultDirection;
    std::vector<int> VariableForceGroups;
};

/** @private  Per-particle record: the System particle index and its displacement (nm). */
class ATMMetaForce::ParticleInfo {
public:
    int particle;
    double dx, dy, dz;
    ParticleInfo() : particle(-1), dx(0.0), dy(0.0), dz(0.0) {}
    explicit ParticleInfo(int particle) : particle(particle), dx(0.0), dy(0.0), dz(0.0) {}
    ParticleInfo(int particle, double dx, double dy, double dz) : particle(particle), dx(dx), dy(dy), dz(dz) {}
};

}  // namespace ATMMetaForcePlugin

#endif /* OPENMM_ATMMETAFORCE_H_ */
