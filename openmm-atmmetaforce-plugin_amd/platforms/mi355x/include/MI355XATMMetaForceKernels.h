#ifndef ATM_MI355X_KERNELS_H_
#define ATM_MI355X_KERNELS_H_
// MI355X implementations of the CalcATMMetaForce and CalcNonbondedForce
// kernels.  Each owns one libatm_mi355x context (include/atm_mi355x.h) holding
// one system; status codes of the C ABI are rethrown as OpenMMException.
#include <string>
#include <vector>
#include "ATMMetaForceKernels.h"
#include "ATMRestraintKernels.h"
#include "atm_mi355x.h"
#include "openmm/NonbondedForce.h"
#include "openmm/Vec3.h"
#include "openmm/kernels.h"

namespace ATMMetaForcePlugin {

// One system of the C ABI built from a NonbondedForce (+ displacements).
class MI355XSystemContext {
public:
    ~MI355XSystemContext();
    void create(int device, const OpenMM::System& system, const OpenMM::NonbondedForce& nb,
                const std::vector<OpenMM::Vec3>* displacements);
    void setParticles(const OpenMM::NonbondedForce& nb, const std::vector<OpenMM::Vec3>* displacements);
    void setDisplacements(const std::vector<OpenMM::Vec3>& displacements);
    // energies of the other variable-group forces at x and at x + d (added to U1, U2)
    void setEnergyOffsets(double state1, double state2);
    // Upload box (re-laying out the cell grid when it changed) and positions.
    void syncState(OpenMM::ContextImpl& context);
    // Evaluate with the given C-ABI flags; re-runs once if a list overflowed.
    void evaluate(int flags);
    void* handle() const { return ctx; }
    int numAtoms() const { return n; }
    double dispersionEnergy(const OpenMM::Vec3 box[3]) const;
    // PME NonbondedForce (atm_pme_*): energies include the dispersion correction
    bool isPME() const { return pme != nullptr; }
    // One PME evaluation at the context's positions and box: with params, the ATM
    // evaluation (U1, U2, combine, hybrid force); without, the force alone at x.
    // Forces are added to the context's force buffer when includeForces.
    void evaluatePME(OpenMM::ContextImpl& context, const atm_params* params, bool includeForces, atm_result* r);

private:
    void createPME(int device, const OpenMM::System& system, const OpenMM::NonbondedForce& nb,
                   const std::vector<OpenMM::Vec3>* displacements);
    void setPMEParticles(const OpenMM::NonbondedForce& nb, const std::vector<OpenMM::Vec3>* displacements);
    void* ctx = nullptr;
    void* pme = nullptr;
    int n = 0;
    double box[3] = {0, 0, 0};
    double tilt[3] = {0, 0, 0};        // (b_x, c_x, c_y) of a triclinic box
    double cutoff = 0, rfDielectric = 0;
    double dispersionCoefficient = 0;  // E_disp = coefficient / volume
    bool haveExceptions = false;       // exceptions were uploaded (atm_set_exceptions)
};

class MI355XCalcATMMetaForceKernel : public CalcATMMetaForceKernel {
public:
    MI355XCalcATMMetaForceKernel(std::string name, const OpenMM::Platform& platform, OpenMM::ContextImpl& context);
    ~MI355XCalcATMMetaForceKernel();
    void initialize(const OpenMM::System& system, const ATMMetaForce& force) override;
    double execute(OpenMM::ContextImpl& context, OpenMM::ContextImpl& innerContext1, OpenMM::ContextImpl& innerContext2,
                   double State1Energy, double State2Energy, bool includeForces, bool includeEnergy) override;
    void copyParametersToContext(OpenMM::ContextImpl& context, const ATMMetaForce& force) override;
    void copyState(OpenMM::ContextImpl& context, OpenMM::ContextImpl& innerContext1,
                   OpenMM::ContextImpl& innerContext2) override;
    double getPerturbationEnergy(void) override { return PerturbationEnergy; }
    bool computesInnerStates() const override { return true; }
    double executeFused(OpenMM::ContextImpl& context, bool includeForces, bool includeEnergy) override;
    // U1, U2 of the last evaluation (for diagnostics and tests)
    double lastState1Energy() const { return U1; }
    double lastState2Energy() const { return U2; }

private:
    OpenMM::ContextImpl& context;
    MI355XSystemContext sys;
    bool hasVariableForces = false;
    std::vector<int> particles;
    std::vector<OpenMM::Vec3> displacements;
    void* variableBonded = nullptr;  // bonded terms of the variable groups (atm_bonded_*), or null
    double PerturbationEnergy = 0.0, U1 = 0.0, U2 = 0.0;
};

class MI355XCalcNonbondedForceKernel : public OpenMM::CalcNonbondedForceKernel {
public:
    MI355XCalcNonbondedForceKernel(std::string name, const OpenMM::Platform& platform, OpenMM::ContextImpl& context)
        : OpenMM::CalcNonbondedForceKernel(name, platform), context(context) {}
    void initialize(const OpenMM::System& system, const OpenMM::NonbondedForce& force) override;
    double execute(OpenMM::ContextImpl& context, bool includeForces, bool includeEnergy) override;
    void copyParametersToContext(OpenMM::ContextImpl& context, const OpenMM::NonbondedForce& force) override;

private:
    OpenMM::ContextImpl& context;
    MI355XSystemContext sys;
};

// Bonded kernels: one libatm_mi355x bonded set each (atm_bonded_*) holding the
// terms of one HarmonicBondForce, HarmonicAngleForce or PeriodicTorsionForce.
template <class F, class Base>
class MI355XCalcBondedKernel : public Base {
public:
    MI355XCalcBondedKernel(std::string name, const OpenMM::Platform& platform, OpenMM::ContextImpl& context)
        : Base(name, platform), context(context) {}
    ~MI355XCalcBondedKernel() override;
    void initialize(const OpenMM::System& system, const F& force) override;
    double execute(OpenMM::ContextImpl& context, bool includeForces, bool includeEnergy) override;
    void copyParametersToContext(OpenMM::ContextImpl& context, const F& force) override;

private:
    void upload(const F& force);
    OpenMM::ContextImpl& context;
    void* set = nullptr;
    int n = 0;
};
using MI355XCalcHarmonicBondForceKernel =
    MI355XCalcBondedKernel<OpenMM::HarmonicBondForce, OpenMM::CalcHarmonicBondForceKernel>;
using MI355XCalcHarmonicAngleForceKernel =
    MI355XCalcBondedKernel<OpenMM::HarmonicAngleForce, OpenMM::CalcHarmonicAngleForceKernel>;
using MI355XCalcPeriodicTorsionForceKernel =
    MI355XCalcBondedKernel<OpenMM::PeriodicTorsionForce, OpenMM::CalcPeriodicTorsionForceKernel>;

// Restraint kernels: one libatm_mi355x restraint set each (atm_restraints_*).
class MI355XCalcCentroidRestraintForceKernel : public CalcCentroidRestraintForceKernel {
public:
    MI355XCalcCentroidRestraintForceKernel(std::string name, const OpenMM::Platform& platform, OpenMM::ContextImpl& context)
        : CalcCentroidRestraintForceKernel(name, platform), context(context) {}
    ~MI355XCalcCentroidRestraintForceKernel() override;
    void initialize(const OpenMM::System& system, const CentroidRestraintForce& force) override;
    double execute(OpenMM::ContextImpl& context, bool includeForces, bool includeEnergy) override;
    void copyParametersToContext(OpenMM::ContextImpl& context, const CentroidRestraintForce& force) override;

private:
    OpenMM::ContextImpl& context;
    void* set = nullptr;
};

class MI355XCalcPositionRestraintForceKernel : public CalcPositionRestraintForceKernel {
public:
    MI355XCalcPositionRestraintForceKernel(std::string name, const OpenMM::Platform& platform, OpenMM::ContextImpl& context)
        : CalcPositionRestraintForceKernel(name, platform), context(context) {}
    ~MI355XCalcPositionRestraintForceKernel() override;
    void initialize(const OpenMM::System& system, const PositionRestraintForce& force) override;
    double execute(OpenMM::ContextImpl& context, bool includeForces, bool includeEnergy) override;
    void copyParametersToContext(OpenMM::ContextImpl& context, const PositionRestraintForce& force) override;

private:
    OpenMM::ContextImpl& context;
    void* set = nullptr;
    bool periodic = false;  // the restraint distances take the (orthorhombic) minimum image
};

}  // namespace ATMMetaForcePlugin

#endif
