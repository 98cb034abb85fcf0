// Minimal OpenMM-compatible core: HarmonicBondForce, HarmonicAngleForce and
// PeriodicTorsionForce (parameter storage and the ForceImpl that calls the
// platform's kernel).
#include "openmm/Context.h"
#include "openmm/HarmonicAngleForce.h"
#include "openmm/HarmonicBondForce.h"
#include "openmm/OpenMMException.h"
#include "openmm/PeriodicTorsionForce.h"
#include "openmm/Platform.h"
#include "openmm/internal/ContextImpl.h"
#include "openmm/internal/ForceImpl.h"
#include "openmm/kernels.h"

namespace OpenMM {

namespace {

// ForceImpl of a bonded force F evaluated by kernel K.
template <class F, class K>
class BondedForceImpl : public ForceImpl {
public:
    explicit BondedForceImpl(const F& owner) : owner(owner) {}
    void initialize(ContextImpl& context) override {
        kernel = context.getPlatform().createKernel(K::Name(), context);
        kernel.getAs<K>().initialize(context.getSystem(), owner);
    }
    const Force& getOwner() const override { return owner; }
    void updateContextState(ContextImpl&, bool&) override {}
    double calcForcesAndEnergy(ContextImpl& context, bool includeForces, bool includeEnergy, int groups) override {
        if ((groups & (1 << owner.getForceGroup())) == 0) return 0.0;
        return kernel.getAs<K>().execute(context, includeForces, includeEnergy);
    }
    std::map<std::string, double> getDefaultParameters() override { return {}; }
    std::vector<std::string> getKernelNames() override { return {K::Name()}; }
    void updateParametersInContext(ContextImpl& context) { kernel.getAs<K>().copyParametersToContext(context, owner); }

private:
    const F& owner;
    Kernel kernel;
};

void checkIndex(int index, size_t n) {
    if (index < 0 || index >= (int)n) throw OpenMMException("Index out of range");
}

void checkNotPeriodic(bool periodic) {
    if (periodic) throw OpenMMException("Periodic bonded terms are not supported (molecules must be whole)");
}

}  // namespace

// ------------------------------------------------------------------ HarmonicBondForce

using BondImpl = BondedForceImpl<HarmonicBondForce, CalcHarmonicBondForceKernel>;

int HarmonicBondForce::addBond(int p1, int p2, double length, double k) {
    bonds.push_back(Bond{p1, p2, length, k});
    return (int)bonds.size() - 1;
}

void HarmonicBondForce::getBondParameters(int index, int& p1, int& p2, double& length, double& k) const {
    checkIndex(index, bonds.size());
    const Bond& b = bonds[index];
    p1 = b.p1; p2 = b.p2; length = b.length; k = b.k;
}

void HarmonicBondForce::setBondParameters(int index, int p1, int p2, double length, double k) {
    checkIndex(index, bonds.size());
    bonds[index] = Bond{p1, p2, length, k};
}

void HarmonicBondForce::setUsesPeriodicBoundaryConditions(bool periodic) { checkNotPeriodic(periodic); }

void HarmonicBondForce::updateParametersInContext(Context& context) {
    static_cast<BondImpl&>(getImplInContext(context)).updateParametersInContext(getContextImpl(context));
}

ForceImpl* HarmonicBondForce::createImpl() const { return new BondImpl(*this); }

// ------------------------------------------------------------------ HarmonicAngleForce

using AngleImpl = BondedForceImpl<HarmonicAngleForce, CalcHarmonicAngleForceKernel>;

int HarmonicAngleForce::addAngle(int p1, int p2, int p3, double angle, double k) {
    angles.push_back(Angle{p1, p2, p3, angle, k});
    return (int)angles.size() - 1;
}

void HarmonicAngleForce::getAngleParameters(int index, int& p1, int& p2, int& p3, double& angle, double& k) const {
    checkIndex(index, angles.size());
    const Angle& a = angles[index];
    p1 = a.p1; p2 = a.p2; p3 = a.p3; angle = a.angle; k = a.k;
}

void HarmonicAngleForce::setAngleParameters(int index, int p1, int p2, int p3, double angle, double k) {
    checkIndex(index, angles.size());
    angles[index] = Angle{p1, p2, p3, angle, k};
}

void HarmonicAngleForce::setUsesPeriodicBoundaryConditions(bool periodic) { checkNotPeriodic(periodic); }

void HarmonicAngleForce::updateParametersInContext(Context& context) {
    static_cast<AngleImpl&>(getImplInContext(context)).updateParametersInContext(getContextImpl(context));
}

ForceImpl* HarmonicAngleForce::createImpl() const { return new AngleImpl(*this); }

// ------------------------------------------------------------------ PeriodicTorsionForce

using TorsionImpl = BondedForceImpl<PeriodicTorsionForce, CalcPeriodicTorsionForceKernel>;

int PeriodicTorsionForce::addTorsion(int p1, int p2, int p3, int p4, int periodicity, double phase, double k) {
    torsions.push_back(Torsion{p1, p2, p3, p4, periodicity, phase, k});
    return (int)torsions.size() - 1;
}

void PeriodicTorsionForce::getTorsionParameters(int index, int& p1, int& p2, int& p3, int& p4, int& periodicity,
                                                double& phase, double& k) const {
    checkIndex(index, torsions.size());
    const Torsion& t = torsions[index];
    p1 = t.p1; p2 = t.p2; p3 = t.p3; p4 = t.p4; periodicity = t.periodicity; phase = t.phase; k = t.k;
}

void PeriodicTorsionForce::setTorsionParameters(int index, int p1, int p2, int p3, int p4, int periodicity,
                                                double phase, double k) {
    checkIndex(index, torsions.size());
    torsions[index] = Torsion{p1, p2, p3, p4, periodicity, phase, k};
}

void PeriodicTorsionForce::setUsesPeriodicBoundaryConditions(bool periodic) { checkNotPeriodic(periodic); }

void PeriodicTorsionForce::updateParametersInContext(Context& context) {
    static_cast<TorsionImpl&>(getImplInContext(context)).updateParametersInContext(getContextImpl(context));
}

ForceImpl* PeriodicTorsionForce::createImpl() const { return new TorsionImpl(*this); }

}  // namespace OpenMM
