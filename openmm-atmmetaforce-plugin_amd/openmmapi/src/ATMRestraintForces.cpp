// Flat-bottom restraint forces and their per-context implementations.
#include "ATMRestraintForces.h"

#include <cmath>
#include <string>
#include "ATMRestraintKernels.h"
#include "openmm/Kernel.h"
#include "openmm/OpenMMException.h"
#include "openmm/Platform.h"
#include "openmm/System.h"
#include "openmm/internal/ContextImpl.h"
#include "openmm/internal/ForceImpl.h"

using namespace OpenMM;

namespace ATMMetaForcePlugin {

namespace {

void checkIndex(int index, int n, const char* what) {
    if (index < 0 || index >= n) throw OpenMMException(std::string(what) + ": index out of range");
}

template <class F, class K>
class RestraintImpl : public ForceImpl {
public:
    explicit RestraintImpl(const F& owner) : owner(owner) {}
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

using CentroidImpl = RestraintImpl<CentroidRestraintForce, CalcCentroidRestraintForceKernel>;
using PositionImpl = RestraintImpl<PositionRestraintForce, CalcPositionRestraintForceKernel>;

}  // namespace

// ------------------------------------------------------------------ CentroidRestraintForce

int CentroidRestraintForce::addGroup(const std::vector<int>& particles, const std::vector<double>& weights) {
    if (particles.empty()) throw OpenMMException("CentroidRestraintForce: a group needs at least one particle");
    if (!weights.empty() && weights.size() != particles.size())
        throw OpenMMException("CentroidRestraintForce: one weight per particle (or none, for masses)");
    for (double w : weights)
        if (!(w > 0.0)) throw OpenMMException("CentroidRestraintForce: weights must be positive");
    groups.push_back(Group{particles, weights});
    return (int)groups.size() - 1;
}

void CentroidRestraintForce::getGroupParameters(int index, std::vector<int>& particles, std::vector<double>& weights) const {
    checkIndex(index, (int)groups.size(), "CentroidRestraintForce::getGroupParameters");
    particles = groups[index].particles;
    weights = groups[index].weights;
}

int CentroidRestraintForce::addBond(int group1, int group2, double kf, double tol, const Vec3& offset) {
    bonds.push_back(Bond{group1, group2, kf, tol, offset});
    return (int)bonds.size() - 1;
}

void CentroidRestraintForce::getBondParameters(int index, int& group1, int& group2, double& kf, double& tol,
                                               Vec3& offset) const {
    checkIndex(index, (int)bonds.size(), "CentroidRestraintForce::getBondParameters");
    const Bond& b = bonds[index];
    group1 = b.g1; group2 = b.g2; kf = b.kf; tol = b.tol; offset = b.offset;
}

void CentroidRestraintForce::setBondParameters(int index, int group1, int group2, double kf, double tol,
                                               const Vec3& offset) {
    checkIndex(index, (int)bonds.size(), "CentroidRestraintForce::setBondParameters");
    bonds[index] = Bond{group1, group2, kf, tol, offset};
}

void CentroidRestraintForce::updateParametersInContext(Context& context) {
    dynamic_cast<CentroidImpl&>(getImplInContext(context)).updateParametersInContext(getContextImpl(context));
}

ForceImpl* CentroidRestraintForce::createImpl() const { return new CentroidImpl(*this); }

// ------------------------------------------------------------------ PositionRestraintForce

int PositionRestraintForce::addParticle(int particle, const Vec3& ref, double kf, double tol) {
    entries.push_back(Entry{particle, ref, kf, tol});
    return (int)entries.size() - 1;
}

void PositionRestraintForce::getParticleParameters(int index, int& particle, Vec3& ref, double& kf, double& tol) const {
    checkIndex(index, (int)entries.size(), "PositionRestraintForce::getParticleParameters");
    const Entry& e = entries[index];
    particle = e.particle; ref = e.ref; kf = e.kf; tol = e.tol;
}

void PositionRestraintForce::setParticleParameters(int index, int particle, const Vec3& ref, double kf, double tol) {
    checkIndex(index, (int)entries.size(), "PositionRestraintForce::setParticleParameters");
    entries[index] = Entry{particle, ref, kf, tol};
}

void PositionRestraintForce::updateParametersInContext(Context& context) {
    dynamic_cast<PositionImpl&>(getImplInContext(context)).updateParametersInContext(getContextImpl(context));
}

ForceImpl* PositionRestraintForce::createImpl() const { return new PositionImpl(*this); }

}  // namespace ATMMetaForcePlugin
