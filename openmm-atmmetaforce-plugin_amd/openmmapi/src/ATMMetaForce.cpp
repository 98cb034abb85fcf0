// ATMMetaForce particle table and context accessors
// (reference: openmmapi/src/ATMMetaForce.cpp:12-44).
#include "ATMMetaForce.h"
#include "internal/ATMMetaForceImpl.h"
#include "openmm/OpenMMException.h"
#include "openmm/internal/AssertionUtilities.h"

using namespace ATMMetaForcePlugin;

int ATMMetaForce::addParticle(int particle, double dx, double dy, double dz) {
    particles.emplace_back(particle, dx, dy, dz);
    return (int)particles.size() - 1;
}

void ATMMetaForce::getParticleParameters(int index, int& particle, double& dx, double& dy, double& dz) const {
    ASSERT_VALID_INDEX(index, particles);
    const ParticleInfo& p = particles[index];
    particle = p.particle;
    dx = p.dx;
    dy = p.dy;
    dz = p.dz;
}

void ATMMetaForce::setParticleParameters(int index, int particle, double dx, double dy, double dz) {
    ASSERT_VALID_INDEX(index, particles);
    particles[index] = ParticleInfo(particle, dx, dy, dz);
}

OpenMM::ForceImpl* ATMMetaForce::createImpl() const { return new ATMMetaForceImpl(*this); }

void ATMMetaForce::updateParametersInContext(OpenMM::Context& context) {
    dynamic_cast<ATMMetaForceImpl&>(getImplInContext(context)).updateParametersInContext(getContextImpl(context));
}

double ATMMetaForce::getPerturbationEnergy(const OpenMM::Context& context) const {
    return dynamic_cast<const ATMMetaForceImpl&>(getImplInContext(context)).getPerturbationEnergy();
}
