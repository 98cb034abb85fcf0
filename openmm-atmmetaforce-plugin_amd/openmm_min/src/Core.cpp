// Minimal OpenMM-compatible core: Force, System, NonbondedForce, State,
// Integrator, Context and ContextImpl.
#include <algorithm>
#include <cmath>
#include <random>
#include <cstdlib>
#include <set>
#include <sstream>

#include "openmm/Context.h"
#include "openmm/Force.h"
#include "openmm/Integrator.h"
#include "openmm/MonteCarloBarostat.h"
#include "openmm/NonbondedForce.h"
#include "openmm/OpenMMException.h"
#include "openmm/Platform.h"
#include "openmm/State.h"
#include "openmm/System.h"
#include "openmm/internal/ContextImpl.h"
#include "openmm/internal/ForceImpl.h"
#include "openmm/kernels.h"

namespace OpenMM {

// ------------------------------------------------------------------ Force

void Force::setForceGroup(int group) {
    if (group < 0 || group > 31) throw OpenMMException("Force group must be between 0 and 31");
    forceGroup = group;
}

const std::string& Force::getName() const {
    if (!name.empty()) return name;
    defaultNameCache = defaultName();
    return defaultNameCache;
}

ForceImpl& Force::getImplInContext(Context& context) {
    for (ForceImpl* impl : context.getImpl().getForceImpls())
        if (&impl->getOwner() == this) return *impl;
    throw OpenMMException("getImplInContext: This Force is not present in the Context");
}

const ForceImpl& Force::getImplInContext(const Context& context) const {
    for (ForceImpl* impl : context.getImpl().getForceImpls())
        if (&impl->getOwner() == this) return *impl;
    throw OpenMMException("getImplInContext: This Force is not present in the Context");
}

ContextImpl& Force::getContextImpl(Context& context) { return context.getImpl(); }

// ------------------------------------------------------------------ System

System::System() {
    box[0] = Vec3(2, 0, 0);
    box[1] = Vec3(0, 2, 0);
    box[2] = Vec3(0, 0, 2);
}

System::~System() {
    for (Force* f : forces) delete f;
}

double System::getParticleMass(int index) const {
    if (index < 0 || index >= (int)masses.size()) throw OpenMMException("Index out of range");
    return masses[index];
}

void System::setParticleMass(int index, double mass) {
    if (index < 0 || index >= (int)masses.size()) throw OpenMMException("Index out of range");
    masses[index] = mass;
}

int System::addConstraint(int particle1, int particle2, double distance) {
    constraints.push_back(Constraint{particle1, particle2, distance});
    return (int)constraints.size() - 1;
}

void System::getConstraintParameters(int index, int& particle1, int& particle2, double& distance) const {
    if (index < 0 || index >= (int)constraints.size()) throw OpenMMException("Index out of range");
    particle1 = constraints[index].p1;
    particle2 = constraints[index].p2;
    distance = constraints[index].d;
}

int System::addForce(Force* force) {
    forces.push_back(force);
    return (int)forces.size() - 1;
}

Force& System::getForce(int index) {
    if (index < 0 || index >= (int)forces.size()) throw OpenMMException("Index out of range");
    return *forces[index];
}

const Force& System::getForce(int index) const {
    if (index < 0 || index >= (int)forces.size()) throw OpenMMException("Index out of range");
    return *forces[index];
}

void System::removeForce(int index) {
    if (index < 0 || index >= (int)forces.size()) throw OpenMMException("Index out of range");
    delete forces[index];
    forces.erase(forces.begin() + index);
}

void System::getDefaultPeriodicBoxVectors(Vec3& a, Vec3& b, Vec3& c) const {
    a = box[0];
    b = box[1];
    c = box[2];
}

// OpenMM's reduced form of periodic box vectors: a along x, b in the x-y plane,
// and each vector's tilt at most half the edge of the ones before it (triclinic
// boxes included).
static void checkReducedBox(const Vec3& a, const Vec3& b, const Vec3& c) {
    if (a[1] != 0.0 || a[2] != 0.0) throw OpenMMException("First periodic box vector must be parallel to x.");
    if (b[2] != 0.0) throw OpenMMException("Second periodic box vector must be in the x-y plane.");
    if (!(a[0] > 0 && b[1] > 0 && c[2] > 0)) throw OpenMMException("Periodic box vectors must be positive");
    if (a[0] < 2 * std::fabs(b[0]) || a[0] < 2 * std::fabs(c[0]) || b[1] < 2 * std::fabs(c[1]))
        throw OpenMMException("Periodic box vectors must be in reduced form.");
}

void System::setDefaultPeriodicBoxVectors(const Vec3& a, const Vec3& b, const Vec3& c) {
    checkReducedBox(a, b, c);
    box[0] = a;
    box[1] = b;
    box[2] = c;
}

bool System::usesPeriodicBoundaryConditions() const {
    for (Force* f : forces)
        if (f->usesPeriodicBoundaryConditions()) return true;
    return false;
}

// ------------------------------------------------------------------ NonbondedForce

class NonbondedForceImpl : public ForceImpl {
public:
    explicit NonbondedForceImpl(const NonbondedForce& owner) : owner(owner) {}
    void initialize(ContextImpl& context) override {
        if (owner.getNumParticles() != context.getSystem().getNumParticles())
            throw OpenMMException("NonbondedForce must have exactly as many particles as the System it belongs to.");
        kernel = context.getPlatform().createKernel(CalcNonbondedForceKernel::Name(), context);
        kernel.getAs<CalcNonbondedForceKernel>().initialize(context.getSystem(), owner);
    }
    const Force& getOwner() const override { return owner; }
    void updateContextState(ContextImpl&, bool&) override {}
    double calcForcesAndEnergy(ContextImpl& context, bool includeForces, bool includeEnergy, int groups) override {
        if ((groups & (1 << owner.getForceGroup())) == 0) return 0.0;
        return kernel.getAs<CalcNonbondedForceKernel>().execute(context, includeForces, includeEnergy);
    }
    std::map<std::string, double> getDefaultParameters() override { return {}; }
    std::vector<std::string> getKernelNames() override { return {CalcNonbondedForceKernel::Name()}; }
    void updateParametersInContext(ContextImpl& context) {
        kernel.getAs<CalcNonbondedForceKernel>().copyParametersToContext(context, owner);
    }

private:
    const NonbondedForce& owner;
    Kernel kernel;
};

// ------------------------------------------------------------------ MonteCarloBarostat

class MonteCarloBarostatImpl : public ForceImpl {
public:
    explicit MonteCarloBarostatImpl(const MonteCarloBarostat& owner) : owner(owner) {}
    void initialize(ContextImpl&) override { step = 0; }
    const Force& getOwner() const override { return owner; }
    void updateContextState(ContextImpl&, bool&) override {
        if (owner.getFrequency() == 0 || ++step < owner.getFrequency()) return;
        throw OpenMMException("MonteCarloBarostat: volume moves are not implemented here; keep its frequency above "
                              "the number of steps (or 0) to use it as a holder of its parameters");
    }
    double calcForcesAndEnergy(ContextImpl&, bool, bool, int) override { return 0.0; }
    std::map<std::string, double> getDefaultParameters() override {
        return {{MonteCarloBarostat::Pressure(), owner.getDefaultPressure()},
                {MonteCarloBarostat::Temperature(), owner.getDefaultTemperature()}};
    }
    std::vector<std::string> getKernelNames() override { return {}; }

private:
    const MonteCarloBarostat& owner;
    long long step = 0;
};

MonteCarloBarostat::MonteCarloBarostat(double defaultPressure, double defaultTemperature, int frequency)
    : pressure(defaultPressure), temperature(defaultTemperature), frequency(0), seed(0) {
    setFrequency(frequency);
}

void MonteCarloBarostat::setFrequency(int freq) {
    if (freq < 0) throw OpenMMException("Frequency cannot be negative");
    frequency = freq;
}

ForceImpl* MonteCarloBarostat::createImpl() const { return new MonteCarloBarostatImpl(*this); }

NonbondedForce::NonbondedForce()
    : method(NoCutoff), cutoff(1.0), rfDielectric(78.3), ewaldTolerance(5e-4), pmeAlpha(0.0), pmeGrid{0, 0, 0},
      useDispersionCorrection(true), recipGroup(-1) {}

void NonbondedForce::setEwaldErrorTolerance(double tol) { ewaldTolerance = tol; }

void NonbondedForce::getPMEParameters(double& alpha, int& nx, int& ny, int& nz) const {
    alpha = pmeAlpha;
    nx = pmeGrid[0];
    ny = pmeGrid[1];
    nz = pmeGrid[2];
}

void NonbondedForce::setPMEParameters(double alpha, int nx, int ny, int nz) {
    pmeAlpha = alpha;
    pmeGrid[0] = nx;
    pmeGrid[1] = ny;
    pmeGrid[2] = nz;
}

static unsigned long long pairKey(int a, int b) {
    const unsigned lo = (unsigned)std::min(a, b), hi = (unsigned)std::max(a, b);
    return ((unsigned long long)lo << 32) | hi;
}

void NonbondedForce::setNonbondedMethod(NonbondedMethod m) {
    if (m < NoCutoff || m > LJPME) throw OpenMMException("NonbondedForce: Illegal value for nonbonded method");
    method = m;
}

void NonbondedForce::setCutoffDistance(double d) { cutoff = d; }
void NonbondedForce::setReactionFieldDielectric(double d) { rfDielectric = d; }

void NonbondedForce::setReciprocalSpaceForceGroup(int group) {
    if (group < -1 || group > 31) throw OpenMMException("Force group must be between -1 and 31");
    recipGroup = group;
}

int NonbondedForce::addParticle(double charge, double sigma, double epsilon) {
    particles.push_back(Particle{charge, sigma, epsilon});
    return (int)particles.size() - 1;
}

void NonbondedForce::getParticleParameters(int index, double& charge, double& sigma, double& epsilon) const {
    if (index < 0 || index >= (int)particles.size()) throw OpenMMException("Index out of range");
    charge = particles[index].q;
    sigma = particles[index].sigma;
    epsilon = particles[index].eps;
}

void NonbondedForce::setParticleParameters(int index, double charge, double sigma, double epsilon) {
    if (index < 0 || index >= (int)particles.size()) throw OpenMMException("Index out of range");
    particles[index] = Particle{charge, sigma, epsilon};
}

int NonbondedForce::addException(int p1, int p2, double qq, double sigma, double eps, bool replace) {
    const unsigned long long key = pairKey(p1, p2);
    auto it = exceptionIndex.find(key);
    if (it != exceptionIndex.end()) {
        if (!replace) throw OpenMMException("NonbondedForce: There is already an exception for those particles");
        exceptions[it->second] = Exception{p1, p2, qq, sigma, eps};
        return it->second;
    }
    exceptions.push_back(Exception{p1, p2, qq, sigma, eps});
    exceptionIndex[key] = (int)exceptions.size() - 1;
    return (int)exceptions.size() - 1;
}

void NonbondedForce::setExceptionParameters(int index, int p1, int p2, double qq, double sigma, double eps) {
    if (index < 0 || index >= (int)exceptions.size()) throw OpenMMException("Index out of range");
    Exception& e = exceptions[index];
    if (pairKey(e.p1, e.p2) != pairKey(p1, p2)) {
        if (exceptionIndex.count(pairKey(p1, p2)))
            throw OpenMMException("NonbondedForce: There is already an exception for those particles");
        exceptionIndex.erase(pairKey(e.p1, e.p2));
        exceptionIndex[pairKey(p1, p2)] = index;
    }
    e = Exception{p1, p2, qq, sigma, eps};
}

void NonbondedForce::getExceptionParameters(int index, int& p1, int& p2, double& qq, double& sigma, double& eps) const {
    if (index < 0 || index >= (int)exceptions.size()) throw OpenMMException("Index out of range");
    const Exception& e = exceptions[index];
    p1 = e.p1;
    p2 = e.p2;
    qq = e.qq;
    sigma = e.sigma;
    eps = e.eps;
}

void NonbondedForce::updateParametersInContext(Context& context) {
    dynamic_cast<NonbondedForceImpl&>(getImplInContext(context)).updateParametersInContext(getContextImpl(context));
}

ForceImpl* NonbondedForce::createImpl() const { return new NonbondedForceImpl(*this); }

// ------------------------------------------------------------------ State

const std::vector<Vec3>& State::getPositions() const {
    if (!(types & Positions)) throw OpenMMException("Invoked getPositions() on a State which does not contain positions.");
    return positions;
}
const std::vector<Vec3>& State::getVelocities() const {
    if (!(types & Velocities)) throw OpenMMException("Invoked getVelocities() on a State which does not contain velocities.");
    return velocities;
}
const std::vector<Vec3>& State::getForces() const {
    if (!(types & Forces)) throw OpenMMException("Invoked getForces() on a State which does not contain forces.");
    return forces;
}
double State::getPotentialEnergy() const {
    if (!(types & Energy)) throw OpenMMException("Invoked getPotentialEnergy() on a State which does not contain energies.");
    return potentialEnergy;
}
double State::getKineticEnergy() const {
    if (!(types & Energy)) throw OpenMMException("Invoked getKineticEnergy() on a State which does not contain energies.");
    return kineticEnergy;
}
const std::map<std::string, double>& State::getParameters() const {
    if (!(types & Parameters)) throw OpenMMException("Invoked getParameters() on a State which does not contain parameters.");
    return parameters;
}

// ------------------------------------------------------------------ Integrator

// LangevinMiddleIntegrator does not apply constraints: a System that holds any
// is refused rather than integrated with unconstrained bonds.
static void requireNoConstraints(const System& sys) {
    if (sys.getNumConstraints() > 0)
        throw OpenMMException("This integrator does not apply constraints; the System has " +
                              std::to_string(sys.getNumConstraints()) + " (use LangevinIntegrator or VerletIntegrator)");
}

// SHAKE (Ryckaert, Ciccotti, Berendsen, J. Comput. Phys. 23, 327 (1977)): move
// the new positions xn along the old bond vectors until every constraint
// |xn_i - xn_j| = d holds to `tol` (relative).  Massless particles do not
// move; a constraint between two of them is left alone.
static void shake(const System& sys, const std::vector<Vec3>& xo, std::vector<Vec3>& xn, double tol) {
    const int nc = sys.getNumConstraints();
    if (nc == 0) return;
    std::vector<int> p1(nc), p2(nc);
    std::vector<double> d2(nc), w1(nc), w2(nc);
    std::vector<Vec3> ref(nc);
    for (int k = 0; k < nc; ++k) {
        double d;
        sys.getConstraintParameters(k, p1[k], p2[k], d);
        d2[k] = d * d;
        const double m1 = sys.getParticleMass(p1[k]), m2 = sys.getParticleMass(p2[k]);
        w1[k] = m1 > 0.0 ? 1.0 / m1 : 0.0;
        w2[k] = m2 > 0.0 ? 1.0 / m2 : 0.0;
        ref[k] = xo[p1[k]] - xo[p2[k]];
    }
    for (int it = 0; it < 1000; ++it) {
        bool done = true;
        for (int k = 0; k < nc; ++k) {
            if (w1[k] + w2[k] == 0.0) continue;
            const Vec3 sv = xn[p1[k]] - xn[p2[k]];
            const double diff = d2[k] - sv.dot(sv);
            if (std::fabs(diff) <= 2.0 * tol * d2[k]) continue;
            done = false;
            double rs = ref[k].dot(sv);
            if (rs < 1e-6 * d2[k]) rs = 1e-6 * d2[k];  // bond turned past 90 degrees: keep the step bounded
            const double g = diff / (2.0 * rs * (w1[k] + w2[k]));
            xn[p1[k]] += ref[k] * (g * w1[k]);
            xn[p2[k]] -= ref[k] * (g * w2[k]);
        }
        if (done) return;
    }
    throw OpenMMException("SHAKE: constraints did not converge in 1000 iterations");
}

void Integrator::setIntegrationForceGroups(const std::set<int>& s) {
    int mask = 0;
    for (int g : s) {
        if (g < 0 || g > 31) throw OpenMMException("Force group must be between 0 and 31");
        mask |= 1 << g;
    }
    groups = mask;
}

void VerletIntegrator::step(int steps) {
    if (!context) throw OpenMMException("This Integrator is not bound to a context!");
    ContextImpl& c = *context;
    const System& sys = c.getSystem();
    const int n = sys.getNumParticles();
    const bool constrained = sys.getNumConstraints() > 0;
    std::vector<Vec3>& x = c.positionsRef();
    std::vector<Vec3>& v = c.velocitiesRef();
    std::vector<Vec3> xn;
    for (int s = 0; s < steps; ++s) {
        c.updateContextState();
        c.calcForcesAndEnergy(true, false, groups);
        const std::vector<Vec3>& f = c.forcesRef();
        xn = x;
        for (int i = 0; i < n; ++i) {
            const double m = sys.getParticleMass(i);
            if (m == 0.0) continue;
            v[i] += f[i] * (stepSize / m);
            xn[i] += v[i] * stepSize;
        }
        if (constrained) {
            shake(sys, x, xn, constraintTolerance);
            for (int i = 0; i < n; ++i)
                if (sys.getParticleMass(i) != 0.0) v[i] = (xn[i] - x[i]) * (1.0 / stepSize);
        }
        x.swap(xn);
        c.setTime(c.getTime() + stepSize);
    }
}

// ------------------------------------------------------------------ Langevin integrators

namespace {
const double kBoltz = 0.00831446261815324;  // kJ/mol/K

// N(0, 1) deviates of step `k` of an integrator seeded with `seed`: a generator
// per step, so a trajectory does not depend on how steps are grouped into calls.
struct StepNormals {
    std::mt19937_64 gen;
    std::normal_distribution<double> dist{0.0, 1.0};
    StepNormals(int seed, unsigned long long k) : gen((unsigned long long)seed * 0x9E3779B97F4A7C15ull + k) {}
    Vec3 operator()() {
        const double a = dist(gen), b = dist(gen), c = dist(gen);
        return Vec3(a, b, c);
    }
};
}  // namespace

void LangevinIntegrator::step(int steps) {
    if (!context) throw OpenMMException("This Integrator is not bound to a context!");
    ContextImpl& c = *context;
    const System& sys = c.getSystem();
    const int n = sys.getNumParticles();
    const bool constrained = sys.getNumConstraints() > 0;
    std::vector<Vec3>& x = c.positionsRef();
    std::vector<Vec3>& v = c.velocitiesRef();
    const double dt = stepSize, a = std::exp(-friction * dt), kT = kBoltz * temperature;
    const double fscale = friction > 0.0 ? (1.0 - a) / friction : dt;
    std::vector<Vec3> xn;
    for (int s = 0; s < steps; ++s) {
        StepNormals R(seed, drawn++);
        c.updateContextState();
        c.calcForcesAndEnergy(true, false, groups);
        const std::vector<Vec3>& f = c.forcesRef();
        xn = x;
        for (int i = 0; i < n; ++i) {
            const double m = sys.getParticleMass(i);
            if (m == 0.0) continue;
            v[i] = v[i] * a + f[i] * (fscale / m) + R() * std::sqrt(kT * (1.0 - a * a) / m);
            xn[i] += v[i] * dt;
        }
        if (constrained) {
            shake(sys, x, xn, constraintTolerance);
            for (int i = 0; i < n; ++i)
                if (sys.getParticleMass(i) != 0.0) v[i] = (xn[i] - x[i]) * (1.0 / dt);
        }
        x.swap(xn);
        c.setTime(c.getTime() + dt);
    }
}

void LangevinMiddleIntegrator::step(int steps) {
    if (!context) throw OpenMMException("This Integrator is not bound to a context!");
    ContextImpl& c = *context;
    const System& sys = c.getSystem();
    requireNoConstraints(sys);
    const int n = sys.getNumParticles();
    std::vector<Vec3>& x = c.positionsRef();
    std::vector<Vec3>& v = c.velocitiesRef();
    const double dt = stepSize, a = std::exp(-friction * dt), kT = kBoltz * temperature;
    for (int s = 0; s < steps; ++s) {
        StepNormals R(seed, drawn++);
        c.updateContextState();
        c.calcForcesAndEnergy(true, false, groups);
        const std::vector<Vec3>& f = c.forcesRef();
        for (int i = 0; i < n; ++i) {
            const double m = sys.getParticleMass(i);
            if (m == 0.0) continue;
            v[i] += f[i] * (dt / m);
            x[i] += v[i] * (0.5 * dt);
            v[i] = v[i] * a + R() * std::sqrt(kT * (1.0 - a * a) / m);
            x[i] += v[i] * (0.5 * dt);
        }
        c.setTime(c.getTime() + dt);
    }
}

// ------------------------------------------------------------------ ContextImpl

ContextImpl::ContextImpl(Context& owner, const System& system, Integrator& integrator, Platform* platform,
                         const std::map<std::string, std::string>& properties)
    : owner(owner), system(system), integrator(integrator), platform(platform), properties(properties), time(0.0),
      platformData(nullptr) {
    const int n = system.getNumParticles();
    if (n == 0) throw OpenMMException("Cannot create a Context for a System with no particles");
    positions.assign(n, Vec3());
    velocities.assign(n, Vec3());
    forces.assign(n, Vec3());
    system.getDefaultPeriodicBoxVectors(box[0], box[1], box[2]);
    platform->contextCreated(*this, properties);
    try {
        for (int i = 0; i < system.getNumForces(); ++i) forceImpls.push_back(system.getForce(i).createImpl());
        for (ForceImpl* f : forceImpls)
            for (auto& kv : f->getDefaultParameters()) parameters[kv.first] = kv.second;
        for (ForceImpl* f : forceImpls) f->initialize(*this);
        integrator.initialize(*this);
    } catch (...) {
        for (ForceImpl* f : forceImpls) delete f;
        forceImpls.clear();
        platform->contextDestroyed(*this);
        throw;
    }
}

ContextImpl::~ContextImpl() {
    for (ForceImpl* f : forceImpls) delete f;
    platform->contextDestroyed(*this);
}

void ContextImpl::setPositions(const std::vector<Vec3>& p) {
    if (p.size() != positions.size())
        throw OpenMMException("Called setPositions() on a Context with the wrong number of positions");
    positions = p;
}

void ContextImpl::setVelocities(const std::vector<Vec3>& v) {
    if (v.size() != velocities.size())
        throw OpenMMException("Called setVelocities() on a Context with the wrong number of velocities");
    velocities = v;
}

double ContextImpl::getParameter(const std::string& name) const {
    auto it = parameters.find(name);
    if (it == parameters.end()) throw OpenMMException("Called getParameter() with invalid parameter name: " + name);
    return it->second;
}

void ContextImpl::setParameter(const std::string& name, double value) {
    auto it = parameters.find(name);
    if (it == parameters.end()) throw OpenMMException("Called setParameter() with invalid parameter name: " + name);
    it->second = value;
}

void ContextImpl::setPeriodicBoxVectors(const Vec3& a, const Vec3& b, const Vec3& c) {
    checkReducedBox(a, b, c);
    box[0] = a;
    box[1] = b;
    box[2] = c;
}

void ContextImpl::updateContextState() {
    bool forcesInvalid = false;
    for (ForceImpl* f : forceImpls) f->updateContextState(*this, forcesInvalid);
}

double ContextImpl::calcForcesAndEnergy(bool includeForces, bool includeEnergy, int groups) {
    for (Vec3& f : forces) f = Vec3();
    double energy = 0.0;
    for (ForceImpl* f : forceImpls) energy += f->calcForcesAndEnergy(*this, includeForces, includeEnergy, groups);
    return energy;
}

double ContextImpl::calcKineticEnergy() const {
    double k = 0.0;
    for (int i = 0; i < system.getNumParticles(); ++i) k += 0.5 * system.getParticleMass(i) * velocities[i].dot(velocities[i]);
    return k;
}

// ------------------------------------------------------------------ Context

static Platform& defaultPlatform() {
    if (Platform::getNumPlatforms() == 0) throw OpenMMException("No Platform is registered");
    Platform* best = &Platform::getPlatform(0);
    for (int i = 1; i < Platform::getNumPlatforms(); ++i)
        if (Platform::getPlatform(i).getSpeed() > best->getSpeed()) best = &Platform::getPlatform(i);
    return *best;
}

Context::Context(const System& system, Integrator& integrator) : impl(nullptr) {
    impl = new ContextImpl(*this, system, integrator, &defaultPlatform(), {});
}

Context::Context(const System& system, Integrator& integrator, Platform& platform) : impl(nullptr) {
    impl = new ContextImpl(*this, system, integrator, &platform, {});
}

Context::Context(const System& system, Integrator& integrator, Platform& platform,
                 const std::map<std::string, std::string>& properties)
    : impl(nullptr) {
    impl = new ContextImpl(*this, system, integrator, &platform, properties);
}

Context::~Context() { delete impl; }

const System& Context::getSystem() const { return impl->getSystem(); }
Integrator& Context::getIntegrator() { return impl->getIntegrator(); }
Platform& Context::getPlatform() { return impl->getPlatform(); }

State Context::getState(int types, bool enforcePeriodicBox, int groups) const {
    State s;
    s.types = types;
    s.time = impl->getTime();
    impl->getPeriodicBoxVectors(s.box[0], s.box[1], s.box[2]);
    if (types & (State::Forces | State::Energy)) {
        const double e = impl->calcForcesAndEnergy((types & State::Forces) != 0, (types & State::Energy) != 0, groups);
        if (types & State::Energy) {
            s.potentialEnergy = e;
            s.kineticEnergy = impl->calcKineticEnergy();
        }
        if (types & State::Forces) impl->getForces(s.forces);
    }
    if (types & State::Parameters) s.parameters = impl->getParameters();
    if (types & State::Positions) {
        impl->getPositions(s.positions);
        if (enforcePeriodicBox)
            for (Vec3& p : s.positions)
                for (int k = 0; k < 3; ++k) p[k] -= s.box[k][k] * std::floor(p[k] / s.box[k][k]);
    }
    if (types & State::Velocities) impl->getVelocities(s.velocities);
    return s;
}

State Context::getState(int types, bool enforcePeriodicBox, const std::set<int>& groups) const {
    int mask = 0;
    for (int g : groups) {
        if (g < 0 || g > 31) throw OpenMMException("Force group must be between 0 and 31");
        mask |= 1 << g;
    }
    return getState(types, enforcePeriodicBox, mask);
}

void Context::setTime(double t) { impl->setTime(t); }
void Context::setPositions(const std::vector<Vec3>& positions) { impl->setPositions(positions); }
void Context::setVelocities(const std::vector<Vec3>& velocities) { impl->setVelocities(velocities); }
double Context::getParameter(const std::string& name) const { return impl->getParameter(name); }
void Context::setParameter(const std::string& name, double value) { impl->setParameter(name, value); }
const std::map<std::string, double>& Context::getParameters() const { return impl->getParameters(); }
void Context::setPeriodicBoxVectors(const Vec3& a, const Vec3& b, const Vec3& c) { impl->setPeriodicBoxVectors(a, b, c); }

void Context::setState(const State& state) {
    setTime(state.getTime());
    Vec3 a, b, c;
    state.getPeriodicBoxVectors(a, b, c);
    setPeriodicBoxVectors(a, b, c);
    if (state.getDataTypes() & State::Positions) setPositions(state.getPositions());
    if (state.getDataTypes() & State::Velocities) setVelocities(state.getVelocities());
    if (state.getDataTypes() & State::Parameters)
        for (auto& kv : state.getParameters()) setParameter(kv.first, kv.second);
}

}  // namespace OpenMM
