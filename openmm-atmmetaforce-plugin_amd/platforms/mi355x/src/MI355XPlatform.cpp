// MI355X platform, its kernel factories and the plugin entry points.
#include <cmath>
#include <cstdlib>
#include <map>
#include <string>
#include <vector>

#include "../../../../include/atm_mi355x.h"
#include "ATMMetaForce.h"
#include "ATMMetaForceKernels.h"
#include "MI355XATMMetaForceKernels.h"
#include "MI355XPlatform.h"
#include "openmm/HarmonicAngleForce.h"
#include "openmm/HarmonicBondForce.h"
#include "openmm/KernelFactory.h"
#include "openmm/NonbondedForce.h"
#include "openmm/PeriodicTorsionForce.h"
#include "openmm/OpenMMException.h"
#include "openmm/internal/ContextImpl.h"

using namespace ATMMetaForcePlugin;
using namespace OpenMM;

static_assert(sizeof(Vec3) == 3 * sizeof(double), "Vec3 must be three packed doubles");

namespace {

void check(int rc, const char* what) {
    if (rc != ATM_OK) throw OpenMMException(std::string("MI355X: ") + what + ": " + atm_last_error());
}

// The NonbondedForce among the forces whose group is in `mask`, with the
// bonded forces of those groups (HarmonicBondForce, HarmonicAngleForce,
// PeriodicTorsionForce) appended to `bonded`; any other force in those groups
// is rejected, as is a second NonbondedForce.
const NonbondedForce* variableNonbonded(const System& system, int mask, const Force* skip,
                                        std::vector<const Force*>* bonded) {
    const NonbondedForce* nb = nullptr;
    for (int i = 0; i < system.getNumForces(); ++i) {
        const Force& f = system.getForce(i);
        if (&f == skip || !(mask & (1 << f.getForceGroup()))) continue;
        if (dynamic_cast<const HarmonicBondForce*>(&f) || dynamic_cast<const HarmonicAngleForce*>(&f) ||
            dynamic_cast<const PeriodicTorsionForce*>(&f)) {
            bonded->push_back(&f);
            continue;
        }
        const NonbondedForce* n = dynamic_cast<const NonbondedForce*>(&f);
        if (!n)
            throw OpenMMException("MI355X ATMMetaForce: the variable force groups may hold one NonbondedForce and "
                                  "harmonic bond, harmonic angle and periodic torsion forces (found '" +
                                  f.getName() + "')");
        if (nb) throw OpenMMException("MI355X ATMMetaForce: more than one NonbondedForce in the variable force groups");
        nb = n;
    }
    return nb;
}

void validateNonbonded(const NonbondedForce& nb) {
    if (nb.getNonbondedMethod() == NonbondedForce::PME) {
        if (nb.getReciprocalSpaceForceGroup() != -1 && nb.getReciprocalSpaceForceGroup() != nb.getForceGroup())
            throw OpenMMException("MI355X platform: the PME reciprocal-space part must stay in the NonbondedForce's group");
        return;
    }
    if (nb.getNonbondedMethod() != NonbondedForce::CutoffPeriodic)
        throw OpenMMException("MI355X platform: NonbondedForce must use CutoffPeriodic (reaction-field) or PME electrostatics");
}

// The NonbondedForce's exceptions as the C ABI takes them.
void exceptionArrays(const NonbondedForce& nb, std::vector<int32_t>& pairs, std::vector<double>& qq,
                     std::vector<double>& es, std::vector<double>& ee) {
    const int ne = nb.getNumExceptions();
    pairs.resize(2 * (size_t)ne);
    qq.resize(ne);
    es.resize(ne);
    ee.resize(ne);
    for (int k = 0; k < ne; ++k) {
        int p1, p2;
        nb.getExceptionParameters(k, p1, p2, qq[k], es[k], ee[k]);
        pairs[2 * k] = p1;
        pairs[2 * k + 1] = p2;
    }
}

class NonbondedFactory : public KernelFactory {
public:
    KernelImpl* createKernelImpl(std::string name, const Platform& platform, ContextImpl& context) const override {
        if (name == CalcNonbondedForceKernel::Name()) return new MI355XCalcNonbondedForceKernel(name, platform, context);
        throw OpenMMException("Tried to create kernel with illegal kernel name '" + name + "'");
    }
};

class BondedFactory : public KernelFactory {
public:
    KernelImpl* createKernelImpl(std::string name, const Platform& platform, ContextImpl& context) const override {
        if (name == CalcHarmonicBondForceKernel::Name()) return new MI355XCalcHarmonicBondForceKernel(name, platform, context);
        if (name == CalcHarmonicAngleForceKernel::Name()) return new MI355XCalcHarmonicAngleForceKernel(name, platform, context);
        if (name == CalcPeriodicTorsionForceKernel::Name())
            return new MI355XCalcPeriodicTorsionForceKernel(name, platform, context);
        throw OpenMMException("Tried to create kernel with illegal kernel name '" + name + "'");
    }
};

class ATMFactory : public KernelFactory {
public:
    KernelImpl* createKernelImpl(std::string name, const Platform& platform, ContextImpl& context) const override {
        if (name == CalcATMMetaForceKernel::Name()) return new MI355XCalcATMMetaForceKernel(name, platform, context);
        if (name == CalcCentroidRestraintForceKernel::Name())
            return new MI355XCalcCentroidRestraintForceKernel(name, platform, context);
        if (name == CalcPositionRestraintForceKernel::Name())
            return new MI355XCalcPositionRestraintForceKernel(name, platform, context);
        throw OpenMMException("Tried to create kernel with illegal kernel name '" + name + "'");
    }
};

// Box edges of a context (the diagonal of its box vectors).
void boxEdges(ContextImpl& ctx, double L[3]) {
    Vec3 a, b, c;
    ctx.getPeriodicBoxVectors(a, b, c);
    L[0] = a[0]; L[1] = b[1]; L[2] = c[2];
}

bool triclinic(const Vec3& b, const Vec3& c) { return b[0] != 0.0 || c[0] != 0.0 || c[1] != 0.0; }

void requireOrthorhombic(const Vec3& b, const Vec3& c, const char* what) {
    if (triclinic(b, c))
        throw OpenMMException(std::string(what) + " supports orthorhombic periodic boxes only on the MI355X platform");
}

}  // namespace

// ------------------------------------------------------------------ platform

MI355XPlatform::MI355XPlatform() {
    platformProperties = {"DeviceIndex", "Precision", "ATMEvaluation"};
    setPropertyDefaultValue("DeviceIndex", "0");
    setPropertyDefaultValue("Precision", "mixed");
    setPropertyDefaultValue("ATMEvaluation", "fused");
    registerKernelFactory(CalcNonbondedForceKernel::Name(), new NonbondedFactory());
    registerKernelFactory(CalcHarmonicBondForceKernel::Name(), new BondedFactory());
    registerKernelFactory(CalcHarmonicAngleForceKernel::Name(), new BondedFactory());
    registerKernelFactory(CalcPeriodicTorsionForceKernel::Name(), new BondedFactory());
}

const std::string& MI355XPlatform::getPropertyValue(const Context& context, const std::string& property) const {
    return Platform::getPropertyValue(context, property);
}

void MI355XPlatform::contextCreated(ContextImpl& context, const std::map<std::string, std::string>& properties) const {
    PlatformData* d = new PlatformData();
    for (const std::string& p : platformProperties) d->properties[p] = getPropertyDefaultValue(p);
    for (auto& kv : properties) {
        bool known = false;
        for (const std::string& p : platformProperties) known |= p == kv.first;
        if (!known) {
            delete d;
            throw OpenMMException("Illegal property name: " + kv.first);
        }
        d->properties[kv.first] = kv.second;
    }
    d->device = std::atoi(d->properties["DeviceIndex"].c_str());
    const std::string& prec = d->properties["Precision"];
    if (prec != "mixed" && prec != "double" && prec != "single") {
        delete d;
        throw OpenMMException("Illegal value for Precision: " + prec);
    }
    d->fp32Pairs = prec == "single";
    const std::string& ev = d->properties["ATMEvaluation"];
    if (ev != "fused" && ev != "plain") {
        delete d;
        throw OpenMMException("Illegal value for ATMEvaluation: " + ev);
    }
    d->plainATM = ev == "plain";
    context.setPlatformData(d);
}

void MI355XPlatform::contextDestroyed(ContextImpl& context) const {
    delete static_cast<PlatformData*>(context.getPlatformData());
    context.setPlatformData(nullptr);
}

MI355XPlatform::PlatformData& MI355XPlatform::data(ContextImpl& context) {
    return *static_cast<PlatformData*>(context.getPlatformData());
}

// ------------------------------------------------------------------ one system of the C ABI

MI355XSystemContext::~MI355XSystemContext() {
    if (ctx) atm_destroy(ctx);
    if (pme) atm_pme_destroy(pme);
}

void MI355XSystemContext::create(int device, const System& system, const NonbondedForce& nb,
                                 const std::vector<Vec3>* displacements) {
    validateNonbonded(nb);
    if (nb.getNonbondedMethod() == NonbondedForce::PME) {
        createPME(device, system, nb, displacements);
        return;
    }
    n = system.getNumParticles();
    if (nb.getNumParticles() != n)
        throw OpenMMException("NonbondedForce must have exactly as many particles as the System it belongs to.");
    Vec3 a, b, c;
    system.getDefaultPeriodicBoxVectors(a, b, c);
    atm_system_desc d;
    d.num_atoms = n;
    d.box[0] = a[0];
    d.box[1] = b[1];
    d.box[2] = c[2];
    d.cutoff = nb.getCutoffDistance();
    d.rf_dielectric = nb.getReactionFieldDielectric();
    cutoff = d.cutoff;
    rfDielectric = d.rf_dielectric;
    for (int k = 0; k < 3; ++k) box[k] = d.box[k];
    check(atm_create(&d, 1, device, &ctx), "atm_create");
    if (triclinic(b, c)) {
        const double v[9] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]};
        check(atm_set_box_vectors(ctx, 0, v), "atm_set_box_vectors");
        tilt[0] = b[0]; tilt[1] = c[0]; tilt[2] = c[1];
    }
    setParticles(nb, displacements);
    std::vector<double> zeros(3 * (size_t)n, 0.0);
    check(atm_set_positions(ctx, 0, zeros.data()), "atm_set_positions");
}

// PME: the NonbondedForce's particles, exceptions and PME parameters on an
// atm_pme object (positions and box are passed at every evaluation).
void MI355XSystemContext::createPME(int device, const System& system, const NonbondedForce& nb,
                                    const std::vector<Vec3>* displacements) {
    n = system.getNumParticles();
    if (nb.getNumParticles() != n)
        throw OpenMMException("NonbondedForce must have exactly as many particles as the System it belongs to.");
    Vec3 a, b, c;
    system.getDefaultPeriodicBoxVectors(a, b, c);
    requireOrthorhombic(b, c, "NonbondedForce with PME");
    atm_pme_desc d{};
    d.num_atoms = n;
    double alpha;
    int nx, ny, nz;
    nb.getPMEParameters(alpha, nx, ny, nz);
    d.grid[0] = alpha > 0.0 ? nx : 0;
    d.grid[1] = alpha > 0.0 ? ny : 0;
    d.grid[2] = alpha > 0.0 ? nz : 0;
    d.box[0] = a[0];
    d.box[1] = b[1];
    d.box[2] = c[2];
    d.cutoff = nb.getCutoffDistance();
    d.ewald_tolerance = nb.getEwaldErrorTolerance();
    d.alpha = alpha > 0.0 ? alpha : 0.0;
    d.dispersion_correction = nb.getUseDispersionCorrection() ? 1 : 0;
    cutoff = d.cutoff;
    for (int k = 0; k < 3; ++k) box[k] = d.box[k];
    check(atm_pme_create(device, &d, &pme), "atm_pme_create");
    setPMEParticles(nb, displacements);
}

void MI355XSystemContext::setPMEParticles(const NonbondedForce& nb, const std::vector<Vec3>* displacements) {
    std::vector<double> q(n), s(n), e(n);
    for (int i = 0; i < n; ++i) nb.getParticleParameters(i, q[i], s[i], e[i]);
    const double* dptr = displacements ? reinterpret_cast<const double*>(displacements->data()) : nullptr;
    check(atm_pme_set_particles(pme, q.data(), s.data(), e.data(), dptr), "atm_pme_set_particles");
    std::vector<int32_t> pairs;
    std::vector<double> qq, es, ee;
    exceptionArrays(nb, pairs, qq, es, ee);
    const int ne = (int)qq.size();
    check(atm_pme_set_exceptions(pme, ne, pairs.data(), qq.data(), es.data(), ee.data()), "atm_pme_set_exceptions");
}

void MI355XSystemContext::evaluatePME(ContextImpl& context, const atm_params* params, bool includeForces, atm_result* r) {
    Vec3 a, b, c;
    context.getPeriodicBoxVectors(a, b, c);
    requireOrthorhombic(b, c, "NonbondedForce with PME");
    const double L[3] = {a[0], b[1], c[2]};
    if (params) check(atm_pme_set_params(pme, params), "atm_pme_set_params");
    double* f = includeForces ? reinterpret_cast<double*>(context.forcesRef().data()) : nullptr;
    check(atm_pme_eval(pme, reinterpret_cast<const double*>(context.positionsRef().data()), L,
                       params ? 0 : ATM_EVAL_STATE1_ONLY, f, r, nullptr),
          "atm_pme_eval");
}

void MI355XSystemContext::setParticles(const NonbondedForce& nb, const std::vector<Vec3>* displacements) {
    if (pme) {
        setPMEParticles(nb, displacements);
        return;
    }
    std::vector<double> q(n), s(n), e(n);
    for (int i = 0; i < n; ++i) nb.getParticleParameters(i, q[i], s[i], e[i]);
    const double* dptr = displacements ? reinterpret_cast<const double*>(displacements->data()) : nullptr;
    check(atm_set_particles(ctx, 0, q.data(), s.data(), e.data(), dptr), "atm_set_particles");
    if (nb.getNumExceptions() > 0 || haveExceptions) {  // (re)lays out the context: only when there are any
        std::vector<int32_t> pairs;
        std::vector<double> qq, es, ee;
        exceptionArrays(nb, pairs, qq, es, ee);
        check(atm_set_exceptions(ctx, 0, (int)qq.size(), pairs.data(), qq.data(), es.data(), ee.data()),
              "atm_set_exceptions");
        haveExceptions = !qq.empty();
    }
    // Long-range dispersion correction: E = 8 pi N^2 / V * (<eps s^12>/(9 rc^9) - <eps s^6>/(3 rc^3)),
    // averaged over unordered pairs i <= j of particle classes with equal (sigma, epsilon).
    dispersionCoefficient = 0.0;
    if (nb.getUseDispersionCorrection()) {
        std::map<std::pair<double, double>, long long> classes;
        for (int i = 0; i < n; ++i) classes[{s[i], e[i]}]++;
        std::vector<std::pair<std::pair<double, double>, long long>> cl(classes.begin(), classes.end());
        double s12 = 0.0, s6 = 0.0;
        for (size_t a = 0; a < cl.size(); ++a)
            for (size_t b = a; b < cl.size(); ++b) {
                const double sig = 0.5 * (cl[a].first.first + cl[b].first.first);
                const double eps = std::sqrt(cl[a].first.second * cl[b].first.second);
                const double w = a == b ? 0.5 * (double)cl[a].second * (cl[a].second + 1) : (double)cl[a].second * cl[b].second;
                const double sig6 = std::pow(sig, 6.0);
                s12 += w * eps * sig6 * sig6;
                s6 += w * eps * sig6;
            }
        const double npairs = 0.5 * (double)n * (n + 1);
        const double rc3 = cutoff * cutoff * cutoff;
        dispersionCoefficient = 8.0 * M_PI * (double)n * n * (s12 / npairs / (9.0 * rc3 * rc3 * rc3) - s6 / npairs / (3.0 * rc3));
    }
}

void MI355XSystemContext::setEnergyOffsets(double state1, double state2) {
    if (pme) check(atm_pme_set_energy_offsets(pme, state1, state2), "atm_pme_set_energy_offsets");
    else check(atm_set_energy_offsets(ctx, 0, state1, state2), "atm_set_energy_offsets");
}

void MI355XSystemContext::setDisplacements(const std::vector<Vec3>& displacements) {
    if (pme) {  // displacements only: the particle parameters of the last upload stay
        check(atm_pme_set_particles(pme, nullptr, nullptr, nullptr, reinterpret_cast<const double*>(displacements.data())),
              "atm_pme_set_particles");
        return;
    }
    check(atm_set_displacements(ctx, 0, reinterpret_cast<const double*>(displacements.data())), "atm_set_displacements");
}

double MI355XSystemContext::dispersionEnergy(const Vec3 b[3]) const {
    return dispersionCoefficient == 0.0 ? 0.0 : dispersionCoefficient / (b[0][0] * b[1][1] * b[2][2]);
}

void MI355XSystemContext::syncState(ContextImpl& context) {
    Vec3 a, b, c;
    context.getPeriodicBoxVectors(a, b, c);
    if (a[0] != box[0] || b[1] != box[1] || c[2] != box[2] || b[0] != tilt[0] || c[0] != tilt[1] ||
        c[1] != tilt[2]) {
        if (triclinic(b, c)) {
            const double v[9] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2]};
            check(atm_set_box_vectors(ctx, 0, v), "atm_set_box_vectors");
        } else {
            const double nb[3] = {a[0], b[1], c[2]};
            check(atm_set_box(ctx, 0, nb), "atm_set_box");
        }
        box[0] = a[0]; box[1] = b[1]; box[2] = c[2];
        tilt[0] = b[0]; tilt[1] = c[0]; tilt[2] = c[1];
    }
    check(atm_set_positions(ctx, 0, reinterpret_cast<const double*>(context.positionsRef().data())), "atm_set_positions");
}

void MI355XSystemContext::evaluate(int flags) {
    for (int attempt = 0; attempt < 3; ++attempt) {
        check(atm_eval(ctx, flags), "atm_eval");
        atm_result r;
        const int rc = atm_get_results(ctx, 0, &r);
        if (rc == ATM_ERR_CAPACITY) continue;  // the list capacity was doubled: evaluate again
        check(rc, "atm_get_results");
        return;
    }
    throw OpenMMException("MI355X: neighbour list capacity could not be satisfied");
}

static void addForces(void* ctx, int n, ContextImpl& context) {
    std::vector<double> f(3 * (size_t)n);
    check(atm_get_forces(ctx, 0, f.data()), "atm_get_forces");
    std::vector<Vec3>& F = context.forcesRef();
    for (int i = 0; i < n; ++i) F[i] += Vec3(f[3 * i], f[3 * i + 1], f[3 * i + 2]);
}

// ------------------------------------------------------------------ NonbondedForce kernel

void MI355XCalcNonbondedForceKernel::initialize(const System& system, const NonbondedForce& force) {
    sys.create(MI355XPlatform::data(context).device, system, force, nullptr);
}

double MI355XCalcNonbondedForceKernel::execute(ContextImpl& ctx, bool includeForces, bool includeEnergy) {
    if (sys.isPME()) {
        atm_result r;
        sys.evaluatePME(ctx, nullptr, includeForces, &r);
        return includeEnergy ? r.energy : 0.0;
    }
    sys.syncState(ctx);
    const MI355XPlatform::PlatformData& pd = MI355XPlatform::data(ctx);
    int flags = ATM_EVAL_STATE1_ONLY | ATM_EVAL_CLEAR_FORCES | (includeForces ? ATM_EVAL_FORCES : 0) |
                (pd.fp32Pairs ? ATM_EVAL_FP32_PAIRS : 0);
    sys.evaluate(flags);
    atm_result r;
    check(atm_get_results(sys.handle(), 0, &r), "atm_get_results");
    if (includeForces) addForces(sys.handle(), sys.numAtoms(), ctx);
    Vec3 b[3];
    ctx.getPeriodicBoxVectors(b[0], b[1], b[2]);
    return includeEnergy ? r.state1_energy + sys.dispersionEnergy(b) : 0.0;
}

void MI355XCalcNonbondedForceKernel::copyParametersToContext(ContextImpl& ctx, const NonbondedForce& force) {
    (void)ctx;
    if (force.getNumParticles() != sys.numAtoms())
        throw OpenMMException("updateParametersInContext: The number of particles has changed");
    validateNonbonded(force);
    sys.setParticles(force, nullptr);
}

namespace {

// Terms of one bonded force as the flat arrays of atm_bonded_set.
struct BondedTerms {
    std::vector<int32_t> bonds, angles, torsions;
    std::vector<double> bprm, aprm, tprm;
};

void termsOf(const HarmonicBondForce& f, BondedTerms& t) {
    for (int k = 0; k < f.getNumBonds(); ++k) {
        int p1, p2;
        double r0, kf;
        f.getBondParameters(k, p1, p2, r0, kf);
        t.bonds.insert(t.bonds.end(), {p1, p2});
        t.bprm.insert(t.bprm.end(), {r0, kf});
    }
}

void termsOf(const HarmonicAngleForce& f, BondedTerms& t) {
    for (int k = 0; k < f.getNumAngles(); ++k) {
        int p1, p2, p3;
        double a0, kf;
        f.getAngleParameters(k, p1, p2, p3, a0, kf);
        t.angles.insert(t.angles.end(), {p1, p2, p3});
        t.aprm.insert(t.aprm.end(), {a0, kf});
    }
}

void termsOf(const PeriodicTorsionForce& f, BondedTerms& t) {
    for (int k = 0; k < f.getNumTorsions(); ++k) {
        int p1, p2, p3, p4, per;
        double ph, kf;
        f.getTorsionParameters(k, p1, p2, p3, p4, per, ph, kf);
        t.torsions.insert(t.torsions.end(), {p1, p2, p3, p4});
        t.tprm.insert(t.tprm.end(), {(double)per, ph, kf});
    }
}

}  // namespace

// ------------------------------------------------------------------ ATMMetaForce kernel

MI355XCalcATMMetaForceKernel::MI355XCalcATMMetaForceKernel(std::string name, const Platform& platform,
                                                           ContextImpl& context)
    : CalcATMMetaForceKernel(name, platform), context(context) {}

// Entry i of the force displaces particle i (the reference's convention,
// ReferenceATMMetaForceKernels.cpp:120-123), so every particle of the System
// must be added, in order.
static std::vector<Vec3> displacementTable(const System& system, const ATMMetaForce& force, std::vector<int>* particles) {
    const int n = system.getNumParticles();
    if (force.getNumParticles() != n)
        throw OpenMMException("ATMMetaForce: every particle of the System must be added to the force (" +
                              std::to_string(force.getNumParticles()) + " of " + std::to_string(n) + ")");
    std::vector<Vec3> d(n);
    if (particles) particles->resize(n);
    for (int i = 0; i < n; ++i) {
        int p;
        double dx, dy, dz;
        force.getParticleParameters(i, p, dx, dy, dz);
        d[i] = Vec3(dx, dy, dz);
        if (particles) (*particles)[i] = p;
    }
    return d;
}

MI355XCalcATMMetaForceKernel::~MI355XCalcATMMetaForceKernel() {
    if (variableBonded) atm_bonded_destroy(variableBonded);
}

void MI355XCalcATMMetaForceKernel::initialize(const System& system, const ATMMetaForce& force) {
    int mask = 0;
    for (int g : force.getVariableForceGroups()) mask |= 1 << g;
    std::vector<const Force*> bonded;
    const NonbondedForce* nb = variableNonbonded(system, mask, &force, &bonded);
    displacements = displacementTable(system, force, &particles);
    hasVariableForces = nb != nullptr;
    if (!bonded.empty()) {
        if (!nb)
            throw OpenMMException("MI355X ATMMetaForce: bonded forces in the variable force groups need the "
                                  "NonbondedForce in those groups too");
        // every bonded term of the variable groups in one set, evaluated at x and at x + d
        BondedTerms t;
        for (const Force* f : bonded) {
            if (auto b = dynamic_cast<const HarmonicBondForce*>(f)) termsOf(*b, t);
            else if (auto a = dynamic_cast<const HarmonicAngleForce*>(f)) termsOf(*a, t);
            else termsOf(*dynamic_cast<const PeriodicTorsionForce*>(f), t);
        }
        check(atm_bonded_create(MI355XPlatform::data(context).device, &variableBonded), "atm_bonded_create");
        check(atm_bonded_set(variableBonded, system.getNumParticles(), (int32_t)t.bonds.size() / 2, t.bonds.data(),
                             t.bprm.data(), (int32_t)t.angles.size() / 3, t.angles.data(), t.aprm.data(),
                             (int32_t)t.torsions.size() / 4, t.torsions.data(), t.tprm.data()),
              "atm_bonded_set");
    }
    if (hasVariableForces) sys.create(MI355XPlatform::data(context).device, system, *nb, &displacements);
}

double MI355XCalcATMMetaForceKernel::executeFused(ContextImpl& ctx, bool includeForces, bool includeEnergy) {
    atm_params p;
    p.lambda1 = ctx.getParameter(ATMMetaForce::Lambda1());
    p.lambda2 = ctx.getParameter(ATMMetaForce::Lambda2());
    p.alpha = ctx.getParameter(ATMMetaForce::Alpha());
    p.u0 = ctx.getParameter(ATMMetaForce::U0());
    p.w0 = ctx.getParameter(ATMMetaForce::W0());
    p.umax = ctx.getParameter(ATMMetaForce::Umax());
    p.ubcore = ctx.getParameter(ATMMetaForce::Ubcore());
    p.acore = ctx.getParameter(ATMMetaForce::Acore());
    p.direction = ctx.getParameter(ATMMetaForce::Direction());
    if (!hasVariableForces) {
        // nothing varies: U1 = U2 = 0; the soft-core and softplus of u = 0
        U1 = U2 = 0.0;
        const double x = -p.alpha * (0.0 - p.u0);
        double w = p.lambda2 * 0.0 + p.w0;
        if (p.alpha > 0.0) w += (p.lambda2 - p.lambda1) / p.alpha * (x > 0.0 ? x + std::log1p(std::exp(-x)) : std::log1p(std::exp(x)));
        PerturbationEnergy = 0.0;
        return includeEnergy ? w : 0.0;
    }
    // bonded terms of the variable groups at x and at x + d: their energies enter U1 and
    // U2 before the combine, their forces are blended with the evaluation's weights
    std::vector<double> fb1, fb2;
    if (variableBonded) {
        const std::vector<Vec3>& x = ctx.positionsRef();
        std::vector<Vec3> xd(x.size());
        for (size_t i = 0; i < x.size(); ++i) xd[i] = x[i] + displacements[i];
        if (includeForces) {
            fb1.assign(3 * x.size(), 0.0);
            fb2.assign(3 * x.size(), 0.0);
        }
        double e1[3], e2[3];
        check(atm_bonded_eval(variableBonded, reinterpret_cast<const double*>(x.data()),
                              includeForces ? fb1.data() : nullptr, e1), "atm_bonded_eval");
        check(atm_bonded_eval(variableBonded, reinterpret_cast<const double*>(xd.data()),
                              includeForces ? fb2.data() : nullptr, e2), "atm_bonded_eval");
        sys.setEnergyOffsets(e1[0] + e1[1] + e1[2], e2[0] + e2[1] + e2[2]);
    }
    auto blendBonded = [&](const atm_result& r) {
        if (fb1.empty()) return;
        const double w2 = p.direction > 0.0 ? r.hybrid_weight : 1.0 - r.hybrid_weight, w1 = 1.0 - w2;
        std::vector<Vec3>& F = ctx.forcesRef();
        for (size_t i = 0; i < F.size(); ++i)
            F[i] += Vec3(w1 * fb1[3 * i] + w2 * fb2[3 * i], w1 * fb1[3 * i + 1] + w2 * fb2[3 * i + 1],
                         w1 * fb1[3 * i + 2] + w2 * fb2[3 * i + 2]);
    };
    if (sys.isPME()) {
        atm_result r;
        sys.evaluatePME(ctx, &p, includeForces, &r);  // dispersion correction inside U1 and U2
        blendBonded(r);
        U1 = r.state1_energy;
        U2 = r.state2_energy;
        PerturbationEnergy = r.perturbation_energy;
        return includeEnergy ? r.energy : 0.0;
    }
    sys.syncState(ctx);
    check(atm_set_params(sys.handle(), 0, &p), "atm_set_params");
    const MI355XPlatform::PlatformData& pd = MI355XPlatform::data(ctx);
    const int flags = ATM_EVAL_CLEAR_FORCES | (includeForces ? ATM_EVAL_FORCES : 0) | (pd.plainATM ? ATM_EVAL_PLAIN : 0) |
                      (pd.fp32Pairs ? ATM_EVAL_FP32_PAIRS : 0);
    sys.evaluate(flags);
    atm_result r;
    check(atm_get_results(sys.handle(), 0, &r), "atm_get_results");
    if (includeForces) addForces(sys.handle(), sys.numAtoms(), ctx);
    blendBonded(r);
    Vec3 b[3];
    ctx.getPeriodicBoxVectors(b[0], b[1], b[2]);
    const double edisp = sys.dispersionEnergy(b);  // present in both states: shifts e0 only
    U1 = r.state1_energy + edisp;
    U2 = r.state2_energy + edisp;
    PerturbationEnergy = r.perturbation_energy;
    return includeEnergy ? r.energy + edisp : 0.0;
}

double MI355XCalcATMMetaForceKernel::execute(ContextImpl&, ContextImpl&, ContextImpl&, double, double, bool, bool) {
    throw OpenMMException("MI355X CalcATMMetaForce evaluates both states itself; use executeFused");
}

void MI355XCalcATMMetaForceKernel::copyState(ContextImpl&, ContextImpl&, ContextImpl&) {
    throw OpenMMException("MI355X CalcATMMetaForce uses no inner contexts");
}

// Displacements are replaced (the behaviour of CommonATMMetaForceKernels.cpp:229-251;
// the reference CPU kernel appends instead, ReferenceATMMetaForceKernels.cpp:155).
void MI355XCalcATMMetaForceKernel::copyParametersToContext(ContextImpl& ctx, const ATMMetaForce& force) {
    (void)ctx;
    if (force.getNumParticles() != (int)particles.size())
        throw OpenMMException("copyParametersToContext: The number of ATMMetaForce particles has changed");
    std::vector<Vec3> d(particles.size());
    for (int i = 0; i < force.getNumParticles(); ++i) {
        int p;
        double dx, dy, dz;
        force.getParticleParameters(i, p, dx, dy, dz);
        if (p != particles[i])
            throw OpenMMException("MI355XCalcATMMetaForceKernel::copyParametersToContext: A particle index has changed");
        d[i] = Vec3(dx, dy, dz);
    }
    displacements = d;
    if (hasVariableForces) sys.setDisplacements(d);
}

// ------------------------------------------------------------------ restraint kernels

namespace {

void setCentroid(void* set, const System& system, const CentroidRestraintForce& force) {
    std::vector<int> members;
    std::vector<double> weights;
    std::vector<int> first(force.getNumGroups()), count(force.getNumGroups());
    const int n = system.getNumParticles();
    for (int g = 0; g < force.getNumGroups(); ++g) {
        std::vector<int> p;
        std::vector<double> w;
        force.getGroupParameters(g, p, w);
        first[g] = (int)members.size();
        count[g] = (int)p.size();
        for (size_t k = 0; k < p.size(); ++k) {
            if (p[k] < 0 || p[k] >= n) throw OpenMMException("CentroidRestraintForce: particle index out of range");
            members.push_back(p[k]);
            // no explicit weights: particle masses, as CustomCentroidBondForce
            weights.push_back(w.empty() ? system.getParticleMass(p[k]) : w[k]);
        }
    }
    std::vector<atm_centroid_restraint> bonds(force.getNumBonds());
    for (int b = 0; b < force.getNumBonds(); ++b) {
        int g1, g2;
        double kf, tol;
        Vec3 off;
        force.getBondParameters(b, g1, g2, kf, tol, off);
        if (g1 < 0 || g1 >= force.getNumGroups() || g2 < 0 || g2 >= force.getNumGroups())
            throw OpenMMException("CentroidRestraintForce: group index out of range");
        bonds[b] = atm_centroid_restraint{first[g1], count[g1], first[g2], count[g2], kf, tol, {off[0], off[1], off[2]}};
    }
    check(atm_restraints_set(set, n, (int)members.size(), members.data(), weights.data(), (int)bonds.size(), bonds.data(), 0,
                             nullptr),
          "atm_restraints_set");
}

void setPosition(void* set, const System& system, const PositionRestraintForce& force) {
    std::vector<atm_position_restraint> q(force.getNumParticles());
    for (int i = 0; i < force.getNumParticles(); ++i) {
        int p;
        Vec3 ref;
        double kf, tol;
        force.getParticleParameters(i, p, ref, kf, tol);
        q[i] = atm_position_restraint{p, force.getPeriodic() ? 1 : 0, {ref[0], ref[1], ref[2]}, kf, tol};
    }
    check(atm_restraints_set(set, system.getNumParticles(), 0, nullptr, nullptr, 0, nullptr, (int)q.size(), q.data()),
          "atm_restraints_set");
}

double evalRestraints(void* set, ContextImpl& ctx, bool includeForces, bool includeEnergy) {
    double L[3], e = 0.0;
    boxEdges(ctx, L);
    check(atm_restraints_eval(set, reinterpret_cast<const double*>(ctx.positionsRef().data()), L,
                              includeForces ? reinterpret_cast<double*>(ctx.forcesRef().data()) : nullptr, &e),
          "atm_restraints_eval");
    return includeEnergy ? e : 0.0;
}

}  // namespace

MI355XCalcCentroidRestraintForceKernel::~MI355XCalcCentroidRestraintForceKernel() {
    if (set) atm_restraints_destroy(set);
}

void MI355XCalcCentroidRestraintForceKernel::initialize(const System& system, const CentroidRestraintForce& force) {
    check(atm_restraints_create(MI355XPlatform::data(context).device, &set), "atm_restraints_create");
    setCentroid(set, system, force);
}

double MI355XCalcCentroidRestraintForceKernel::execute(ContextImpl& ctx, bool includeForces, bool includeEnergy) {
    return evalRestraints(set, ctx, includeForces, includeEnergy);
}

void MI355XCalcCentroidRestraintForceKernel::copyParametersToContext(ContextImpl& ctx, const CentroidRestraintForce& force) {
    setCentroid(set, ctx.getSystem(), force);
}

MI355XCalcPositionRestraintForceKernel::~MI355XCalcPositionRestraintForceKernel() {
    if (set) atm_restraints_destroy(set);
}

void MI355XCalcPositionRestraintForceKernel::initialize(const System& system, const PositionRestraintForce& force) {
    check(atm_restraints_create(MI355XPlatform::data(context).device, &set), "atm_restraints_create");
    setPosition(set, system, force);
    periodic = force.getPeriodic();
}

double MI355XCalcPositionRestraintForceKernel::execute(ContextImpl& ctx, bool includeForces, bool includeEnergy) {
    if (periodic) {
        Vec3 a, b, c;
        ctx.getPeriodicBoxVectors(a, b, c);
        requireOrthorhombic(b, c, "A periodic PositionRestraintForce");
    }
    return evalRestraints(set, ctx, includeForces, includeEnergy);
}

void MI355XCalcPositionRestraintForceKernel::copyParametersToContext(ContextImpl& ctx, const PositionRestraintForce& force) {
    setPosition(set, ctx.getSystem(), force);
    periodic = force.getPeriodic();
}

// ------------------------------------------------------------------ plugin entry points

extern "C" __attribute__((visibility("default"))) void registerPlatforms() {
    try {
        Platform::getPlatformByName("MI355X");
    } catch (...) {
        Platform::registerPlatform(new MI355XPlatform());
    }
}

// Registers the ATMMetaForce kernel factory with the MI355X platform; errors are
// ignored, as the reference's factories do (HipATMMetaForceKernelFactory.cpp:18-27).
// ------------------------------------------------------------------ bonded kernels

template <class F, class Base>
MI355XCalcBondedKernel<F, Base>::~MI355XCalcBondedKernel() {
    if (set) atm_bonded_destroy(set);
}

template <class F, class Base>
void MI355XCalcBondedKernel<F, Base>::upload(const F& force) {
    BondedTerms t;
    termsOf(force, t);
    check(atm_bonded_set(set, n, (int32_t)t.bonds.size() / 2, t.bonds.data(), t.bprm.data(), (int32_t)t.angles.size() / 3,
                         t.angles.data(), t.aprm.data(), (int32_t)t.torsions.size() / 4, t.torsions.data(),
                         t.tprm.data()),
          "atm_bonded_set");
}

template <class F, class Base>
void MI355XCalcBondedKernel<F, Base>::initialize(const System& system, const F& force) {
    n = system.getNumParticles();
    check(atm_bonded_create(MI355XPlatform::data(context).device, &set), "atm_bonded_create");
    upload(force);
}

template <class F, class Base>
double MI355XCalcBondedKernel<F, Base>::execute(ContextImpl& ctx, bool includeForces, bool includeEnergy) {
    double e[3];
    check(atm_bonded_eval(set, reinterpret_cast<const double*>(ctx.positionsRef().data()),
                          includeForces ? reinterpret_cast<double*>(ctx.forcesRef().data()) : nullptr, e),
          "atm_bonded_eval");
    return includeEnergy ? e[0] + e[1] + e[2] : 0.0;
}

template <class F, class Base>
void MI355XCalcBondedKernel<F, Base>::copyParametersToContext(ContextImpl& ctx, const F& force) {
    (void)ctx;
    upload(force);
}

template class MI355XCalcBondedKernel<HarmonicBondForce, CalcHarmonicBondForceKernel>;
template class MI355XCalcBondedKernel<HarmonicAngleForce, CalcHarmonicAngleForceKernel>;
template class MI355XCalcBondedKernel<PeriodicTorsionForce, CalcPeriodicTorsionForceKernel>;

extern "C" __attribute__((visibility("default"))) void registerKernelFactories() {
    try {
        Platform& platform = Platform::getPlatformByName("MI355X");
        platform.registerKernelFactory(CalcATMMetaForceKernel::Name(), new ATMFactory());
        platform.registerKernelFactory(CalcCentroidRestraintForceKernel::Name(), new ATMFactory());
        platform.registerKernelFactory(CalcPositionRestraintForceKernel::Name(), new ATMFactory());
    } catch (...) {
    }
}

extern "C" __attribute__((visibility("default"))) void registerATMMetaForceMI355XKernelFactories() {
    registerPlatforms();
    registerKernelFactories();
}

namespace {
struct AutoRegister {
    AutoRegister() { registerATMMetaForceMI355XKernelFactories(); }
} autoRegister;
}  // namespace
