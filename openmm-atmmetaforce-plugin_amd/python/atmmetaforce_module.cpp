// Python module _atmmetaforce: the reference's SWIG surface
// (python/atmmetaforceplugin.i:64-124) -- ATMMetaForce with its static
// parameter-name methods, cast/isinstance and ATMMETAFORCE_VERSION -- plus the
// minimal OpenMM-compatible classes a script needs around it (System,
// NonbondedForce, Context, State, Platform, VerletIntegrator, XmlSerializer).
// C++ exceptions surface as Python exceptions (atmmetaforceplugin.i:54-61).
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "ATMMetaForce.h"
#include "ATMRestraintForces.h"
#include "ATMMetaForceVersion.h"
#include "internal/ATMMetaForceImpl.h"
#include "openmm/Context.h"
#include "openmm/HarmonicAngleForce.h"
#include "openmm/HarmonicBondForce.h"
#include "openmm/Integrator.h"
#include "openmm/MonteCarloBarostat.h"
#include "openmm/NonbondedForce.h"
#include "openmm/OpenMMException.h"
#include "openmm/PeriodicTorsionForce.h"
#include "openmm/Platform.h"
#include "openmm/State.h"
#include "openmm/System.h"
#include "openmm/serialization/XmlSerializer.h"

namespace py = pybind11;
using namespace OpenMM;
using ATMMetaForcePlugin::ATMMetaForce;

namespace {

Vec3 toVec3(const py::handle& h) {
    py::sequence s = py::reinterpret_borrow<py::sequence>(h);
    if (py::len(s) != 3) throw OpenMMException("expected a 3-vector");
    return Vec3(s[0].cast<double>(), s[1].cast<double>(), s[2].cast<double>());
}

std::vector<Vec3> toVec3List(const py::object& o) {
    std::vector<Vec3> v;
    if (py::isinstance<py::array>(o)) {
        py::array_t<double, py::array::c_style | py::array::forcecast> a(o);
        if (a.ndim() != 2 || a.shape(1) != 3) throw OpenMMException("expected an (N, 3) array");
        auto r = a.unchecked<2>();
        v.reserve(a.shape(0));
        for (py::ssize_t i = 0; i < a.shape(0); ++i) v.emplace_back(r(i, 0), r(i, 1), r(i, 2));
        return v;
    }
    for (py::handle h : o) v.push_back(toVec3(h));
    return v;
}

py::object fromVec3List(const std::vector<Vec3>& v, bool asNumpy) {
    if (asNumpy) {
        py::array_t<double> a({(py::ssize_t)v.size(), (py::ssize_t)3});
        auto w = a.mutable_unchecked<2>();
        for (size_t i = 0; i < v.size(); ++i)
            for (int k = 0; k < 3; ++k) w(i, k) = v[i][k];
        return a;
    }
    py::list l;
    for (const Vec3& x : v) l.append(py::make_tuple(x[0], x[1], x[2]));
    return l;
}

int groupMask(const py::object& groups) {
    if (groups.is_none()) return -1;
    if (py::isinstance<py::int_>(groups)) return groups.cast<int>();
    int mask = 0;
    for (py::handle h : groups) {
        int g = h.cast<int>();
        if (g < 0 || g > 31) throw OpenMMException("Force group must be between 0 and 31");
        mask |= 1 << g;
    }
    return mask;
}

}  // namespace

PYBIND11_MODULE(_atmmetaforce, m) {
    m.doc() = "MI355X-native ATMMetaForce (C++ API binding)";
    py::register_exception<OpenMMException>(m, "OpenMMException", PyExc_Exception);
    m.attr("ATMMETAFORCE_VERSION") = ATMMETAFORCE_VERSION;

    // ---------------------------------------------------------------- core
    py::class_<Platform, std::unique_ptr<Platform, py::nodelete>>(m, "Platform")
        .def_static("getNumPlatforms", &Platform::getNumPlatforms)
        .def_static("getPlatform", &Platform::getPlatform, py::return_value_policy::reference)
        .def_static("getPlatformByName", &Platform::getPlatformByName, py::return_value_policy::reference)
        .def_static("loadPluginsFromDirectory", &Platform::loadPluginsFromDirectory)
        .def_static("getPluginLoadFailures", &Platform::getPluginLoadFailures)
        .def_static("getDefaultPluginsDirectory", &Platform::getDefaultPluginsDirectory)
        .def_static("getOpenMMVersion", &Platform::getOpenMMVersion)
        .def("getName", &Platform::getName)
        .def("getSpeed", &Platform::getSpeed)
        .def("supportsDoublePrecision", &Platform::supportsDoublePrecision)
        .def("getPropertyNames", &Platform::getPropertyNames)
        .def("getPropertyDefaultValue", &Platform::getPropertyDefaultValue)
        .def("setPropertyDefaultValue", &Platform::setPropertyDefaultValue);

    py::class_<Force, std::unique_ptr<Force, py::nodelete>>(m, "Force")
        .def("getForceGroup", &Force::getForceGroup)
        .def("setForceGroup", &Force::setForceGroup)
        .def("getName", &Force::getName)
        .def("setName", &Force::setName)
        .def("usesPeriodicBoundaryConditions", &Force::usesPeriodicBoundaryConditions);

    py::class_<MonteCarloBarostat, Force, std::unique_ptr<MonteCarloBarostat, py::nodelete>>(m, "MonteCarloBarostat")
        .def(py::init([](double p, double t, int freq) { return new MonteCarloBarostat(p, t, freq); }),
             py::arg("defaultPressure"), py::arg("defaultTemperature"), py::arg("frequency") = 25)
        .def_static("Pressure", &MonteCarloBarostat::Pressure)
        .def_static("Temperature", &MonteCarloBarostat::Temperature)
        .def("getDefaultPressure", &MonteCarloBarostat::getDefaultPressure)
        .def("setDefaultPressure", &MonteCarloBarostat::setDefaultPressure)
        .def("getDefaultTemperature", &MonteCarloBarostat::getDefaultTemperature)
        .def("setDefaultTemperature", &MonteCarloBarostat::setDefaultTemperature)
        .def("getFrequency", &MonteCarloBarostat::getFrequency)
        .def("setFrequency", &MonteCarloBarostat::setFrequency)
        .def("getRandomNumberSeed", &MonteCarloBarostat::getRandomNumberSeed)
        .def("setRandomNumberSeed", &MonteCarloBarostat::setRandomNumberSeed);

    py::class_<NonbondedForce, Force, std::unique_ptr<NonbondedForce, py::nodelete>> nb(m, "NonbondedForce");
    nb.def(py::init([]() { return new NonbondedForce(); }))
        .def("getNumParticles", &NonbondedForce::getNumParticles)
        .def("getNumExceptions", &NonbondedForce::getNumExceptions)
        .def("addParticle", &NonbondedForce::addParticle, py::arg("charge"), py::arg("sigma"), py::arg("epsilon"))
        .def("getParticleParameters", [](const NonbondedForce& f, int i) {
            double q, s, e;
            f.getParticleParameters(i, q, s, e);
            return py::make_tuple(q, s, e);
        })
        .def("setParticleParameters", &NonbondedForce::setParticleParameters)
        .def("addException", &NonbondedForce::addException, py::arg("particle1"), py::arg("particle2"),
             py::arg("chargeProd"), py::arg("sigma"), py::arg("epsilon"), py::arg("replace") = false)
        // methods are plain ints, as NonbondedForce.CutoffPeriodic etc. are in OpenMM's Python API
        .def("getNonbondedMethod", [](const NonbondedForce& f) { return (int)f.getNonbondedMethod(); })
        .def("setNonbondedMethod", [](NonbondedForce& f, int m) {
            f.setNonbondedMethod((NonbondedForce::NonbondedMethod)m);
        })
        .def("getCutoffDistance", &NonbondedForce::getCutoffDistance)
        .def("setCutoffDistance", &NonbondedForce::setCutoffDistance)
        .def("getReactionFieldDielectric", &NonbondedForce::getReactionFieldDielectric)
        .def("setReactionFieldDielectric", &NonbondedForce::setReactionFieldDielectric)
        .def("getUseDispersionCorrection", &NonbondedForce::getUseDispersionCorrection)
        .def("setUseDispersionCorrection", &NonbondedForce::setUseDispersionCorrection)
        .def("getReciprocalSpaceForceGroup", &NonbondedForce::getReciprocalSpaceForceGroup)
        .def("setReciprocalSpaceForceGroup", &NonbondedForce::setReciprocalSpaceForceGroup)
        .def("getExceptionParameters", [](const NonbondedForce& f, int i) {
            int p1, p2;
            double qq, s, e;
            f.getExceptionParameters(i, p1, p2, qq, s, e);
            return py::make_tuple(p1, p2, qq, s, e);
        })
        .def("setExceptionParameters", &NonbondedForce::setExceptionParameters)
        .def("getEwaldErrorTolerance", &NonbondedForce::getEwaldErrorTolerance)
        .def("setEwaldErrorTolerance", &NonbondedForce::setEwaldErrorTolerance)
        .def("getPMEParameters", [](const NonbondedForce& f) {
            double a;
            int x, y, z;
            f.getPMEParameters(a, x, y, z);
            return py::make_tuple(a, x, y, z);
        })
        .def("setPMEParameters", &NonbondedForce::setPMEParameters)
        .def("updateParametersInContext", &NonbondedForce::updateParametersInContext);
    nb.attr("NoCutoff") = (int)NonbondedForce::NoCutoff;
    nb.attr("CutoffNonPeriodic") = (int)NonbondedForce::CutoffNonPeriodic;
    nb.attr("CutoffPeriodic") = (int)NonbondedForce::CutoffPeriodic;
    nb.attr("Ewald") = (int)NonbondedForce::Ewald;
    nb.attr("PME") = (int)NonbondedForce::PME;
    nb.attr("LJPME") = (int)NonbondedForce::LJPME;

    py::class_<HarmonicBondForce, Force, std::unique_ptr<HarmonicBondForce, py::nodelete>>(m, "HarmonicBondForce")
        .def(py::init([]() { return new HarmonicBondForce(); }))
        .def("getNumBonds", &HarmonicBondForce::getNumBonds)
        .def("addBond", &HarmonicBondForce::addBond, py::arg("particle1"), py::arg("particle2"), py::arg("length"),
             py::arg("k"))
        .def("getBondParameters", [](const HarmonicBondForce& f, int i) {
            int a, b;
            double r0, k;
            f.getBondParameters(i, a, b, r0, k);
            return py::make_tuple(a, b, r0, k);
        })
        .def("setBondParameters", &HarmonicBondForce::setBondParameters)
        .def("setUsesPeriodicBoundaryConditions", &HarmonicBondForce::setUsesPeriodicBoundaryConditions)
        .def("updateParametersInContext", &HarmonicBondForce::updateParametersInContext);

    py::class_<HarmonicAngleForce, Force, std::unique_ptr<HarmonicAngleForce, py::nodelete>>(m, "HarmonicAngleForce")
        .def(py::init([]() { return new HarmonicAngleForce(); }))
        .def("getNumAngles", &HarmonicAngleForce::getNumAngles)
        .def("addAngle", &HarmonicAngleForce::addAngle, py::arg("particle1"), py::arg("particle2"),
             py::arg("particle3"), py::arg("angle"), py::arg("k"))
        .def("getAngleParameters", [](const HarmonicAngleForce& f, int i) {
            int a, b, c;
            double t0, k;
            f.getAngleParameters(i, a, b, c, t0, k);
            return py::make_tuple(a, b, c, t0, k);
        })
        .def("setAngleParameters", &HarmonicAngleForce::setAngleParameters)
        .def("setUsesPeriodicBoundaryConditions", &HarmonicAngleForce::setUsesPeriodicBoundaryConditions)
        .def("updateParametersInContext", &HarmonicAngleForce::updateParametersInContext);

    py::class_<PeriodicTorsionForce, Force, std::unique_ptr<PeriodicTorsionForce, py::nodelete>>(m, "PeriodicTorsionForce")
        .def(py::init([]() { return new PeriodicTorsionForce(); }))
        .def("getNumTorsions", &PeriodicTorsionForce::getNumTorsions)
        .def("addTorsion", &PeriodicTorsionForce::addTorsion, py::arg("particle1"), py::arg("particle2"),
             py::arg("particle3"), py::arg("particle4"), py::arg("periodicity"), py::arg("phase"), py::arg("k"))
        .def("getTorsionParameters", [](const PeriodicTorsionForce& f, int i) {
            int a, b, c, d, n;
            double ph, k;
            f.getTorsionParameters(i, a, b, c, d, n, ph, k);
            return py::make_tuple(a, b, c, d, n, ph, k);
        })
        .def("setTorsionParameters", &PeriodicTorsionForce::setTorsionParameters)
        .def("setUsesPeriodicBoundaryConditions", &PeriodicTorsionForce::setUsesPeriodicBoundaryConditions)
        .def("updateParametersInContext", &PeriodicTorsionForce::updateParametersInContext);

    py::class_<System>(m, "System")
        .def(py::init<>())
        .def("getNumParticles", &System::getNumParticles)
        .def("addParticle", &System::addParticle)
        .def("getParticleMass", &System::getParticleMass)
        .def("getNumConstraints", &System::getNumConstraints)
        .def("addConstraint", &System::addConstraint)
        .def("getConstraintParameters", [](const System& s, int i) {
            int a, b;
            double d;
            s.getConstraintParameters(i, a, b, d);
            return py::make_tuple(a, b, d);
        })
        .def("getNumForces", &System::getNumForces)
        // the System takes ownership of the force, as in OpenMM
        .def("addForce", [](System& s, Force* f) { return s.addForce(f); }, py::keep_alive<1, 2>())
        .def("getForce", (Force & (System::*)(int)) & System::getForce, py::return_value_policy::reference_internal)
        .def("getDefaultPeriodicBoxVectors", [](const System& s) {
            Vec3 a, b, c;
            s.getDefaultPeriodicBoxVectors(a, b, c);
            return py::make_tuple(py::make_tuple(a[0], a[1], a[2]), py::make_tuple(b[0], b[1], b[2]),
                                  py::make_tuple(c[0], c[1], c[2]));
        })
        .def("setDefaultPeriodicBoxVectors", [](System& s, py::object a, py::object b, py::object c) {
            s.setDefaultPeriodicBoxVectors(toVec3(a), toVec3(b), toVec3(c));
        })
        .def("usesPeriodicBoundaryConditions", &System::usesPeriodicBoundaryConditions);

    py::class_<Integrator>(m, "Integrator")
        .def("step", &Integrator::step)
        .def("getStepSize", &Integrator::getStepSize)
        .def("getConstraintTolerance", &Integrator::getConstraintTolerance)
        .def("setConstraintTolerance", &Integrator::setConstraintTolerance)
        .def("setStepSize", &Integrator::setStepSize)
        .def("getIntegrationForceGroups", &Integrator::getIntegrationForceGroups)
        .def("setIntegrationForceGroups", [](Integrator& i, py::object g) { i.setIntegrationForceGroups(groupMask(g)); });
    py::class_<VerletIntegrator, Integrator>(m, "VerletIntegrator").def(py::init<double>());
    py::class_<LangevinIntegrator, Integrator>(m, "LangevinIntegrator")
        .def(py::init<double, double, double>(), py::arg("temperature"), py::arg("frictionCoeff"), py::arg("stepSize"))
        .def("getTemperature", &LangevinIntegrator::getTemperature)
        .def("setTemperature", &LangevinIntegrator::setTemperature)
        .def("getFriction", &LangevinIntegrator::getFriction)
        .def("setFriction", &LangevinIntegrator::setFriction)
        .def("getRandomNumberSeed", &LangevinIntegrator::getRandomNumberSeed)
        .def("setRandomNumberSeed", &LangevinIntegrator::setRandomNumberSeed);
    py::class_<LangevinMiddleIntegrator, LangevinIntegrator>(m, "LangevinMiddleIntegrator")
        .def(py::init<double, double, double>(), py::arg("temperature"), py::arg("frictionCoeff"), py::arg("stepSize"));

    py::class_<State>(m, "State")
        .def("getTime", &State::getTime)
        .def("getPotentialEnergy", &State::getPotentialEnergy)
        .def("getKineticEnergy", &State::getKineticEnergy)
        .def("getForces", [](const State& s, bool asNumpy) { return fromVec3List(s.getForces(), asNumpy); },
             py::arg("asNumpy") = false)
        .def("getPositions", [](const State& s, bool asNumpy) { return fromVec3List(s.getPositions(), asNumpy); },
             py::arg("asNumpy") = false)
        .def("getVelocities", [](const State& s, bool asNumpy) { return fromVec3List(s.getVelocities(), asNumpy); },
             py::arg("asNumpy") = false)
        .def("getParameters", &State::getParameters)
        .def("getPeriodicBoxVectors", [](const State& s) {
            Vec3 a, b, c;
            s.getPeriodicBoxVectors(a, b, c);
            return py::make_tuple(py::make_tuple(a[0], a[1], a[2]), py::make_tuple(b[0], b[1], b[2]),
                                  py::make_tuple(c[0], c[1], c[2]));
        });

    py::class_<Context>(m, "Context")
        .def(py::init([](const System& s, Integrator& i) { return new Context(s, i); }), py::keep_alive<1, 2>(),
             py::keep_alive<1, 3>())
        .def(py::init([](const System& s, Integrator& i, Platform& p) { return new Context(s, i, p); }),
             py::keep_alive<1, 2>(), py::keep_alive<1, 3>())
        .def(py::init([](const System& s, Integrator& i, Platform& p, const std::map<std::string, std::string>& props) {
                 return new Context(s, i, p, props);
             }),
             py::keep_alive<1, 2>(), py::keep_alive<1, 3>())
        .def("getSystem", &Context::getSystem, py::return_value_policy::reference_internal)
        .def("getIntegrator", &Context::getIntegrator, py::return_value_policy::reference_internal)
        .def("getPlatform", &Context::getPlatform, py::return_value_policy::reference)
        .def("getState",
             [](const Context& c, bool getPositions, bool getVelocities, bool getForces, bool getEnergy,
                bool getParameters, bool enforcePeriodicBox, py::object groups) {
                 int types = (getPositions ? State::Positions : 0) | (getVelocities ? State::Velocities : 0) |
                             (getForces ? State::Forces : 0) | (getEnergy ? State::Energy : 0) |
                             (getParameters ? State::Parameters : 0);
                 return c.getState(types, enforcePeriodicBox, groupMask(groups));
             },
             py::arg("getPositions") = false, py::arg("getVelocities") = false, py::arg("getForces") = false,
             py::arg("getEnergy") = false, py::arg("getParameters") = false, py::arg("enforcePeriodicBox") = false,
             py::arg("groups") = py::none())
        .def("setPositions", [](Context& c, py::object x) { c.setPositions(toVec3List(x)); })
        .def("setVelocities", [](Context& c, py::object v) { c.setVelocities(toVec3List(v)); })
        .def("setTime", &Context::setTime)
        .def("getParameter", &Context::getParameter)
        .def("setParameter", &Context::setParameter)
        .def("getParameters", &Context::getParameters)
        .def("setState", &Context::setState)
        .def("setPeriodicBoxVectors", [](Context& c, py::object a, py::object b, py::object d) {
            c.setPeriodicBoxVectors(toVec3(a), toVec3(b), toVec3(d));
        });

    // ---------------------------------------------------------------- restraints
    using ATMMetaForcePlugin::CentroidRestraintForce;
    using ATMMetaForcePlugin::PositionRestraintForce;
    py::class_<CentroidRestraintForce, Force, std::unique_ptr<CentroidRestraintForce, py::nodelete>>(
        m, "CentroidRestraintForce")
        .def(py::init([]() { return new CentroidRestraintForce(); }))
        .def("getNumGroups", &CentroidRestraintForce::getNumGroups)
        .def("getNumBonds", &CentroidRestraintForce::getNumBonds)
        .def("addGroup", &CentroidRestraintForce::addGroup, py::arg("particles"),
             py::arg("weights") = std::vector<double>())
        .def("getGroupParameters", [](const CentroidRestraintForce& f, int i) {
            std::vector<int> p;
            std::vector<double> w;
            f.getGroupParameters(i, p, w);
            return py::make_tuple(p, w);
        })
        .def("addBond", [](CentroidRestraintForce& f, int g1, int g2, double kf, double tol, const py::object& off) {
            return f.addBond(g1, g2, kf, tol, off.is_none() ? Vec3() : toVec3(off));
        }, py::arg("group1"), py::arg("group2"), py::arg("kf"), py::arg("tol"), py::arg("offset") = py::none())
        .def("getBondParameters", [](const CentroidRestraintForce& f, int i) {
            int g1, g2;
            double kf, tol;
            Vec3 o;
            f.getBondParameters(i, g1, g2, kf, tol, o);
            return py::make_tuple(g1, g2, kf, tol, py::make_tuple(o[0], o[1], o[2]));
        })
        .def("setBondParameters", [](CentroidRestraintForce& f, int i, int g1, int g2, double kf, double tol,
                                     const py::object& off) {
            f.setBondParameters(i, g1, g2, kf, tol, toVec3(off));
        })
        .def("updateParametersInContext", &CentroidRestraintForce::updateParametersInContext);

    py::class_<PositionRestraintForce, Force, std::unique_ptr<PositionRestraintForce, py::nodelete>>(
        m, "PositionRestraintForce")
        .def(py::init([](bool periodic) { return new PositionRestraintForce(periodic); }), py::arg("periodic") = true)
        .def("getPeriodic", &PositionRestraintForce::getPeriodic)
        .def("getNumParticles", &PositionRestraintForce::getNumParticles)
        .def("addParticle", [](PositionRestraintForce& f, int p, const py::object& ref, double kf, double tol) {
            return f.addParticle(p, toVec3(ref), kf, tol);
        }, py::arg("particle"), py::arg("ref"), py::arg("kf"), py::arg("tol"))
        .def("getParticleParameters", [](const PositionRestraintForce& f, int i) {
            int p;
            Vec3 r;
            double kf, tol;
            f.getParticleParameters(i, p, r, kf, tol);
            return py::make_tuple(p, py::make_tuple(r[0], r[1], r[2]), kf, tol);
        })
        .def("setParticleParameters", [](PositionRestraintForce& f, int i, int p, const py::object& ref, double kf,
                                         double tol) { f.setParticleParameters(i, p, toVec3(ref), kf, tol); })
        .def("updateParametersInContext", &PositionRestraintForce::updateParametersInContext);

    // ---------------------------------------------------------------- ATMMetaForce
    py::class_<ATMMetaForce, Force, std::unique_ptr<ATMMetaForce, py::nodelete>>(m, "ATMMetaForce")
        .def(py::init([](double l1, double l2, double a, double u0, double w0, double umax, double ub, double ac,
                         double dir, const std::vector<int>& groups) {
                 return new ATMMetaForce(l1, l2, a, u0, w0, umax, ub, ac, dir, groups);
             }),
             py::arg("Lambda1"), py::arg("Lambda2"), py::arg("Alpha"), py::arg("U0"), py::arg("W0"), py::arg("Umax"),
             py::arg("Ubcore"), py::arg("Acore"), py::arg("direction"), py::arg("VariableForceGroups"))
        .def("getNumParticles", &ATMMetaForce::getNumParticles)
        .def("addParticle", &ATMMetaForce::addParticle)
        .def("setParticleParameters", &ATMMetaForce::setParticleParameters)
        .def("getParticleParameters", [](const ATMMetaForce& f, int i) {
            int p;
            double dx, dy, dz;
            f.getParticleParameters(i, p, dx, dy, dz);
            return py::make_tuple(p, dx, dy, dz);
        })
        .def("updateParametersInContext", &ATMMetaForce::updateParametersInContext)
        .def("getPerturbationEnergy", &ATMMetaForce::getPerturbationEnergy)
        .def_static("Lambda1", &ATMMetaForce::Lambda1)
        .def_static("Lambda2", &ATMMetaForce::Lambda2)
        .def_static("Alpha", &ATMMetaForce::Alpha)
        .def_static("U0", &ATMMetaForce::U0)
        .def_static("W0", &ATMMetaForce::W0)
        .def_static("Umax", &ATMMetaForce::Umax)
        .def_static("Ubcore", &ATMMetaForce::Ubcore)
        .def_static("Acore", &ATMMetaForce::Acore)
        .def_static("Direction", &ATMMetaForce::Direction)
        .def_static("Version", &ATMMetaForce::Version)
        .def("getDefaultLambda1", &ATMMetaForce::getDefaultLambda1)
        .def("getDefaultLambda2", &ATMMetaForce::getDefaultLambda2)
        .def("getDefaultAlpha", &ATMMetaForce::getDefaultAlpha)
        .def("getDefaultU0", &ATMMetaForce::getDefaultU0)
        .def("getDefaultW0", &ATMMetaForce::getDefaultW0)
        .def("getDefaultUmax", &ATMMetaForce::getDefaultUmax)
        .def("getDefaultUbcore", &ATMMetaForce::getDefaultUbcore)
        .def("getDefaultAcore", &ATMMetaForce::getDefaultAcore)
        .def("getDefaultDirection", &ATMMetaForce::getDefaultDirection)
        .def("getVariableForceGroups", &ATMMetaForce::getVariableForceGroups)
        .def("usesPeriodicBoundaryConditions", &ATMMetaForce::usesPeriodicBoundaryConditions)
        .def_static("cast", [](Force& f) -> ATMMetaForce& { return dynamic_cast<ATMMetaForce&>(f); },
                    py::return_value_policy::reference)
        .def_static("isinstance", [](Force& f) { return dynamic_cast<ATMMetaForce*>(&f) != nullptr; });

    // ---------------------------------------------------------------- XML
    py::class_<XmlSerializer>(m, "XmlSerializer")
        .def_static("serialize", [](const ATMMetaForce& f) {
            std::stringstream s;
            XmlSerializer::serialize<ATMMetaForce>(&f, "Force", s);
            return s.str();
        })
        .def_static("serialize", [](const State& st) {
            std::stringstream s;
            XmlSerializer::serialize<State>(&st, "State", s);
            return s.str();
        })
        // The root element's type attribute selects the class, as in OpenMM.
        .def_static("deserialize", [](const std::string& xml) -> py::object {
            std::stringstream s(xml);
            SerializationNode node;
            XmlSerializer::readXml(s, node);
            const std::string type = node.getStringProperty("type");
            const SerializationProxy& proxy = SerializationProxy::getProxy(type);
            if (type == "State")
                return py::cast(static_cast<State*>(proxy.deserialize(node)), py::return_value_policy::take_ownership);
            if (type == "ATMMetaForce")
                return py::cast(static_cast<ATMMetaForce*>(proxy.deserialize(node)),
                                py::return_value_policy::take_ownership);
            throw OpenMMException("XmlSerializer.deserialize: unsupported type " + type);
        });
}
