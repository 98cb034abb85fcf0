// Round trip of ATMMetaForce through XmlSerializer with the values of the
// reference's serialization test (serialization/tests/TestSerializeATMMetaForce.cpp:13-24,
// particles :40-41): every field must come back exactly.
#include <iostream>
#include <sstream>
#include <vector>
#include "ATMMetaForce.h"
#include "openmm/internal/AssertionUtilities.h"
#include "openmm/serialization/XmlSerializer.h"

using namespace ATMMetaForcePlugin;
using namespace OpenMM;

static void roundTrip() {
    const std::vector<int> groups = {1};
    ATMMetaForce force(0.0, 0.1, 0.25, 0.5, 0.6, 200.0, 100.0, 0.07, 0.0, groups);
    force.setForceGroup(30);
    force.setName("MyATMMetaForce");
    force.addParticle(0, 0.1, 0.2, 0.3);
    force.addParticle(2, 0.4, 0.5, 0.6);

    std::stringstream xml;
    XmlSerializer::serialize<ATMMetaForce>(&force, "Force", xml);
    ATMMetaForce* copy = XmlSerializer::deserialize<ATMMetaForce>(xml);

    ASSERT_EQUAL(30, copy->getForceGroup());
    ASSERT_EQUAL(std::string("MyATMMetaForce"), copy->getName());
    ASSERT_EQUAL(0.0, copy->getDefaultLambda1());
    ASSERT_EQUAL(0.1, copy->getDefaultLambda2());
    ASSERT_EQUAL(0.25, copy->getDefaultAlpha());
    ASSERT_EQUAL(0.5, copy->getDefaultU0());
    ASSERT_EQUAL(0.6, copy->getDefaultW0());
    ASSERT_EQUAL(200.0, copy->getDefaultUmax());
    ASSERT_EQUAL(100.0, copy->getDefaultUbcore());
    ASSERT_EQUAL(0.07, copy->getDefaultAcore());
    ASSERT_EQUAL(0.0, copy->getDefaultDirection());
    ASSERT(groups == copy->getVariableForceGroups());
    ASSERT_EQUAL(2, copy->getNumParticles());
    int p;
    double dx, dy, dz;
    copy->getParticleParameters(0, p, dx, dy, dz);
    ASSERT_EQUAL(0, p);
    ASSERT_EQUAL(0.1, dx);
    ASSERT_EQUAL(0.2, dy);
    ASSERT_EQUAL(0.3, dz);
    copy->getParticleParameters(1, p, dx, dy, dz);
    ASSERT_EQUAL(2, p);
    ASSERT_EQUAL(0.4, dx);
    ASSERT_EQUAL(0.5, dy);
    ASSERT_EQUAL(0.6, dz);
    delete copy;
}

static void badVersion() {
    std::stringstream xml("<?xml version=\"1.0\" ?>\n<Force type=\"ATMMetaForce\" version=\"1\"/>\n");
    bool threw = false;
    try {
        delete XmlSerializer::deserialize<ATMMetaForce>(xml);
    } catch (const OpenMMException& e) {
        threw = std::string(e.what()) == "Unsupported version";
    }
    ASSERT(threw);
}

int main() {
    try {
        roundTrip();
        badVersion();
    } catch (const std::exception& e) {
        std::cout << "exception: " << e.what() << std::endl;
        return 1;
    }
    std::cout << "Done" << std::endl;
    return 0;
}
