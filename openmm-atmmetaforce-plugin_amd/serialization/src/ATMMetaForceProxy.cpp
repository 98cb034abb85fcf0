// Node layout (the reference's, ATMMetaForceProxy.cpp:11-40):
//   attributes version=0, forceGroup, name, lambda1, lambda2, alpha, u0, w0,
//   uMax, ubCore, aCore, direction;
//   <VariableForceGroups> with one <Parameter group=".."/> per group;
//   <Particles> with one <Particle particle=".." dx=".." dy=".." dz=".."/> per particle.
#include "ATMMetaForceProxy.h"
#include <memory>
#include <vector>
#include "ATMMetaForce.h"
#include "openmm/OpenMMException.h"

using namespace ATMMetaForcePlugin;
using OpenMM::SerializationNode;

ATMMetaForceProxy::ATMMetaForceProxy() : OpenMM::SerializationProxy("ATMMetaForce") {}

void ATMMetaForceProxy::serialize(const void* object, SerializationNode& node) const {
    const ATMMetaForce& f = *reinterpret_cast<const ATMMetaForce*>(object);
    node.setIntProperty("version", 0);
    node.setIntProperty("forceGroup", f.getForceGroup());
    node.setStringProperty("name", f.getName());
    node.setDoubleProperty("lambda1", f.getDefaultLambda1());
    node.setDoubleProperty("lambda2", f.getDefaultLambda2());
    node.setDoubleProperty("alpha", f.getDefaultAlpha());
    node.setDoubleProperty("u0", f.getDefaultU0());
    node.setDoubleProperty("w0", f.getDefaultW0());
    node.setDoubleProperty("uMax", f.getDefaultUmax());
    node.setDoubleProperty("ubCore", f.getDefaultUbcore());
    node.setDoubleProperty("aCore", f.getDefaultAcore());
    node.setDoubleProperty("direction", f.getDefaultDirection());
    SerializationNode& groups = node.createChildNode("VariableForceGroups");
    for (int g : f.getVariableForceGroups()) groups.createChildNode("Parameter").setIntProperty("group", g);
    SerializationNode& parts = node.createChildNode("Particles");
    for (int i = 0; i < f.getNumParticles(); ++i) {
        int p;
        double dx, dy, dz;
        f.getParticleParameters(i, p, dx, dy, dz);
        parts.createChildNode("Particle")
            .setIntProperty("particle", p)
            .setDoubleProperty("dx", dx)
            .setDoubleProperty("dy", dy)
            .setDoubleProperty("dz", dz);
    }
}

void* ATMMetaForceProxy::deserialize(const SerializationNode& node) const {
    if (node.getIntProperty("version") != 0) throw OpenMM::OpenMMException("Unsupported version");
    std::vector<int> groups;
    for (const SerializationNode& c : node.getChildNode("VariableForceGroups").getChildren())
        groups.push_back(c.getIntProperty("group"));
    std::unique_ptr<ATMMetaForce> f(new ATMMetaForce(
        node.getDoubleProperty("lambda1"), node.getDoubleProperty("lambda2"), node.getDoubleProperty("alpha"),
        node.getDoubleProperty("u0"), node.getDoubleProperty("w0"), node.getDoubleProperty("uMax"),
        node.getDoubleProperty("ubCore"), node.getDoubleProperty("aCore"), node.getDoubleProperty("direction"), groups));
    f->setForceGroup(node.getIntProperty("forceGroup", 0));
    f->setName(node.getStringProperty("name", f->getName()));
    for (const SerializationNode& c : node.getChildNode("Particles").getChildren())
        f->addParticle(c.getIntProperty("particle"), c.getDoubleProperty("dx"), c.getDoubleProperty("dy"),
                       c.getDoubleProperty("dz"));
    return f.release();
}
