#ifndef OPENMM_ATMMETAFORCE_PROXY_H_
#define OPENMM_ATMMETAFORCE_PROXY_H_
// XML proxy of ATMMetaForce, type name "ATMMetaForce", version 0
// (reference: serialization/src/ATMMetaForceProxy.cpp:8-80).
#include "internal/windowsExportATMMetaForce.h"
#include "openmm/serialization/SerializationProxy.h"

namespace ATMMetaForcePlugin {

class OPENMM_EXPORT_ATMMETAFORCE ATMMetaForceProxy : public OpenMM::SerializationProxy {
public:
    ATMMetaForceProxy();
    void serialize(const void* object, OpenMM::SerializationNode& node) const override;
    void* deserialize(const OpenMM::SerializationNode& node) const override;
};

}  // namespace ATMMetaForcePlugin

#endif
