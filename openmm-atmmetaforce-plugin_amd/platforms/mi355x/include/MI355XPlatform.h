#ifndef ATM_MI355X_PLATFORM_H_
#define ATM_MI355X_PLATFORM_H_
// The "MI355X" platform: evaluates NonbondedForce (CutoffPeriodic, reaction
// field) and ATMMetaForce through libatm_mi355x.so on one AMD Instinct MI355X.
//
// Properties:
//   DeviceIndex    GPU index (default "0")
//   Precision      "mixed" (default) or "double": FP64 scalar pair function;
//                  "single": FP32 scalar pair function.  Coordinates, force
//                  vectors and their accumulation are FP32, energies FP64 and
//                  forces 64-bit fixed point in every mode.
//   ATMEvaluation  "fused" (default): x and x + d evaluated in one captured
//                  HIP graph, combine on the GPU; "plain": the reference's
//                  structure -- one launch sequence per state with host energy
//                  read-backs and the soft-core/softplus on the host.
#include <map>
#include <string>
#include "openmm/Platform.h"

namespace ATMMetaForcePlugin {

class MI355XPlatform : public OpenMM::Platform {
public:
    struct PlatformData {
        int device = 0;
        bool fp32Pairs = false;
        bool plainATM = false;
        std::map<std::string, std::string> properties;
    };
    MI355XPlatform();
    const std::string& getName() const override {
        static const std::string name = "MI355X";
        return name;
    }
    double getSpeed() const override { return 100.0; }
    bool supportsDoublePrecision() const override { return true; }
    const std::string& getPropertyValue(const OpenMM::Context& context, const std::string& property) const override;
    void contextCreated(OpenMM::ContextImpl& context, const std::map<std::string, std::string>& properties) const override;
    void contextDestroyed(OpenMM::ContextImpl& context) const override;
    static PlatformData& data(OpenMM::ContextImpl& context);
};

}  // namespace ATMMetaForcePlugin

#endif
