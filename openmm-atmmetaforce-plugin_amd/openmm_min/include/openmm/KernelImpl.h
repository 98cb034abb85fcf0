// Minimal OpenMM-compatible core: a platform-specific implementation of a
// named calculation.
#pragma once
#include <string>
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class Platform;

class OPENMM_EXPORT KernelImpl {
public:
    KernelImpl(const std::string& name, const Platform& platform) : name(name), platform(&platform) {}
    virtual ~KernelImpl() {}
    const std::string& getName() const { return name; }
    const Platform& getPlatform() { return *platform; }

private:
    std::string name;
    const Platform* platform;
};

}  // namespace OpenMM
