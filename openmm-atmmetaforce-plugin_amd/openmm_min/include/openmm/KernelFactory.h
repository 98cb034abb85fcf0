// Minimal OpenMM-compatible core: creates KernelImpls for a Platform.
#pragma once
#include <string>
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class ContextImpl;
class KernelImpl;
class Platform;

class OPENMM_EXPORT KernelFactory {
public:
    virtual ~KernelFactory() {}
    virtual KernelImpl* createKernelImpl(std::string name, const Platform& platform, ContextImpl& context) const = 0;
};

}  // namespace OpenMM
