// Minimal OpenMM-compatible core: shared handle to a KernelImpl.
#pragma once
#include <memory>
#include <string>
#include "openmm/KernelImpl.h"
#include "openmm/OpenMMException.h"

namespace OpenMM {

class Kernel {
public:
    Kernel() {}
    explicit Kernel(KernelImpl* impl) : impl(impl) {}
    const std::string& getName() const { return impl->getName(); }
    KernelImpl& getImpl() { return *impl; }
    template <class T> T& getAs() {
        T* t = dynamic_cast<T*>(impl.get());
        if (!t) throw OpenMMException("Kernel has the wrong type");
        return *t;
    }
    template <class T> const T& getAs() const {
        const T* t = dynamic_cast<const T*>(impl.get());
        if (!t) throw OpenMMException("Kernel has the wrong type");
        return *t;
    }

private:
    std::shared_ptr<KernelImpl> impl;
};

}  // namespace OpenMM
