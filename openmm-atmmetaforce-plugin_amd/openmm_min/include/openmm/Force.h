// Minimal OpenMM-compatible core: base class of every force term.
#pragma once
#include <string>
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class Context;
class ContextImpl;
class ForceImpl;

class OPENMM_EXPORT Force {
public:
    Force() : forceGroup(0) {}
    virtual ~Force() {}
    int getForceGroup() const { return forceGroup; }
    // Force groups are 0..31: a context evaluates subsets of them by bit mask.
    void setForceGroup(int group);
    const std::string& getName() const;
    void setName(const std::string& n) { name = n; }
    virtual bool usesPeriodicBoundaryConditions() const = 0;

protected:
    friend class ContextImpl;
    virtual ForceImpl* createImpl() const = 0;
    // Default name when none was set (OpenMM uses the class name).
    virtual const char* defaultName() const { return "Force"; }
    ForceImpl& getImplInContext(Context& context);
    const ForceImpl& getImplInContext(const Context& context) const;
    ContextImpl& getContextImpl(Context& context);

private:
    int forceGroup;
    std::string name;
    mutable std::string defaultNameCache;
};

}  // namespace OpenMM
