// Minimal OpenMM-compatible core: a compute platform with named kernel
// factories, a global registry and plugin loading (registerPlatforms /
// registerKernelFactories entry points of shared libraries).
#pragma once
#include <map>
#include <string>
#include <vector>
#include "openmm/Kernel.h"
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class Context;
class ContextImpl;
class KernelFactory;

class OPENMM_EXPORT Platform {
public:
    virtual ~Platform();
    virtual const std::string& getName() const = 0;
    virtual double getSpeed() const = 0;
    virtual bool supportsDoublePrecision() const = 0;
    const std::vector<std::string>& getPropertyNames() const { return platformProperties; }
    virtual const std::string& getPropertyValue(const Context& context, const std::string& property) const;
    void setPropertyDefaultValue(const std::string& property, const std::string& value);
    const std::string& getPropertyDefaultValue(const std::string& property) const;
    virtual void contextCreated(ContextImpl& context, const std::map<std::string, std::string>& properties) const {}
    virtual void contextDestroyed(ContextImpl& context) const {}
    // The Platform takes ownership of the factory.
    void registerKernelFactory(const std::string& name, KernelFactory* factory);
    bool supportsKernels(const std::vector<std::string>& kernelNames) const;
    Kernel createKernel(const std::string& name, ContextImpl& context) const;
    static void registerPlatform(Platform* platform);
    static int getNumPlatforms();
    static Platform& getPlatform(int index);
    static Platform& getPlatformByName(const std::string& name);
    // dlopen()s every shared library in `directory` and calls its
    // registerPlatforms() then registerKernelFactories(); returns the libraries loaded.
    static std::vector<std::string> loadPluginsFromDirectory(const std::string& directory);
    static std::vector<std::string> getPluginLoadFailures();
    static std::string getDefaultPluginsDirectory();
    static const std::string& getOpenMMVersion();

protected:
    std::vector<std::string> platformProperties;

private:
    std::map<std::string, KernelFactory*> kernelFactories;
    std::map<std::string, std::string> defaultProperties;
};

}  // namespace OpenMM
