// Minimal OpenMM-compatible core: Platform registry, kernel factories and
// plugin loading.
#include <dirent.h>
#include <dlfcn.h>

#include <algorithm>
#include <cstdlib>

#include "openmm/Context.h"
#include "openmm/KernelFactory.h"
#include "openmm/OpenMMException.h"
#include "openmm/Platform.h"

namespace OpenMM {

static std::vector<Platform*>& registry() {
    static std::vector<Platform*> platforms;
    return platforms;
}

static std::vector<std::string>& loadFailures() {
    static std::vector<std::string> failures;
    return failures;
}

Platform::~Platform() {
    std::vector<KernelFactory*> unique;
    for (auto& kv : kernelFactories)
        if (std::find(unique.begin(), unique.end(), kv.second) == unique.end()) unique.push_back(kv.second);
    for (KernelFactory* f : unique) delete f;
}

const std::string& Platform::getPropertyValue(const Context& context, const std::string& property) const {
    (void)context;
    return getPropertyDefaultValue(property);
}

void Platform::setPropertyDefaultValue(const std::string& property, const std::string& value) {
    if (std::find(platformProperties.begin(), platformProperties.end(), property) == platformProperties.end())
        throw OpenMMException("setPropertyDefaultValue: Illegal property name: " + property);
    defaultProperties[property] = value;
}

const std::string& Platform::getPropertyDefaultValue(const std::string& property) const {
    auto it = defaultProperties.find(property);
    if (it == defaultProperties.end()) throw OpenMMException("getPropertyDefaultValue: Illegal property name: " + property);
    return it->second;
}

void Platform::registerKernelFactory(const std::string& name, KernelFactory* factory) { kernelFactories[name] = factory; }

bool Platform::supportsKernels(const std::vector<std::string>& kernelNames) const {
    for (const std::string& n : kernelNames)
        if (kernelFactories.find(n) == kernelFactories.end()) return false;
    return true;
}

Kernel Platform::createKernel(const std::string& name, ContextImpl& context) const {
    auto it = kernelFactories.find(name);
    if (it == kernelFactories.end())
        throw OpenMMException("Called createKernel() on a Platform which does not support the requested kernel: " + name);
    return Kernel(it->second->createKernelImpl(name, *this, context));
}

void Platform::registerPlatform(Platform* platform) {
    for (Platform* p : registry())
        if (p->getName() == platform->getName()) return;  // first registration wins
    registry().push_back(platform);
}

int Platform::getNumPlatforms() { return (int)registry().size(); }

Platform& Platform::getPlatform(int index) {
    if (index < 0 || index >= (int)registry().size()) throw OpenMMException("Invalid platform index");
    return *registry()[index];
}

Platform& Platform::getPlatformByName(const std::string& name) {
    for (Platform* p : registry())
        if (p->getName() == name) return *p;
    throw OpenMMException("There is no registered Platform called \"" + name + "\"");
}

std::vector<std::string> Platform::loadPluginsFromDirectory(const std::string& directory) {
    std::vector<std::string> loaded;
    DIR* dir = opendir(directory.c_str());
    if (!dir) return loaded;
    std::vector<std::string> files;
    while (dirent* e = readdir(dir)) {
        std::string f = e->d_name;
        if (f.size() > 3 && f.compare(f.size() - 3, 3, ".so") == 0) files.push_back(directory + "/" + f);
    }
    closedir(dir);
    std::sort(files.begin(), files.end());
    std::vector<void*> handles;
    for (const std::string& f : files) {
        void* h = dlopen(f.c_str(), RTLD_LAZY | RTLD_GLOBAL);
        if (!h) {
            loadFailures().push_back(dlerror());
            continue;
        }
        handles.push_back(h);
        loaded.push_back(f);
    }
    for (void* h : handles)
        if (auto fn = (void (*)())dlsym(h, "registerPlatforms")) fn();
    for (void* h : handles)
        if (auto fn = (void (*)())dlsym(h, "registerKernelFactories")) fn();
    return loaded;
}

std::vector<std::string> Platform::getPluginLoadFailures() { return loadFailures(); }

std::string Platform::getDefaultPluginsDirectory() {
    const char* d = std::getenv("OPENMM_PLUGIN_DIR");
    return d ? d : "";
}

const std::string& Platform::getOpenMMVersion() {
    static const std::string v = "7.7 (minimal compatible core)";
    return v;
}

}  // namespace OpenMM
