// Minimal OpenMM-compatible core: MonteCarloBarostat as a holder of its two
// global parameters (MonteCarloPressure in bar, MonteCarloTemperature in K).
// ATM workflows keep one in the System with a frequency larger than the run
// (example/abfe/abfe.py:96-98), so that State checkpoints carry and restore
// both parameters.  Volume moves are not implemented: reaching `frequency`
// steps raises OpenMMException.
#pragma once
#include <string>
#include "openmm/Force.h"
#include "openmm/internal/windowsExport.h"

namespace OpenMM {

class OPENMM_EXPORT MonteCarloBarostat : public Force {
public:
    static const std::string& Pressure() {
        static const std::string key("MonteCarloPressure");
        return key;
    }
    static const std::string& Temperature() {
        static const std::string key("MonteCarloTemperature");
        return key;
    }
    MonteCarloBarostat(double defaultPressure, double defaultTemperature, int frequency = 25);
    double getDefaultPressure() const { return pressure; }
    void setDefaultPressure(double p) { pressure = p; }
    double getDefaultTemperature() const { return temperature; }
    void setDefaultTemperature(double t) { temperature = t; }
    int getFrequency() const { return frequency; }
    void setFrequency(int freq);
    int getRandomNumberSeed() const { return seed; }
    void setRandomNumberSeed(int s) { seed = s; }
    bool usesPeriodicBoundaryConditions() const override { return false; }

protected:
    ForceImpl* createImpl() const override;
    const char* defaultName() const override { return "MonteCarloBarostat"; }

private:
    double pressure, temperature;
    int frequency, seed;
};

}  // namespace OpenMM
