#ifndef OPENMM_WINDOW_EXPORT_ATMMETAFORCE_H_
#define OPENMM_WINDOW_EXPORT_ATMMETAFORCE_H_
// Export macro of the ATMMetaForce library (Linux build: default visibility).
#define OPENMM_EXPORT_ATMMETAFORCE __attribute__((visibility("default")))
#endif
