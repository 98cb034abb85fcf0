#ifndef OPENMM_ATMMETAFORCE_VERSION_H_
#define OPENMM_ATMMETAFORCE_VERSION_H_
// Version of the ATMMetaForce API this build implements
// (openmmapi/include/ATMMetaForceVersion.h:4 in the reference).
#define ATMMETAFORCE_VERSION "0.3.1"
#endif
