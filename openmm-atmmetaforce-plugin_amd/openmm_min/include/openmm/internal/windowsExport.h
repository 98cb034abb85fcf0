// Minimal OpenMM-compatible core: symbol export macro (Linux build).
#pragma once
#define OPENMM_EXPORT __attribute__((visibility("default")))
