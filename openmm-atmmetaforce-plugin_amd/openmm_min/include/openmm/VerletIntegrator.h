#pragma once
#include "openmm/Integrator.h"
