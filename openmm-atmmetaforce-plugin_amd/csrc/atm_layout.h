// Host-side layout of a coordinate set (cells, columns, clusters, groups and
// list capacities), shared by the context API and the no-allocation entry point.
#pragma once
#include "atm_internal.h"
#include "../../include/atm_mi355x.h"

namespace atm {

// Launch configuration.  Defaults: the autotuner's usual choice on a
// 22 x 23k-atom batch.
struct Tuning {
    double colScale = 1.3;     // column edge over the edge of a cube holding 8 atoms
    int pairWaves = 2;         // waves per pair-kernel block
    int zbinsPerColumn = 4;    // fine z-bins per column edge
    int pairKernel = 0;        // pair-kernel variant (launch_pairs)
    int listCapacity = 0;      // list entries per group to start from (0: sized from the density)
    int pairSplit = 1;         // waves sharing one group's list in the tile kernel (1, 2, 4 or 8)
    int fullState2 = 0;        // 1: bin and pair the x + d set of every system (no state difference)
    int pairOrder = 2;         // the tile kernel's group order (kOrder* in atm_kernels.h)
};

// Throws AtmError(ATM_ERR_INVALID_ARGUMENT) for an unusable system description.
void validate_desc(const atm_system_desc& d, int s);

// Geometry and capacities of coordinate set `state` of system `s`; the offsets
// into the batch arrays are left zero (the caller places the set).
// tilt: (b_x, c_x, c_y) of a triclinic box in reduced form, or null / all zero for an
// orthorhombic one.
SetDesc make_set(const atm_system_desc& d, const Tuning& tu, int s, int state, const double* tilt = nullptr);

// Largest partial force (kJ/mol/nm) that enters a set's 64-bit fixed-point sums
// when each slot may also receive extraTerms exception terms.
float fixed_point_limit(const SetDesc& D, int extraTerms = 0);

}  // namespace atm
