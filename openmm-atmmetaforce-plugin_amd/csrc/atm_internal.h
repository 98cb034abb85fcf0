// Internal layout shared by the gfx950 kernels and the host side of the C ABI.
//
// Batch layout in HBM
//   * A context holds S systems.  System s owns atoms [atomOff_s, atomOff_s+N_s)
//     of the unsorted per-atom arrays (positions+charge, LJ parameters,
//     displacements) and of the accumulated fixed-point ATM force (SoA x|y|z).
//   * Every system has two coordinate sets ("states"): state 0 is x, state 1 is
//     x + d.  Set c = 2 s + state is binned, sorted and paired independently;
//     each set owns a slot range of the sorted arrays, a cell range, a column
//     range, a cluster range, a group range and an entry range.
//   * Sorting: atoms are binned into (column, z-bin) fine cells.  Columns tile
//     the x-y plane of the box; each column's atom count is padded to a multiple
//     of 8 with dummy slots (sidx = -1), so 8-slot clusters never straddle
//     columns.  Up to 8 consecutive clusters of one column form a group; a group
//     is the unit of work of the pair kernel (one wave64).
//   * Exceptions (NonbondedForce exclusions and scaled pairs) of system s: a
//     64-bit near mask per atom a, bit (b - a + 32) set for each excepted
//     partner b with |b - a| <= 31 and bit 32 (b = a, never a pair) flagging
//     partners further away, which are listed per atom (CSR over the batch's
//     atoms); the pair kernels drop every excepted pair.  Exception terms with
//     a nonzero charge product or epsilon are evaluated by a kernel of their own
//     (one block per set) at the plain, non-periodic distance.
//   * State difference: x + d differs from x only at the displaced atoms
//     (d != 0).  When they are few, state 1 is not binned or paired: the
//     state-0 set is, and a displaced-atom kernel evaluates every pair with a
//     displaced atom at x and at x + d (neighbours from the state-0 cells), so
//     U(x + d) = U(x) - E_d(x) + E_d(x + d) and F(x + d) = F(x) + dF, dF held in
//     atom order in the state-1 set's fixed-point sums.
#pragma once
#include <stdint.h>

namespace atm {

constexpr int kClusterSize = 8;      // atoms per cluster
constexpr int kGroupClusters = 8;    // clusters per group (one wave64 = 8 j x 8 i lanes)
constexpr double kForceScale = 16777216.0;  // 2^24 fixed-point units per kJ/mol/nm
// Upper bound of the fixed-point partial-force threshold (kJ/mol/nm).  Each set
// uses fixMax <= this: partial forces at or above it go to the FP64 side
// accumulator, and fixMax is chosen so that the at most maxTerms integer terms
// one slot can receive cannot sum past 2^63 (no wrap, by construction).
constexpr float kFixedMax = 134217728.0f;  // 2^27

struct SetDesc {
    // geometry (FP32 copies of the FP64 host values)
    float L[3], invL[3];
    float ce[2], invCe[2];   // column edges in x, y
    float ze, invZe;         // z-bin edge
    float rc, rcList2;       // cutoff and the (slightly enlarged) list-build cutoff^2
    float rc2lo, rc2hi;      // FP32 window around rc^2 where the FP64 test decides
    float krf, crf;          // reaction field constants (k_e folded into the charges)
    double rc2;              // exact FP64 cutoff^2
    double krf64, crf64;     // reaction field constants in FP64 (applied to the energy sums)
    int nx, ny, nz;          // columns in x, y; z-bins
    int Rx, Ry;              // column search radius (columns) in x, y
    int fullX, fullY;        // iterate every column once in that dimension
    int shiftOK;             // x/y image shift known from the column offset
    int sys, state;
    int natoms, atomOff;     // system atoms in the unsorted arrays
    int setAtomOff;          // this set's atoms in the per-set binned arrays
    int cellOff, colOff;
    int slotOff, slotCap;
    int clusterOff, clusterCap;
    int groupOff, groupCap;
    long long entryOff;      // first entry of group 0 of this set
    int entryCap;            // entries per group
    int alistCap;            // atom-level list entries per cluster (multiple of 8)
    long long alistOff;      // first atom-level list entry of cluster 0 of this set
    float fixMax;            // partial forces >= fixMax (kJ/mol/nm) go to the FP64 side sums
    int excCount;            // exception terms of this set's system (nonzero charge product or epsilon)
    int excOff;              // its first exception term in the batch arrays
    int ligCount;            // displaced atoms (d != 0) of the system
    int ligOff;              // their first entry in the batch's displaced-atom list
    int lxCount;             // exception terms with a displaced atom
    int lxOff;               // their first entry in the batch's list of such terms
    int tri;                 // triclinic box: vectors a = (L0, 0, 0), b = (tilt0, L1, 0),
    float tilt[3];           // c = (tilt1, tilt2, L2) (reduced form); every column is a
                             // candidate (fullX, fullY) and pairs take the image of
                             // OpenMM's triclinic rule (c, then b, then a)
};

struct SetRuntime {
    int nslots, nclusters, ngroups;
    int overflow;  // bit 0: a group's list exceeded entryCap, bit 1: an atom list exceeded
                   // alistCap; sticky until the host relays out
    int big;       // a partial force of this evaluation went to the FP64 side sums
    int maxEntries;  // largest untruncated list length of one group (this evaluation)
    unsigned long long npairs;  // in-cutoff pairs the pair kernel evaluated (this evaluation)
    int maxAtoms;  // largest untruncated atom-level list of one cluster (this evaluation)
    int pad_;
    unsigned long long lpairs[2];  // state-difference evaluation: in-cutoff pairs with a displaced
                                   // atom at x ([0]) and at x + d ([1]) (state-1 set of a system)
};

// Per-system state of the last evaluation on the device (mirrors atm_result).
struct SysResult {
    double W, usc, U1, U2, u, fp, bfp, sp;
};

struct AtmParamsDev {
    double lambda1, lambda2, alpha, u0, w0, umax, ubcore, acore, direction;
};

}  // namespace atm
