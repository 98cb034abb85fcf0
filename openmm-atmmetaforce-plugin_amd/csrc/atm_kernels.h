// Device buffer table and kernel launchers of the ATM evaluation.
#pragma once
#include <hip/hip_runtime.h>
#include <algorithm>
#include "atm_internal.h"
#include "atm_mi355x.h"

namespace atm {

// Every device array of a context, passed by value to the kernels.
struct DevPtrs {
    const SetDesc* desc;   // [2 S]
    SetRuntime* rt;        // [2 S]
    // per system, unsorted atom order
    const float4* posq;    // (x, y, z, q sqrt(k_e))
    const float4* displ;   // (dx, dy, dz, 0)
    const float2* ljp;     // (sigma/2, 2 sqrt(eps))
    const double4* prm64;  // (sigma/2, 2 sqrt(eps), q sqrt(k_e), 0) in FP64
    const AtmParamsDev* params;  // [S]
    SysResult* res;        // [S]
    double* outF;          // accumulated ATM force (kJ/mol/nm), SoA per system at 3*atomOff
    // per set, unsorted order (setAtomOff)
    float4* bpos;
    int* bcell;
    int* brank;
    int* slotOf;
    // cells and columns
    int* cellCount;
    int* cellStart;
    int* colStart;
    int* colCount;
    // sorted slots
    float4* spos;
    float2* sprm;
    double4* sprm64;
    int* sidx;             // atom of each slot in its final order (-1: padding)
    int2* sslot;           // (atom, cell) in the counting scatter's provisional order ((-1, -1): padding)
    long long* sf;         // 64-bit fixed-point pair-force sums: x of slot s at sf[s], y at
    long long sfn;         // sf[sfn + s], z at sf[2 sfn + s] (one array per component keeps
                           // the atomics of consecutive slots on shared cache lines)
    double* sd;            // FP64 side sums, (x, y, z) per slot, of partial forces beyond the
                           // fixed-point range; read and cleared only when rt[c].big is set
    unsigned long long* sdMask;  // bit per slot: its side sums are nonzero (set slot ranges
                                 // start on multiples of 64)
    // clusters
    float4* bbc;
    float4* bbh;
    // groups
    int4* ginfo;
    int* gcount;
    double* genergy;       // pair-kernel energy slots: one per group, or one per i-cluster wave (8 per group)
    uint32_t* entries;
    uint32_t* alist;       // atom-level j lists per i-cluster (pair kernel 5)
    int* acount;           // entries (whole 8-atom blocks) per i-cluster
    double* setEnergy;     // [2 S] (reference-structured path)
    // exceptions (null when the batch has none)
    const unsigned long long* xmask;  // per atom (batch order): near excepted partners, bit 32: far ones
    const int* xfarStart;  // [total atoms + 1]: far partners of atom a at xfar[xfarStart[a], xfarStart[a + 1])
    const int* xfar;       // system-local atom indices
    const int2* xpair;     // exception terms: system-local atom pair
    const double4* xprm;   // (k_e q'q, sigma, 4 epsilon, 0)
    double* xenergy;       // [2 S] exception energy per set
    const double* eoff;    // [2 S] energies of other variable-group forces added to U per set (may be null)
    // state difference (state-0 sets paired, displaced atoms evaluated at both states)
    int setStride = 1;     // 1: every set; 2: the state-0 sets only (set c = setBegin + 2 k)
    const int* ligAtoms;   // displaced atoms of every system (system-local, at SetDesc::ligOff)
    const unsigned char* ligMark;  // per atom (batch order): 1 if displaced
    double* lenergy;       // [2 x displaced atoms x kDisplacedParts]: pair energy of each at x and at x + d
    const int* lxidx;      // exception terms (batch index) with a displaced atom, at SetDesc::lxOff
    // group order of the tile kernel, longest lists first per set (k_order, run by the pair
    // launchers; the set's range at groupOff); null: launch order
    int* lptItems;
    int lptMode;           // kOrder*: how the tile kernel's launch ranks map to the groups
    int anyTri = 0;        // some system has a triclinic box (SetDesc::tri)
    // backward
    double* gq;            // [2 S sets x slots] gradient accumulators in slot order
    double* gsig;
    double* geps;
};

// Clears the cell counters and pair counters of sets setBegin + k P.setStride, k < nsets
// (maxCells: the largest cell count among them) and, when nf > 0, the doubles f[0, nf).
hipError_t launch_zero_counts(const DevPtrs& P, int setBegin, int nsets, int maxCells, double* f, long long nf,
                              hipStream_t st);
hipError_t launch_build(const DevPtrs& P, int setBegin, int nsets, int maxAtoms, int maxCells, int maxSlots,
                        hipStream_t st);
hipError_t launch_pairlist(const DevPtrs& P, int setBegin, int nsets, int maxGroups, int listWaves, hipStream_t st);
// atom-level j lists from the cluster list (needed by pair kernel variant kPairKernelAtomLists)
constexpr int kPairKernelAtomLists = 5;
hipError_t launch_alist(const DevPtrs& P, int setBegin, int nsets, int maxGroups, hipStream_t st);
// variant 0: 8x8 tile kernel; 1-4: compacted kernel (FP64 pairs only) with (list entries per batch,
// copies of the j-force sums, drain after every test step) = (4, 2, yes), (4, 2, no), (4, 4, yes),
// (8, 2, yes); 5: atom-level j lists (launch_alist first), 8 i x 8 j lanes per step
// 6: filtered atom queues (one wave per i-cluster, FP64 pairs only; FP32 pairs run variant 0)
constexpr int kPairKernelFiltered = 6;
constexpr int kPairKernelCount = 7;
// A batch with exceptions or a triclinic box runs the tile kernel whatever the
// variant asked for: only it applies the exclusion masks and the triclinic images.
inline int effective_pair_variant(const DevPtrs& P, int variant) { return (P.xmask || P.anyTri) ? 0 : variant; }
// split (1-8): the tile kernel divides each group's list between that many waves.
hipError_t launch_pairs(const DevPtrs& P, int setBegin, int nsets, int maxGroups, int waves, bool forces, bool fp64,
                        int variant, int split, hipStream_t st);
// Energy slots the pair kernel fills per group in genergy: 8 for the filtered
// kernel, `split` for the tile kernel, 1 for the others.
inline int pair_energy_stride(int variant, bool fp64, int split) {
    if (fp64 && variant == kPairKernelFiltered) return 8;
    return (variant == 0 || !fp64) ? split : 1;
}
// Exception terms of sets [setBegin, setBegin + nsets): energy into xenergy[c], with
// `forces` their forces added to the set's slot sums (call after the pair kernel
// and before launch_energy / launch_set_energy, which add xenergy when excCount > 0).
hipError_t launch_exceptions(const DevPtrs& P, int setBegin, int nsets, bool forces, hipStream_t st);
// The tile kernel's group order (atm_tuning.pair_order): launch order; every set's groups
// longest list first, interleaved over all sets / over windows of kOrderWindowSets sets;
// each set's groups longest first, the sets one after the other.
constexpr int kOrderLaunch = 0, kOrderAllSets = 1, kOrderWindows = 2, kOrderSetMajor = 3;
constexpr int kOrderWindowSets = 8;
// Length buckets of the tile kernel's group order (entries / kLptShift, capped).
constexpr int kLptBuckets = 64;
constexpr int kLptShift = 4;
// Waves per (displaced atom, state) in the displaced-atom kernel (each its own energy slot).
constexpr int kDisplacedParts = 1;
// delta: the state-difference evaluation (state-0 sets paired, displaced atoms by launch_displaced)
hipError_t launch_energy(const DevPtrs& P, int sysBegin, int nsys, int estride, hipStream_t st, bool delta = false);
// withOffsets: add eoff[c] (an ATM evaluation; not a plain state-1 one)
hipError_t launch_set_energy(const DevPtrs& P, int c, int estride, bool withOffsets, hipStream_t st);
// Backward pass of the pair terms (the tile kernel's backward mode): adds wset[c] dE/d(q', h, e')
// of every in-cutoff pair of sets [setBegin, setBegin + nsets) to gq, gsig, geps (slot order).
// wset null: unit weights, and the pass also records the energies and pair counts (an evaluation
// with ATM_EVAL_PARAM_GRADS).
hipError_t launch_backward_pairs(const DevPtrs& P, int setBegin, int nsets, int maxGroups, const double* wset,
                                 int split, hipStream_t st);
// clear: write the blended force instead of adding it to outF.
hipError_t launch_blend(const DevPtrs& P, int sysBegin, int nsys, int maxAtoms, int hostMode, double sp, double dir,
                        bool clear, hipStream_t st, bool delta = false);
// State difference: the pairs with a displaced atom of systems [sysBegin, sysBegin + nsys) at
// x and at x + d (after the state-0 sets' build; maxLig: most displaced atoms of one system),
// and the exception terms with a displaced atom (their difference).
// grads: the backward mode (energies, counts and the parameter derivatives of the displaced pairs,
// x + d minus x, into gq/gsig/geps of the x + d sets indexed by atom; no forces)
hipError_t launch_displaced(const DevPtrs& P, int sysBegin, int nsys, int maxLig, bool forces, hipStream_t st,
                            bool grads = false);
hipError_t launch_displaced_exceptions(const DevPtrs& P, int sysBegin, int nsys, bool forces, hipStream_t st);
// The pairs the displaced-atom kernel evaluates for system `sys` at `state` (lower index first),
// appended to out[ctr[0]++] while below cap.
hipError_t launch_displaced_dump(const DevPtrs& P, int sys, int state, int nLig, int2* out, long long cap,
                                 unsigned long long* ctr, hipStream_t st);
hipError_t launch_fill_i64(long long* p, long long n, hipStream_t st);
hipError_t launch_fill_zero(void* p, long long n8, hipStream_t st);
hipError_t launch_pairs_dump(const DevPtrs& P, int c, int maxGroups, int2* out, long long cap,
                             unsigned long long* counters, hipStream_t st);

// Flat-bottom restraints (atm_restraints.hip).
struct RestraintPtrs {
    const double* pos;           // atoms x 3
    unsigned long long* fsum;    // atoms x 3, 64-bit fixed point
    const int* members;
    const double* weights;
    const atm_centroid_restraint* cent;
    const atm_position_restraint* posr;
    double* energy;              // one slot per restraint: centroid ones first
    int ncent, npos;
    double box[3];
};
hipError_t launch_restraints(const RestraintPtrs& R, hipStream_t st);

// Batched Langevin dynamics (atm_md.hip): FP64 positions/velocities, one per atom of the context.
struct MDPtrs {
    double* x;          // total atoms x 3
    double* v;          // total atoms x 3
    const double* invm; // total atoms; 0 = not integrated
    float4* posq;       // the FP32 positions (+ charge) the pair kernels read
};
hipError_t launch_md_step(const DevPtrs& P, const MDPtrs& M, int nsys, int maxAtoms, double dt, double a, double kT,
                          uint64_t seed, uint64_t step, hipStream_t st);
hipError_t launch_md_init(const DevPtrs& P, const MDPtrs& M, long long total, double kT, uint64_t seed, hipStream_t st);
hipError_t launch_md_posq(const DevPtrs& P, const MDPtrs& M, long long total, hipStream_t st);

}  // namespace atm
