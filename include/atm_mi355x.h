/*
 * atm_mi355x.h -- C ABI of the MI355X ATMMetaForce evaluation library
 * (libatm_mi355x.so).
 *
 * One atm_context_t owns a batch of independent ATM systems resident on one
 * GPU.  A system is a periodic box (orthorhombic, or triclinic through
 * atm_set_box_vectors) of N atoms with per-atom
 * charge, Lennard-Jones sigma/epsilon and an ATM displacement vector, plus the
 * nine ATM global parameters.  One evaluation computes, for every system,
 *
 *   U1, F1 = cutoff LJ + reaction-field energy and forces at x        (state 1)
 *   U2, F2 = the same at x + d                                        (state 2)
 *   u      = direction > 0 ? U2 - U1 : U1 - U2, soft-core mapped -> u_sc, f'
 *   W      = e0 + [alpha>0] (l2-l1)/alpha ln(1+exp(-alpha(u_sc-u0))) + l2 u_sc + w0
 *   F     += (1-sp) F_e0 + sp F_other,   sp = f' ((l2-l1)/(1+exp(-alpha(u_sc-u0))) + l1)
 *
 * which is what the reference computes with two inner OpenMM contexts
 * (openmmapi/src/ATMMetaForceImpl.cpp:90-128) and the kernel interface
 * CalcATMMetaForceKernel (openmmapi/include/ATMMetaForceKernels.h:15-59).
 *
 * Conventions
 *  - Units: nm, kJ/mol, e (elementary charge), kJ/mol/nm.
 *  - Every call returns ATM_OK (0) or a negative status; atm_last_error()
 *    returns the message of the last failure on the calling thread.  No C++
 *    exception crosses this boundary.
 *  - Host arrays are owned by the caller and only read/written during the call.
 *    Device memory is owned by the context.
 *  - A context is not thread-safe: calls on one context must not overlap,
 *    matching the single-threaded use of an OpenMM Context by the reference.
 *  - Pair forces accumulate in 64-bit fixed point (ATM_FORCE_SCALE units per
 *    kJ/mol/nm), so their sums do not depend on summation order.  Partial
 *    forces at or above a per-set limit (at most 2^27 kJ/mol/nm, lowered so
 *    that the number of terms one atom's sum can receive times the limit stays
 *    below the int64 range: the sums cannot wrap) go to an FP64 side
 *    accumulator instead.  The
 *    blended ATM force of a system is accumulated in FP64: it is the sum over
 *    evaluations since its last atm_clear_forces() (ATM_EVAL_CLEAR_FORCES
 *    clears it first), mirroring the reference's "+=" into the context force
 *    buffer (platforms/reference/src/ReferenceATMMetaForceKernels.cpp:100-110).
 */
#ifndef ATM_MI355X_H_
#define ATM_MI355X_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ATM_MI355X_ABI_VERSION 6

/* status codes */
#define ATM_OK 0
#define ATM_ERR_INVALID_ARGUMENT (-1)
#define ATM_ERR_DEVICE (-2)          /* HIP runtime failure */
#define ATM_ERR_CAPACITY (-3)        /* a neighbour structure overflowed twice */
#define ATM_ERR_STATE (-4)           /* call out of order (e.g. results before eval) */
#define ATM_ERR_UNSUPPORTED (-5)

/* fixed-point force units per kJ/mol/nm (2^24) */
#define ATM_FORCE_SCALE 16777216.0

/* Coulomb constant 1/(4 pi eps0) in kJ nm/(mol e^2), OpenMM 7.7's ONE_4PI_EPS0 */
#define ATM_ONE_4PI_EPS0 138.935456

/* atm_eval flags */
#define ATM_EVAL_FORCES        0x1  /* accumulate the hybrid ATM force */
#define ATM_EVAL_CLEAR_FORCES  0x2  /* zero the accumulated force first */
#define ATM_EVAL_PLAIN         0x4  /* reference-structured path: one system and one
                                       coordinate set per launch sequence, U1/U2 read
                                       back to the host, soft-core/softplus on the host,
                                       hybrid-force kernel with a host scalar */
#define ATM_EVAL_NO_GRAPH      0x8  /* batched path launched kernel by kernel instead
                                       of through the captured HIP graph */
#define ATM_EVAL_FP32_PAIRS    0x10 /* scalar pair function in FP32 (default: FP64 scalars
                                       from the exact FP64 r^2; coordinates, force vectors
                                       and their accumulation are FP32 in both modes) */
#define ATM_EVAL_STATE1_ONLY   0x20 /* plain cutoff LJ + reaction-field evaluation at x:
                                       energy = state1_energy = U1, forces += F1; no ATM
                                       combine (a NonbondedForce outside the ATM force) */
#define ATM_EVAL_PARAM_GRADS   0x40 /* energies (no forces) of the batched FP64 path, and in
                                       the same pass over the pairs dU1/dp and dU2/dp for
                                       the per-atom charges, sigmas and epsilons p: a
                                       following atm_backward / atm_backward_device only
                                       weights and gathers them (no second pair pass).
                                       Not with ATM_EVAL_FORCES, ATM_EVAL_FP32_PAIRS,
                                       ATM_EVAL_PLAIN or ATM_EVAL_STATE1_ONLY */

/* One system of the batch.  Box edges must satisfy cutoff <= box/2. */
typedef struct atm_system_desc {
    int32_t num_atoms;
    double box[3];          /* orthorhombic periodic box edge lengths, nm */
    double cutoff;          /* nm */
    double rf_dielectric;   /* reaction-field dielectric constant */
} atm_system_desc;

/* The nine ATM global parameters, named as in ATMMetaForce.h:151-218. */
typedef struct atm_params {
    double lambda1, lambda2, alpha, u0, w0, umax, ubcore, acore, direction;
} atm_params;

/* Per-system results of the last evaluation. */
typedef struct atm_result {
    double energy;               /* W: the ATM alchemical potential energy */
    double perturbation_energy;  /* u_sc: soft-core perturbation energy
                                    (what ATMMetaForce::getPerturbationEnergy returns) */
    double state1_energy;        /* U1 (variable-group energy at x) */
    double state2_energy;        /* U2 (variable-group energy at x + d) */
    double raw_perturbation;     /* u before the soft-core map */
    double softcore_derivative;  /* f' = du_sc/du */
    double softplus_derivative;  /* bfp = dW/du_sc */
    double hybrid_weight;        /* sp = bfp f' (weight of the non-e0 state's force) */
} atm_result;

/* Autotuned launch configuration (read back with atm_get_tuning). */
typedef struct atm_tuning {
    double column_scale;      /* column edge / (8/density)^(1/3) */
    int32_t pair_waves_per_block;
    int32_t zbins_per_column; /* z-bins per column edge */
    double best_ms;           /* measured time of the chosen configuration */
    int32_t pair_kernel;      /* 0: 8x8 cluster-pair tiles; 1-4: compacted in-cutoff pairs
                                 (FP64 pair function only) with (list entries per batch,
                                 copies of the j-force sums, drain after every test step)
                                 = (4,2,yes), (4,2,no), (4,4,yes), (8,2,yes); 5: atom-level
                                 j lists per i-cluster (j atoms within the list cutoff of
                                 the cluster's box), 8 i x 8 j lanes per step (FP64 only);
                                 6: one wave per i-cluster filtering its clusters' j atoms
                                 against its box into an LDS queue, 8 i x 8 j lanes per
                                 queued step (FP64 only; FP32 pairs run variant 0) */
    int32_t list_capacity;    /* neighbour-list entries per group to start from; 0 = sized
                                 from the density.  A list that overflows grows it (the
                                 evaluation reports ATM_ERR_CAPACITY and is re-run) */
    int32_t pair_split;       /* waves sharing one group's list in the tile kernel (variant 0
                                 and FP32 pairs): 1, 2, 4 or 8; more, shorter waves shorten
                                 the end of the launch where few waves are left */
    int32_t full_state2;      /* 0: when every system displaces at most a quarter of its atoms,
                                 x + d is evaluated as the difference from x (the pairs and
                                 exception terms with a displaced atom at both states; the
                                 other atoms do not move), so only x is binned and paired;
                                 1: x + d is binned and paired in full like x */
    int32_t pair_order;       /* the tile kernel's group order: 0 launch order; 1 every set's
                                 groups longest list first, interleaved over all sets; 2 the
                                 same over windows of 8 sets (default); 3 each set longest
                                 first, the sets one after the other.  The autotuner times
                                 all four for the tile kernel */
    int32_t reserved_;        /* 0 */
} atm_tuning;

/* Library version string and ABI number. */
const char* atm_version(void);
int32_t atm_abi_version(void);
const char* atm_last_error(void);

/* Number of visible GPUs (0 when none). */
int atm_device_count(int32_t* count);

/* Create a context holding `num_systems` systems on GPU `device`.
 * Replaces CalcATMMetaForceKernel::initialize (ATMMetaForceKernels.h:28) and the
 * inner-context setup of ATMMetaForceImpl::initialize / calcForcesAndEnergy
 * (ATMMetaForceImpl.cpp:69-101). */
int atm_create(const atm_system_desc* systems, int32_t num_systems, int32_t device,
               void** out_context);

int atm_destroy(void* context);

/* Per-atom nonbonded parameters and ATM displacements of system `sys`
 * (arrays of num_atoms; displacements num_atoms*3).  Entry i displaces atom i,
 * as the reference does (ReferenceATMMetaForceKernels.cpp:120-123).
 * Replaces CalcATMMetaForceKernel::copyParametersToContext (ATMMetaForceKernels.h:46). */
int atm_set_particles(void* context, int32_t sys, const double* charges,
                      const double* sigmas, const double* epsilons,
                      const double* displacements);

/* Energies of the other forces of the variable groups (e.g. bonded terms
 * evaluated beside this library) at x (state1) and at x + d (state2) of
 * system `sys`: added to U1 and U2 before the soft-core/softplus combine of
 * every following evaluation, until changed (0, 0 at creation).  Their forces
 * are the caller's to blend with the returned hybrid weight: F += w1 F1' +
 * w2 F2' with (w1, w2) = (1 - sp, sp) for direction > 0, (sp, 1 - sp)
 * otherwise.  The reference's inner contexts evaluate every force of the
 * variable groups (ATMMetaForceImpl.cpp:52-62, :113, :116). */
int atm_set_energy_offsets(void* context, int32_t sys, double state1, double state2);

/* Exceptions of system `sys`'s NonbondedForce: `count` atom pairs (pairs[2k],
 * pairs[2k+1]) with charge product (e^2), sigma (nm) and epsilon (kJ/mol), as
 * NonbondedForce::addException takes them; charge product 0 and epsilon 0 is an
 * exclusion.  Every listed pair leaves the cutoff sum; the others add
 * k_e qq/r + 4 eps ((sigma/r)^12 - (sigma/r)^6) at the plain distance (no
 * cutoff, no periodic image: NonbondedForce's default for exceptions) in both
 * coordinate sets.  count 0 removes them.  Rejects the call (nothing changed)
 * for an index out of range, i == j, a pair listed twice, a non-finite value or
 * a negative sigma/epsilon.  Relays out the context (a setup call: buffers and
 * graphs are rebuilt).  The reference's inner contexts evaluate clones of the
 * system's forces, exceptions included (ATMMetaForceImpl.cpp:52-62); this is
 * the exception half of what those clones carry to the nonbonded kernel. */
int atm_set_exceptions(void* context, int32_t sys, int32_t count, const int32_t* pairs,
                       const double* charge_prod, const double* sigma, const double* epsilon);

/* Displacements only (num_atoms*3), the part copyParametersToContext updates in
 * the reference (CommonATMMetaForceKernels.cpp:229-251). */
int atm_set_displacements(void* context, int32_t sys, const double* displacements);

/* Positions (num_atoms*3, nm) of system `sys`; the device path stores them in FP32.
 * Replaces the position half of CalcATMMetaForceKernel::copyState
 * (ATMMetaForceKernels.h:55; ReferenceATMMetaForceKernels.cpp:116-124). */
int atm_set_positions(void* context, int32_t sys, const double* positions);

/* Positions of every system at once from a device array (float32, 3 per atom,
 * systems concatenated in order); the copy is enqueued on the context stream. */
int atm_set_positions_device(void* context, const float* positions_dev);

/* Charges, sigmas and epsilons of every system at once from device arrays
 * (doubles, systems concatenated in order, num_atoms of the whole context
 * each), converted on the device as atm_set_particles converts host arrays;
 * the displacements are kept.  The arrays are read on the context stream
 * (the caller orders its writes before the call, e.g. by synchronizing or
 * with atm_set_stream).  Rejects the whole update (nothing changed) if any
 * entry is non-finite or a sigma/epsilon is negative.  The device path of
 * copyParametersToContext for per-atom parameters (ATMMetaForceKernels.h:46),
 * used by the autograd entry point. */
int atm_set_particles_device(void* context, const double* charges, const double* sigmas, const double* epsilons);

/* New box edges for system `sys`: an orthorhombic box (copyState copies the
 * box vectors, ReferenceATMMetaForceKernels.cpp:126-129). */
int atm_set_box(void* context, int32_t sys, const double* box3);

/* Periodic box vectors a, b, c of system `sys` (vectors[9] = ax ay az bx by bz
 * cx cy cz, nm), triclinic boxes included: the box vectors copyState copies into
 * both inner contexts (ReferenceATMMetaForceKernels.cpp:126-129).  They must be in
 * OpenMM's reduced form (ay = az = bz = 0, positive diagonal, ax >= 2|bx|,
 * ax >= 2|cx|, by >= 2|cy|) with cutoff <= half of ax, by and cz.  A pair takes
 * the image OpenMM's triclinic rule picks (shift by c to the nearest z, then by b
 * to the nearest y, then by a to the nearest x), in FP32 on the device: with the
 * positions and the box vectors on the 2^-20 nm grid every shift is exact.  A batch with a triclinic system
 * runs the 8x8 tile kernel. */
int atm_set_box_vectors(void* context, int32_t sys, const double* vectors);

/* The nine ATM global parameters of system `sys` (read by execute,
 * ReferenceATMMetaForceKernels.cpp:66-80). */
int atm_set_params(void* context, int32_t sys, const atm_params* params);

/* Evaluate every system (ATM_EVAL_* flags).  Enqueues on the context stream and
 * returns; atm_get_results/atm_get_forces synchronize.
 * Replaces ATMMetaForceImpl::calcForcesAndEnergy (ATMMetaForceImpl.cpp:90-128)
 * with CalcATMMetaForceKernel::copyState + execute (ATMMetaForceKernels.h:37-55). */
int atm_eval(void* context, int32_t flags);

/* Wait for all work enqueued on the context stream. */
int atm_synchronize(void* context);

/* Results of the last evaluation of system `sys`.
 * perturbation_energy is CalcATMMetaForceKernel::getPerturbationEnergy
 * (ATMMetaForceKernels.h:57). */
int atm_get_results(void* context, int32_t sys, atm_result* out);

/* Accumulated ATM force of system `sys` (num_atoms*3 doubles, kJ/mol/nm). */
int atm_get_forces(void* context, int32_t sys, double* forces);

/* Accumulated ATM force in fixed point (num_atoms*3, ATM_FORCE_SCALE units,
 * rounded to nearest, saturated at the int64 range). */
int atm_get_forces_fixed(void* context, int32_t sys, int64_t* forces);

/* Zero the accumulated force of every system. */
int atm_clear_forces(void* context);

/* Parity: the in-cutoff pair set of system `sys`, coordinate set `state`
 * (0: x, 1: x + d) found by the last evaluation.  Writes min(count, capacity)
 * pairs as (i, j) atom indices with i < j, sorted, and the total count. */
int atm_get_pairs(void* context, int32_t sys, int32_t state, int32_t* pairs,
                  int64_t capacity, int64_t* count);

/* Gradient of sum_s gW[s] W_s with respect to per-atom charges, sigmas and
 * epsilons (each sum_s num_atoms doubles, systems concatenated) and the
 * per-system softplus parameters (num_systems x 5 doubles in the order
 * lambda1, lambda2, alpha, u0, w0).  Recomputes the pair terms at the last
 * evaluated positions, unless that evaluation ran with ATM_EVAL_PARAM_GRADS
 * (then it only weights the stored per-state derivatives); any output pointer
 * may be NULL. */
int atm_backward(void* context, const double* grad_energy, double* d_charges,
                 double* d_sigmas, double* d_epsilons, double* d_params);

/* atm_backward with the per-atom gradients written to device arrays
 * (d_charges, d_sigmas, d_epsilons: num_atoms of the whole context each, any
 * may be NULL) on the context stream, with no host copy; grad_energy and
 * d_params are host arrays as in atm_backward.  Returns after the gradients
 * are written. */
int atm_backward_device(void* context, const double* grad_energy, double* d_charges, double* d_sigmas,
                        double* d_epsilons, double* d_params);

/* Use a caller-provided HIP stream (hipStream_t) instead of the context's own
 * (NULL restores the context stream). */
int atm_set_stream(void* context, void* hip_stream);

/* Time the candidate launch configurations on the current inputs and keep the
 * fastest (cell shape, z-binning, pair-kernel variant and waves per block; the
 * four fastest of the sweep are timed again with 4x the repeats).  A context
 * starts at column scale 1.3, 2 waves per block, 4 z-bins, pair kernel 0. */
int atm_autotune(void* context, int32_t repeats, atm_tuning* out);
int atm_get_tuning(void* context, atm_tuning* out);
int atm_set_tuning(void* context, const atm_tuning* tuning);

/* Per-stage device times (ms, averaged over `reps` batched evaluations launched
 * kernel by kernel with HIP events between stages on the context stream).
 * build = clear + binning + scan + scatter + cell sort + cluster boxes; pairs =
 * the pair kernel alone; displaced = the exception terms and, when x + d is
 * evaluated as the difference from x, the displaced-atom kernels. */
typedef struct atm_stage_times {
    double build_ms, pairlist_ms, pairs_ms, displaced_ms, energy_ms, blend_ms, total_ms;
    int32_t reps;
} atm_stage_times;
int atm_profile(void* context, int32_t flags, int32_t reps, atm_stage_times* out);

/* Counters of the last evaluation, for measurement: number of in-cutoff pairs
 * (both states, all systems) as counted by the pair kernel that ran, candidate
 * pairs in the cluster-pair tiles of the lists (a traversal of the lists), and
 * neighbour-list entries.  Any pointer may be NULL. */
int atm_get_stats(void* context, int64_t* pairs_in_cutoff, int64_t* pairs_evaluated,
                  int64_t* list_entries);

/* Per coordinate set of the last evaluation (arrays of 2 * num_systems; set
 * 2 s + state is system s at x (state 0) or x + d (state 1)): the in-cutoff
 * pairs the pair kernel evaluated (each i < j once; counted from the lane masks
 * of its exact r < rc decisions), the longest neighbour list of one group
 * (before truncation to the capacity) and the list capacity per group.  Any
 * pointer may be NULL.  The parity counterpart of the pair count is the size
 * of the FP64 oracle's pair set, which OpenMM's NonbondedForce evaluates at
 * ATMMetaForceImpl.cpp:113,116. */
int atm_get_list_stats(void* context, int64_t* pairs, int32_t* max_entries, int32_t* capacity);

/* ------------------------------------------------------------------------
 * No-allocation evaluation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).
 *
 * One ATM evaluation of one system (both coordinate sets, the combine and,
 * with ATM_EVAL_FORCES, the hybrid force) from caller-owned device arrays,
 * enqueued on `hip_stream` (hipStream_t; NULL = the null stream) on the
 * current device.  It calls no allocating routine: every intermediate buffer
 * lives in a static device array of ATM_NOALLOC_SCRATCH_MIB MiB that each GPU
 * the library runs on holds until the process ends.  The kernels and launch
 * configuration are atm_eval's defaults, so the results equal those of a
 * one-system context (atm_create + atm_eval) on the same inputs.
 *   positions      num_atoms x 3 doubles (nm), device
 *   charges, sigmas, epsilons   num_atoms doubles each, device
 *   displacements  num_atoms x 3 doubles (nm), device
 *   forces         num_atoms x 3 doubles, device: the hybrid ATM force is
 *                  added (written with ATM_EVAL_CLEAR_FORCES); may be NULL
 *                  without ATM_EVAL_FORCES
 *   result         one atm_result, device
 *   status         one int32, device: ATM_OK, ATM_ERR_INVALID_ARGUMENT (a
 *                  non-finite input or a negative sigma/epsilon) or
 *                  ATM_ERR_CAPACITY (a neighbour list outgrew its capacity,
 *                  sized from the density with 1.5x headroom); on either error
 *                  the result fields are NaN and the forces are untouched.
 * flags: ATM_EVAL_FORCES, ATM_EVAL_CLEAR_FORCES, ATM_EVAL_FP32_PAIRS.
 * Limits: num_atoms <= ATM_NOALLOC_MAX_ATOMS and the system's buffers within
 * the scratch (ATM_ERR_UNSUPPORTED otherwise, checked on the host before any
 * launch).  Every call on one GPU shares that GPU's scratch: calls that
 * overlap on one GPU (other streams, other threads) race; calls on one
 * stream are ordered and each starts from cleared state.
 * ------------------------------------------------------------------------ */
#define ATM_NOALLOC_MAX_ATOMS 32768
#define ATM_NOALLOC_SCRATCH_MIB 32
int atm_eval_noalloc(const atm_system_desc* system, const atm_params* params, const double* positions,
                     const double* charges, const double* sigmas, const double* epsilons,
                     const double* displacements, int32_t flags, double* forces, atm_result* result,
                     int32_t* status, void* hip_stream);

/* ------------------------------------------------------------------------
 * Flat-bottom restraints evaluated beside the ATM force every step (force
 * group 0 in ATM workflows, python/tests/test_abfe.py:65-79).
 *
 * Centroid restraint (replaces the CustomCentroidBondForce built by
 * ATMMetaForceUtils.addVsiteRestraintForceCMCM, python/ATMMetaForceUtils.py:79-138):
 *   c_g = sum_i w_i x_i / sum_i w_i over the members of group g (no periodic
 *   wrapping, as CustomCentroidBondForce by default), d = |c1 - offset - c2|,
 *   E = kf/2 (d - tol)^2 for d > tol, else 0 (expression at :113-114).
 * Position restraint (replaces the CustomExternalForce of
 * ATMMetaForceUtils.addPosRestraints, python/ATMMetaForceUtils.py:707-753):
 *   d = |x - x0| (minimum image in the orthorhombic box when periodic),
 *   E = kf/2 (d - tol)^2 for d > tol, else 0 (expressions at :733-736).
 * Units: nm, kJ/mol, kJ/mol/nm^2.  Forces are summed in 64-bit fixed point
 * (2^-24 kJ/mol/nm), so results do not depend on the order of the sums.
 * ------------------------------------------------------------------------ */
typedef struct atm_centroid_restraint {
    int32_t group1_first, group1_count; /* members [first, first + count) of the member arrays */
    int32_t group2_first, group2_count;
    double kf, tol;
    double offset[3];
} atm_centroid_restraint;

typedef struct atm_position_restraint {
    int32_t atom;     /* index into the positions passed to atm_restraints_eval */
    int32_t periodic; /* nonzero: minimum-image distance */
    double ref[3];
    double kf, tol;
} atm_position_restraint;

/* A restraint set on GPU `device`; its evaluations take positions of
 * `num_atoms` atoms (any indexing the caller chooses). */
int atm_restraints_create(int32_t device, void** out);
int atm_restraints_destroy(void* restraints);

/* Replace the restraints.  member_atoms/member_weights: the concatenated
 * groups the centroid restraints refer to (weights > 0; ATMMetaForceUtils
 * weights by mass). */
int atm_restraints_set(void* restraints, int32_t num_atoms, int32_t num_members,
                       const int32_t* member_atoms, const double* member_weights,
                       int32_t num_centroid, const atm_centroid_restraint* centroid,
                       int32_t num_position, const atm_position_restraint* position);

/* Energy (kJ/mol) and forces at `positions` (num_atoms x 3, nm) in a box with
 * edges `box` (nm; used by periodic position restraints).  Forces are added
 * to `forces` (num_atoms x 3) when it is not NULL. */
int atm_restraints_eval(void* restraints, const double* positions, const double box[3], double* forces,
                        double* energy);

/* ------------------------------------------------------------------------
 * Batched Langevin dynamics: every system of a context integrated together on
 * its ATM force W (its variable-group LJ + reaction-field terms through the
 * soft-core/softplus combine), one lambda-state per system -- the MD loop of
 * an ATM run (example/abfe/abfe.py:101-102, 149-160: a LangevinIntegrator
 * over the force groups {0, ATM group}) for a whole lambda schedule at once.
 * Splitting as OpenMM's LangevinMiddleIntegrator (BAOAB); FP64 positions and
 * velocities on the device; reproducible Philox4x32-10 noise keyed by `seed`.
 * ------------------------------------------------------------------------ */

/* Start: masses (amu) and positions (nm) of all atoms, systems concatenated;
 * Maxwell-Boltzmann velocities at `temperature` (K); forces of the initial
 * positions are evaluated.  Massless atoms stay fixed. */
int atm_md_init(void* context, const double* masses, const double* positions, double temperature, uint64_t seed);

/* Advance every system by `steps` steps of `dt` (ps) with friction (1/ps) at
 * `temperature` (K).  Afterwards the results (atm_get_results) and forces are
 * those of the new positions.  A neighbour-list overflow during the steps is
 * handled inside: the steps since the last check are replayed with doubled
 * capacity, bit for bit. */
int atm_md_step(void* context, int32_t steps, double dt, double friction, double temperature);

/* Positions (nm) and velocities (nm/ps) of system `sys` (num_atoms x 3 each;
 * either may be NULL) and its kinetic energy (kJ/mol; may be NULL). */
int atm_md_get(void* context, int32_t sys, double* positions, double* velocities, double* kinetic_energy);

/* ------------------------------------------------------------------------
 * PME variable group.  The ATM evaluation when the variable group is an
 * OpenMM NonbondedForce with particle-mesh Ewald, as in the reference's
 * end-to-end test (python/tests/test_abfe.py:42-43: AMBER system, PME, 1 nm
 * cutoff; the NonbondedForce is the variable group, :86-96).  Replaces the
 * inner-context evaluations of ATMMetaForceImpl::calcForcesAndEnergy
 * (openmmapi/src/ATMMetaForceImpl.cpp:110-122) for that force and the
 * combine of ReferenceATMMetaForceKernels.cpp:56-110.  All arithmetic FP64;
 * forces summed in 64-bit fixed point.
 *
 * Model (OpenMM's documented NonbondedForce PME): alpha = sqrt(-ln(2 tol))/rc
 * unless given; grid per axis ceil(2 alpha L / (3 tol^(1/5))) raised to the
 * next 2,3,5,7-smooth size (at least 6) unless given; order-5 B-splines;
 * direct space erfc(alpha r)/r Coulomb + Lorentz-Berthelot LJ for
 * minimum-image pairs with r < rc that are not exceptions; each exception
 * adds k_e qq'/r + 4 eps' [(s'/r)^12 - (s'/r)^6] at the plain (non-periodic)
 * distance and removes the reciprocal part k_e q_i q_j erf(alpha r)/r; self
 * energy; optional long-range dispersion correction.  Rectangular boxes with
 * at least three cutoff lengths per edge.
 * ------------------------------------------------------------------------ */
typedef struct atm_pme_desc {
    int32_t num_atoms;
    int32_t grid[3];               /* PME grid; 0 = from ewald_tolerance */
    double box[3];                 /* nm (initial; atm_pme_eval may pass another) */
    double cutoff;                 /* nm */
    double ewald_tolerance;        /* OpenMM's default 5e-4 */
    double alpha;                  /* 1/nm; 0 = from ewald_tolerance */
    int32_t dispersion_correction; /* nonzero: add the long-range LJ correction */
    int32_t reserved;
} atm_pme_desc;

/* Replaces the NonbondedForce clone of ATMMetaForceImpl.cpp:58-63 and the
 * inner contexts created at :91-101. */
int atm_pme_create(int32_t device, const atm_pme_desc* desc, void** out);
int atm_pme_destroy(void* pme);

/* The Ewald alpha (1/nm) and grid the force uses. */
int atm_pme_get_info(void* pme, double* alpha, int32_t grid[3]);

/* Per-particle charge (e), sigma (nm), epsilon (kJ/mol) -- NonbondedForce
 * addParticle -- and the ATM displacement d (num_atoms x 3, nm; NULL keeps
 * the current one), ATMMetaForce::addParticle / setParticleParameters
 * (ATMMetaForce.h:79-297).  charges, sigmas and epsilons all NULL: only the
 * displacements change (ATMMetaForce::updateParametersInContext). */
int atm_pme_set_particles(void* pme, const double* charges, const double* sigmas, const double* epsilons,
                          const double* displacements);

/* NonbondedForce exceptions: `num` atom pairs (2*num indices) with their
 * charge product (e^2), sigma (nm) and epsilon (kJ/mol).  Every listed pair
 * leaves the direct-space sum; a pair may appear once. */
int atm_pme_set_exceptions(void* pme, int32_t num, const int32_t* pairs, const double* charge_prod,
                           const double* sigma, const double* epsilon);

/* The nine ATM parameters (ATMMetaForceImpl.cpp:130-142 names and defaults). */
int atm_pme_set_params(void* pme, const atm_params* params);

/* One evaluation: U1 at x, U2 at x + d, combine.  positions num_atoms x 3
 * (nm); box NULL keeps the last one.  flags: 0, or ATM_EVAL_STATE1_ONLY for
 * the NonbondedForce alone at x (energy = state1_energy = U1, forces F1).
 * forces (num_atoms x 3, may be NULL) are ADDED the hybrid ATM force (F1 with
 * ATM_EVAL_STATE1_ONLY); result (may be NULL) as atm_get_results; parts (may
 * be NULL, 10 doubles): per state (x, then x + d) the direct-space,
 * exception, reciprocal, self and dispersion energies. */
int atm_pme_eval(void* pme, const double* positions, const double box[3], int32_t flags, double* forces,
                 atm_result* result, double* parts);

/* atm_set_energy_offsets for a PME NonbondedForce: added to U1 and U2 before the
 * combine of every following ATM evaluation (not to a state-1-only one's U1). */
int atm_pme_set_energy_offsets(void* pme, double state1, double state2);

/* ------------------------------------------------------------------------
 * Bonded terms of group 0 beside the ATM force (the AMBER HarmonicBondForce,
 * HarmonicAngleForce and PeriodicTorsionForce of test_abfe.py's system).
 * bond_params (r0 nm, k kJ/mol/nm^2) per bond, energy k/2 (r - r0)^2;
 * angle_params (theta0 rad, k kJ/mol/rad^2) per angle (atoms i, vertex j, k),
 * energy k/2 (theta - theta0)^2; torsion_params (periodicity, phase rad,
 * k kJ/mol) per torsion, energy k (1 + cos(n phi - phase)), IUPAC phi.
 * ------------------------------------------------------------------------ */
int atm_bonded_create(int32_t device, void** out);
int atm_bonded_destroy(void* bonded);
int atm_bonded_set(void* bonded, int32_t num_atoms, int32_t num_bonds, const int32_t* bond_atoms,
                   const double* bond_params, int32_t num_angles, const int32_t* angle_atoms,
                   const double* angle_params, int32_t num_torsions, const int32_t* torsion_atoms,
                   const double* torsion_params);
/* energies[3] = bond, angle, torsion energy; forces (may be NULL) are ADDED. */
int atm_bonded_eval(void* bonded, const double* positions, double* forces, double energies[3]);

#ifdef __cplusplus
}
#endif

#endif /* ATM_MI355X_H_ */
