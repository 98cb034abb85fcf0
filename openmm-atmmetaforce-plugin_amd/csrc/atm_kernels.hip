// gfx950 kernels of the ATM evaluation.
//
// Per evaluation, for every coordinate set c = 2*system + state of the batch:
//   k_zero_counts  clear fine-cell counters, the per-set pair counters and
//                  (batched path) the accumulated ATM forces
//   k_bin          x (+ d for state 1), wrap into the box, fine-cell id and rank
//   k_scan         per-set column/cell scans, dummy padding, group table
//   k_scatter_idx  counting-sort scatter of (atom, cell) into provisional slots
//   k_gather_bbox  order each fine cell by atom index (deterministic results),
//                  gather positions/parameters into cluster order, clear the
//                  set's force accumulators, per-cluster bounding boxes
//   k_pairlist     per group of 8 clusters: j-clusters (from the group's
//                  z-window) whose boxes are within the cutoff of at least one
//                  i-cluster (half shell, i <= j)
//   k_pairs        LJ + reaction-field pair kernel, one wave64 per group,
//                  lanes = 8 j-atoms x 8 i-atoms, FP32 geometry, FP64 pair
//                  function (or FP32), FP64 energy, 64-bit fixed-point forces;
//                  k_pairs_f: one wave per i-cluster over its atom-filtered j
//                  atoms
// and per system:
//   k_energy       U1, U2 (fixed-order FP64 sums), soft-core, softplus -> W, sp
//   k_blend        F += (1-sp) F1 + sp F2 (direction-aware), gathered back to
//                  the system's atom order
//
// The ATM math follows platforms/reference/src/ReferenceATMMetaForceKernels.cpp
// (SoftCoreF :26-37, execute :56-113); the pair term is OpenMM's NonbondedForce
// with CutoffPeriodic + reaction field (U = 4 eps [(s/r)^12 - (s/r)^6]
// + k_e q_i q_j (1/r + k_rf r^2 - c_rf), pairs with r < rc under minimum image).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>
#include "atm_internal.h"
#include "atm_kernels.h"
#include "atm_math.h"

namespace atm {

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ float wrap_coord(float x, float L, float invL) {
    // x - L floor(x/L) is exact on the 2^-20 grid (single rounding of an
    // exactly representable result); one correction step fixes floor() misses.
    float w = fmaf(-L, floorf(x * invL), x);
    if (w < 0.0f) w += L;
    if (w >= L) w -= L;
    return w;
}

__device__ __forceinline__ float min_image(float d, float L, float invL) {
    return fmaf(-L, rintf(d * invL), d);
}

// Triclinic box in OpenMM's reduced form (a = (L0, 0, 0), b = (tilt0, L1, 0),
// c = (tilt1, tilt2, L2)): wrap a position into the brick [0, L0) x [0, L1) x
// [0, L2), which tiles space under the lattice -- by c to 0 <= z < L2, then by b
// to 0 <= y < L1, then by a.  Each step is exact on the 2^-20 grid below 16 nm.
__device__ __forceinline__ void wrap_tri(const SetDesc& D, float4& p) {
    float n = floorf(p.z * D.invL[2]);
    const float z = fmaf(-D.L[2], n, p.z);
    n += z < 0.f ? -1.f : (z >= D.L[2] ? 1.f : 0.f);
    p.z = fmaf(-D.L[2], n, p.z);
    p.y = fmaf(-D.tilt[2], n, p.y);
    p.x = fmaf(-D.tilt[1], n, p.x);
    n = floorf(p.y * D.invL[1]);
    const float y = fmaf(-D.L[1], n, p.y);
    n += y < 0.f ? -1.f : (y >= D.L[1] ? 1.f : 0.f);
    p.y = fmaf(-D.L[1], n, p.y);
    p.x = fmaf(-D.tilt[0], n, p.x);
    p.x = wrap_coord(p.x, D.L[0], D.invL[0]);
}

// The x and y parts of OpenMM's triclinic image of a pair difference (shift by c
// to the nearest z, then by b to the nearest y, then by a to the nearest x); the
// caller takes dz = min_image(dz, L2), which applies the same c shift to z.
// The b and a steps of the rule (what is left when the c image is n_c = 0).
__device__ __forceinline__ void tri_image_ab(const SetDesc& D, float& dx, float& dy) {
    const float nb = rintf(dy * D.invL[1]);
    dy = fmaf(-D.L[1], nb, dy);
    dx = fmaf(-D.tilt[0], nb, dx);
    dx = min_image(dx, D.L[0], D.invL[0]);
}

__device__ __forceinline__ void tri_image_xy(const SetDesc& D, float& dx, float& dy, float dz) {
    const float nc = rintf(dz * D.invL[2]);
    dy = fmaf(-D.tilt[2], nc, dy);
    dx = fmaf(-D.tilt[1], nc, dx);
    tri_image_ab(D, dx, dy);
}

// The image a pair kernel takes for a difference vector: explicit column shifts
// (shiftOK) leave x and y as they are; otherwise the orthorhombic minimum image or
// the triclinic rule; z always to its nearest image.
__device__ __forceinline__ void pair_image(const SetDesc& D, float& dx, float& dy, float& dz) {
    if (!D.shiftOK) {
        if (D.tri) {
            tri_image_xy(D, dx, dy, dz);
        } else {
            dx = min_image(dx, D.L[0], D.invL[0]);
            dy = min_image(dy, D.L[1], D.invL[1]);
        }
    }
    dz = min_image(dz, D.L[2], D.invL[2]);
}

// Two boxes with centre difference d and summed half-extents h hold points closer
// than sqrt(r2) under some lattice image n_a a + n_b b + n_c c of a triclinic box
// (every image whose shifted box can reach is tested, so the pair kernel's image
// of any pair inside the cutoff is among them).
__device__ __forceinline__ bool tri_boxes_near(const SetDesc& D, float dx, float dy, float dz, float hx, float hy,
                                               float hz, float r, float r2) {
    const int c0 = (int)floorf((dz - hz - r) * D.invL[2]), c1 = (int)ceilf((dz + hz + r) * D.invL[2]);
    for (int nc = c0; nc <= c1; ++nc) {
        const float az = fmaxf(0.f, fabsf(fmaf(-D.L[2], (float)nc, dz)) - hz);
        const float remz = r2 - az * az;
        if (remz <= 0.f) continue;
        const float dy1 = fmaf(-D.tilt[2], (float)nc, dy), dx1 = fmaf(-D.tilt[1], (float)nc, dx);
        const int b0 = (int)floorf((dy1 - hy - r) * D.invL[1]), b1 = (int)ceilf((dy1 + hy + r) * D.invL[1]);
        for (int nb = b0; nb <= b1; ++nb) {
            const float ay = fmaxf(0.f, fabsf(fmaf(-D.L[1], (float)nb, dy1)) - hy);
            const float remy = remz - ay * ay;
            if (remy <= 0.f) continue;
            const float dx2 = fmaf(-D.tilt[0], (float)nb, dx1);
            const int a0 = (int)floorf((dx2 - hx - r) * D.invL[0]), a1 = (int)ceilf((dx2 + hx + r) * D.invL[0]);
            for (int na = a0; na <= a1; ++na) {
                const float ax = fmaxf(0.f, fabsf(fmaf(-D.L[0], (float)na, dx2)) - hx);
                if (ax * ax < remy) return true;
            }
        }
    }
    return false;
}

// A column's x-y rectangle and a group's x-y box (centre difference d, summed
// half-extents h) come closer than r in x-y under some image n_a a + n_b b +
// n_c c of a triclinic box with |n_c| <= ncSpan, n_c = 0 excluded when skip0
// (z is left to the cluster test).
__device__ __forceinline__ bool tri_rect_near(const SetDesc& D, float dx, float dy, float hx, float hy, int ncSpan,
                                              float r, float r2, bool skip0 = false) {
    for (int nc = -ncSpan; nc <= ncSpan; ++nc) {
        if (skip0 && nc == 0) continue;
        const float dy1 = fmaf(-D.tilt[2], (float)nc, dy), dx1 = fmaf(-D.tilt[1], (float)nc, dx);
        const int b0 = (int)floorf((dy1 - hy - r) * D.invL[1]), b1 = (int)ceilf((dy1 + hy + r) * D.invL[1]);
        for (int nb = b0; nb <= b1; ++nb) {
            const float ay = fmaxf(0.f, fabsf(fmaf(-D.L[1], (float)nb, dy1)) - hy);
            const float remy = r2 - ay * ay;
            if (remy <= 0.f) continue;
            const float dx2 = fmaf(-D.tilt[0], (float)nb, dx1);
            const int a0 = (int)floorf((dx2 - hx - r) * D.invL[0]), a1 = (int)ceilf((dx2 + hx + r) * D.invL[0]);
            for (int na = a0; na <= a1; ++na) {
                const float ax = fmaxf(0.f, fabsf(fmaf(-D.L[0], (float)na, dx2)) - hx);
                if (ax * ax < remy) return true;
            }
        }
    }
    return false;
}

// float -> 64-bit fixed point (round to nearest).  |f * scale| < 2^51 takes the
// magic-number path (one FP64 FMA + a 64-bit subtract); larger values use the
// general conversion.
__device__ __forceinline__ long long to_fixed(float f) {
    double d = (double)f * kForceScale;
    if (fabs(d) < 2251799813685248.0) {
        double t = d + 6755399441055744.0;  // 1.5 * 2^52
        return __double_as_longlong(t) - 0x4338000000000000LL;
    }
    return __double2ll_rn(d);
}

__device__ __forceinline__ void atomic_add_fixed(long long* p, long long v) {
    atomicAdd(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(v));
}

// Mark a slot's FP64 side sums as nonzero (the blend reads only marked slots).
__device__ __forceinline__ void mark_side(const DevPtrs& P, int c, long long slot) {
    atomicOr(P.sdMask + (slot >> 6), 1ull << (slot & 63));
    P.rt[c].big = 1;
}

// Add a partial force to a slot: 64-bit fixed point below the set's fixMax
// (<= 2^27 kJ/mol/nm, so the magic-number conversion is exact), FP64 side
// accumulator at or above it (clashing pairs of the displaced state).
__device__ __forceinline__ void add_force(const DevPtrs& P, int c, float fixMax, int slot, float fx, float fy, float fz) {
    long long* f = P.sf + slot;
    double* d = P.sd + 3ll * slot;
    bool big = false;
    if (fabsf(fx) < fixMax) atomic_add_fixed(f, to_fixed(fx)); else { atomicAdd(d, (double)fx); big = true; }
    if (fabsf(fy) < fixMax) atomic_add_fixed(f + P.sfn, to_fixed(fy)); else { atomicAdd(d + 1, (double)fy); big = true; }
    if (fabsf(fz) < fixMax) atomic_add_fixed(f + 2 * P.sfn, to_fixed(fz)); else { atomicAdd(d + 2, (double)fz); big = true; }
    if (big) mark_side(P, c, slot);
}

// DPP-fused sum over each aligned group of 8 lanes (xor 1, xor 2, half-row mirror).
__device__ __forceinline__ float oct_sum(float v) {
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, true));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x141, 0xF, 0xF, true));
    return v;
}

// x, y and z summed over each aligned group of 8 lanes, scattered: lanes 0 and 7
// of the group return the x sum, 1 and 6 the y sum, 2 and 5 the z sum (3 and 4:
// partial sums).  Pairs and quads of all three, then one half-row mirror of the
// value each lane keeps; selX / selY: this lane keeps x / y.
__device__ __forceinline__ float oct_sum3(float x, float y, float z, bool selX, bool selY) {
    x += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0xB1, 0xF, 0xF, true));
    y += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, y), 0xB1, 0xF, 0xF, true));
    z += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, z), 0xB1, 0xF, 0xF, true));
    x += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), 0x4E, 0xF, 0xF, true));
    y += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, y), 0x4E, 0xF, 0xF, true));
    z += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, z), 0x4E, 0xF, 0xF, true));
    const float v = selX ? x : (selY ? y : z);
    return v + __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x141, 0xF, 0xF, true));
}

// Sum over the 8 lanes {l & 7 + 8 k}: the same i-atom across the 8 j-columns.
__device__ __forceinline__ float stride8_sum(float v) {
    v += __shfl_xor(v, 8, 64);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

__device__ __forceinline__ double wave_sum_d(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// In-cutoff decision: FP32 r^2 outside a narrow window around rc^2 decides;
// inside it the exact FP64 r^2 (the squares of 2^-20-grid differences are
// exact) decides, so the pair set equals the FP64 one bit for bit.
__device__ __forceinline__ bool in_cutoff(const SetDesc& D, float dx, float dy, float dz, float r2) {
    bool in = r2 < D.rc2lo;
    if (r2 >= D.rc2lo && r2 < D.rc2hi) {
        double r2d = (double)dx * dx + (double)dy * dy + (double)dz * dz;
        in = r2d < D.rc2;
    }
    return in;
}

// 1/sqrt(x): hardware estimate plus one Newton step written on the residual
// e = 1 - x y^2 (an FMA, exact before its rounding), which keeps the result
// unbiased; the form y (1.5 - 0.5 x y^2) rounds near 1.0 and drops the
// correction of low estimates.
__device__ __forceinline__ float rsqrt_nr(float x) {
    float y = __builtin_amdgcn_rsqf(x);
    const float t = x * y;
    const float e = fmaf(-t, y, 1.0f);
    return fmaf(0.5f * y, e, y);
}

// Masked LJ + reaction-field pair term.  Particle data: pos.w = q sqrt(k_e);
// prm = (sigma/2, 2 sqrt(eps)).  fs = F_i / (x_i - x_j); the energy is returned
// as elj (LJ) and the Coulomb pieces qq/r, qq and qq r^2, so that
// E_C = qq/r + k_rf qq r^2 - c_rf qq is formed in FP64 from the sums.  Lanes
// with in == false return zeros (their r^2 may be 0: self pairs, dummies).
struct PairE {
    float fs, elj, qr, qq, qr2;
};

__device__ __forceinline__ PairE pair_term(bool in, float r2, float qq, float sig, float eps4, float krf) {
    PairE o;
    const float rinv = rsqrt_nr(r2);
    const float r2inv = rinv * rinv;
    const float s2 = sig * sig * r2inv;
    const float s6 = s2 * s2 * s2;
    const float elj = eps4 * fmaf(s6, s6, -s6);
    const float flj = eps4 * s6 * fmaf(12.0f, s6, -6.0f);
    const float fc = qq * fmaf(-2.0f * krf, r2, rinv);
    const float fs = (flj + fc) * r2inv;
    const float qqm = in ? qq : 0.f;
    o.fs = in ? fs : 0.f;
    o.elj = in ? elj : 0.f;
    o.qr = qqm * (in ? rinv : 0.f);
    o.qq = qqm;
    o.qr2 = qqm * r2;
    return o;
}

// FP64 scalar pair function from the exact FP64 r^2 and FP64 per-atom
// parameters (h = sigma/2, e' = 2 sqrt(eps), q' = q sqrt(k_e)).  On CDNA4 the
// FP64 FMA issues at the rate of a non-packed FP32 FMA.  Returns the force
// magnitude over r (FP64), the energy without the reaction-field constants
// (e = E_LJ + q'q'/r) and q'q'.
struct PairE64 {
    double fs, e, qq;
};

__device__ __forceinline__ PairE64 pair_term64(bool in, double r2, double sig, double eps4, double qq, double krf) {
    double y = __builtin_amdgcn_rsq(r2);  // ~2^-22 estimate
    {
        const double t = r2 * y;
        const double e = fma(-t, y, 1.0);
        y = fma(0.5 * y, e, y);            // Newton on the residual: ~1e-13
    }
    const double r2inv = y * y;
    const double s2 = sig * sig * r2inv;
    const double s6 = s2 * s2 * s2;
    // E_LJ = t (s6 - 1) and r dE/dr = -(12 t s6 - 6 t) = -(12 E_LJ + 6 t), t = 4 eps s6
    const double t = eps4 * s6;
    const double elj = fma(t, s6, -t);
    const double flj = fma(12.0, elj, 6.0 * t);
    const double fc = qq * fma(-2.0 * krf, r2, y);
    PairE64 o;
    o.fs = in ? (flj + fc) * r2inv : 0.0;
    o.e = in ? fma(qq, y, elj) : 0.0;
    o.qq = in ? qq : 0.0;
    return o;
}

// The same FP64 pair function returning the whole pair energy,
// E_LJ + q'q' (1/r + k_rf r^2 - c_rf), so one FP64 sum per lane suffices.
struct PairEF64 {
    double fs, e;
};

__device__ __forceinline__ PairEF64 pair_ef64(bool in, double r2, double sig, double eps4, double qq,
                                             double krf, double crf) {
    double y = __builtin_amdgcn_rsq(r2);
    {
        const double t = r2 * y;
        const double e = fma(-t, y, 1.0);
        y = fma(0.5 * y, e, y);
    }
    const double r2inv = y * y;
    const double s2 = sig * sig * r2inv;
    const double s6 = s2 * s2 * s2;
    const double t = eps4 * s6;
    const double elj = fma(t, s6, -t);
    const double flj = fma(12.0, elj, 6.0 * t);
    const double fc = qq * fma(-2.0 * krf, r2, y);
    PairEF64 o;
    o.fs = in ? (flj + fc) * r2inv : 0.0;
    o.e = in ? fma(qq, fma(krf, r2, y) - crf, elj) : 0.0;
    return o;
}

// LDS writes of a wave visible to all its lanes (wavefront-scope fences around a wave barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// ---------------------------------------------------------------- binning

__global__ void k_zero_counts(DevPtrs P, int setBegin, double* f, long long nf) {
    // blockIdx.y: set setBegin + blockIdx.y * setStride (its cell counters and pair counters)
    const int c = setBegin + (int)blockIdx.y * P.setStride;
    const SetDesc& D = P.desc[c];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncells = D.nx * D.ny * D.nz;
    for (int k = i; k < ncells; k += gridDim.x * blockDim.x) P.cellCount[D.cellOff + k] = 0;
    // outF (when the evaluation clears it without a blend): the blocks of the first set
    if (blockIdx.y == 0)
        for (long long k = i; k < nf; k += (long long)gridDim.x * blockDim.x) f[k] = 0.0;
    // the overflow flag stays set until the host has seen it (it relays out and clears it)
    if (i == 0) {
        P.rt[c].npairs = 0ull;
        P.rt[c].maxEntries = 0;
        P.rt[c].maxAtoms = 0;
        P.rt[c].lpairs[0] = 0ull;
        P.rt[c].lpairs[1] = 0ull;
        if (P.setStride == 2 && (c & 1) == 0) {
            // state difference: the state-1 set's counters hold the derived state-1 values
            P.rt[c + 1].npairs = 0ull;
            P.rt[c + 1].maxEntries = 0;
            P.rt[c + 1].maxAtoms = 0;
        }
    }
}

__global__ void k_bin(DevPtrs P, int setBegin) {
    const int c = setBegin + (int)blockIdx.y * P.setStride;
    const SetDesc& D = P.desc[c];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.natoms) return;
    float4 p = P.posq[D.atomOff + i];
    if (D.state) {
        float4 d = P.displ[D.atomOff + i];
        p.x += d.x; p.y += d.y; p.z += d.z;  // exact on the 2^-20 grid
    }
    if (D.tri) {
        wrap_tri(D, p);
    } else {
        p.x = wrap_coord(p.x, D.L[0], D.invL[0]);
        p.y = wrap_coord(p.y, D.L[1], D.invL[1]);
        p.z = wrap_coord(p.z, D.L[2], D.invL[2]);
    }
    // clamp in float before the conversion: every index stays inside the grid
    const int cx = (int)fminf(fmaxf(p.x * D.invCe[0], 0.f), (float)(D.nx - 1));
    const int cy = (int)fminf(fmaxf(p.y * D.invCe[1], 0.f), (float)(D.ny - 1));
    const int cz = (int)fminf(fmaxf(p.z * D.invZe, 0.f), (float)(D.nz - 1));
    int cell = (cx * D.ny + cy) * D.nz + cz;
    int r = atomicAdd(P.cellCount + D.cellOff + cell, 1);
    const int k = D.setAtomOff + i;
    P.bpos[k] = p;
    P.bcell[k] = cell;
    P.brank[k] = r;
    if (P.setStride == 2 && (c & 1) == 0) {
        // state difference: clear dF of atom i (the state-1 set's sums, by atom) and,
        // after an evaluation that used them, its FP64 side sums and their mask word
        // (the 64 atoms of a word are one wave's: set slot ranges start on multiples of 64)
        const long long k2 = P.desc[c + 1].slotOff + i;
        if (P.rt[c + 1].big) {
            const unsigned long long m = P.sdMask[k2 >> 6];
            if ((m >> (k2 & 63)) & 1ull) {
                double* d = P.sd + 3 * k2;
                d[0] = 0.0; d[1] = 0.0; d[2] = 0.0;
            }
            if ((k2 & 63) == 0 && m != 0ull) P.sdMask[k2 >> 6] = 0ull;
        }
        P.sf[k2] = 0; P.sf[P.sfn + k2] = 0; P.sf[2 * P.sfn + k2] = 0;
    }
}

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// One block of 1024 threads (16 waves) per set.
//  A: one wave per column (16 columns in flight), wave prefix sums over its
//     z-bins -> cell starts relative to their column, and the column count;
//  B: after a block barrier, each thread scans a contiguous chunk of columns
//     (padded slot counts and group counts from A's column counts), a block
//     scan joins the chunks; column starts, dummy slots and the group table
//     are written.
// A cell's first slot is colStart[column] + cellStart[cell].
__global__ void __launch_bounds__(1024) k_scan(DevPtrs P, int setBegin) {
    constexpr int NT = 1024, NW = NT / 64;
    const int c = setBegin + (int)blockIdx.x * P.setStride;
    const SetDesc& D = P.desc[c];
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int ncols = D.nx * D.ny, nz = D.nz;
    if (nz <= 64) {
        // a column in one wave step: U columns' counts loaded before their scans,
        // so U loads per wave are in flight instead of one
        constexpr int U = 8;
        for (int col0 = wave; col0 < ncols; col0 += U * NW) {
            int v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int col = col0 + u * NW;
                v[u] = (col < ncols && lane < nz) ? P.cellCount[D.cellOff + (size_t)col * nz + lane] : 0;
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int col = col0 + u * NW;
                const int incl = wave_incl_scan(v[u]);  // lane 63: the column count (lanes >= nz add 0)
                if (col < ncols) {
                    if (lane < nz) P.cellStart[D.cellOff + (size_t)col * nz + lane] = incl - v[u];
                    if (lane == 63) P.colCount[D.colOff + col] = incl;
                }
            }
        }
    }
    for (int col = nz <= 64 ? ncols : wave; col < ncols; col += NW) {
        const int* cnt = P.cellCount + D.cellOff + (size_t)col * nz;
        int* st = P.cellStart + D.cellOff + (size_t)col * nz;
        int run = 0;
        for (int z0 = 0; z0 < nz; z0 += 64) {
            const int z = z0 + lane;
            const int v = z < nz ? cnt[z] : 0;
            const int incl = wave_incl_scan(v);
            if (z < nz) st[z] = run + incl - v;
            run += __shfl(incl, 63, 64);
        }
        if (lane == 0) P.colCount[D.colOff + col] = run;
    }
    // the column counts written above by other waves of this block: the barrier
    // orders them (workgroup scope covers global memory of the same block)
    __syncthreads();
    const int* colCount = P.colCount + D.colOff;
    const int chunk = (ncols + NT - 1) / NT;
    const int b = min(t * chunk, ncols), e = min(b + chunk, ncols);
    __shared__ int sS[NW], sG[NW];
    int accS = 0, accG = 0;
    for (int col = b; col < e; ++col) {
        const int n = colCount[col];
        const int padded = (n + 7) & ~7;
        accS += padded;
        accG += ((padded >> 3) + kGroupClusters - 1) / kGroupClusters;
    }
    // block exclusive scan: wave scans, then a scan of the 16 wave totals
    const int inS = wave_incl_scan(accS), inG = wave_incl_scan(accG);
    if (lane == 63) { sS[wave] = inS; sG[wave] = inG; }
    __syncthreads();
    int baseS = 0, baseG = 0, totS = 0, totG = 0;
    for (int w = 0; w < NW; ++w) {
        const int vs = sS[w], vg = sG[w];
        if (w < wave) { baseS += vs; baseG += vg; }
        totS += vs; totG += vg;
    }
    int offS = baseS + inS - accS, offG = baseG + inG - accG;
    for (int col = b; col < e; ++col) {
        const int cnt = colCount[col];
        const int padded = (cnt + 7) & ~7;
        const int ncl = padded >> 3;
        const int ng = (ncl + kGroupClusters - 1) / kGroupClusters;
        P.colStart[D.colOff + col] = offS;
        for (int s2 = offS + cnt; s2 < offS + padded; ++s2) P.sslot[D.slotOff + s2] = make_int2(-1, -1);
        for (int gg = 0; gg < ng; ++gg)
            P.ginfo[D.groupOff + offG + gg] =
                make_int4((offS >> 3) + gg * kGroupClusters, min(kGroupClusters, ncl - gg * kGroupClusters), col, 0);
        offS += padded;
        offG += ng;
    }
    if (t == 0) {
        P.rt[c].nslots = totS;
        P.rt[c].nclusters = totS >> 3;
        P.rt[c].ngroups = totG;
        // state difference: dF's side sums were cleared by k_bin (the displaced-atom kernels,
        // which may run beside the pair list, set the flag again)
        if (P.setStride == 2 && (c & 1) == 0) P.rt[c + 1].big = 0;
    }
}

// Counting-sort scatter: each atom's (index, cell) into its provisional slot
// (cell start + the atomic rank of k_bin, whose order is not deterministic).
__global__ void k_scatter_idx(DevPtrs P, int setBegin) {
    const int c = setBegin + (int)blockIdx.y * P.setStride;
    const SetDesc& D = P.desc[c];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.natoms) return;
    const int k = D.setAtomOff + i;
    const int cell = P.bcell[k];
    const int slot = P.colStart[D.colOff + cell / D.nz] + P.cellStart[D.cellOff + cell] + P.brank[k];
    P.sslot[D.slotOff + slot] = make_int2(i, cell);
}

// The atom that belongs at rank r of its fine cell: the cell's atoms ordered
// by index (fixed cluster composition makes the FP32 partial sums, and so
// every result of an evaluation, bit-reproducible).  mem: the cell's
// provisional (atom, cell) slots.  Cells of up to 8 atoms (nearly all at the
// default binning) are sorted in registers by an optimal 19-comparator
// network, up to 16 by odd-even transposition, larger ones by rank counting.
__device__ __forceinline__ int cell_rank_atom(const int2* mem, int cnt, int r) {
    if (cnt <= 8) {
        int v[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) v[a] = a < cnt ? mem[a].x : 0x7FFFFFFF;
        auto cx = [&](int a, int b) {
            const int lo = min(v[a], v[b]), hi = max(v[a], v[b]);
            v[a] = lo;
            v[b] = hi;
        };
        cx(0, 2); cx(1, 3); cx(4, 6); cx(5, 7);
        cx(0, 4); cx(1, 5); cx(2, 6); cx(3, 7);
        cx(0, 1); cx(2, 3); cx(4, 5); cx(6, 7);
        cx(2, 4); cx(3, 5);
        cx(1, 4); cx(3, 6);
        cx(1, 2); cx(3, 4); cx(5, 6);
        int out = v[0];
#pragma unroll
        for (int a = 1; a < 8; ++a) out = r == a ? v[a] : out;
        return out;
    }
    if (cnt <= 16) {
        int v[16];
#pragma unroll
        for (int a = 0; a < 16; ++a) v[a] = a < cnt ? mem[a].x : 0x7FFFFFFF;
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
#pragma unroll
            for (int a = rr & 1; a + 1 < 16; a += 2) {
                const int lo = min(v[a], v[a + 1]), hi = max(v[a], v[a + 1]);
                v[a] = lo;
                v[a + 1] = hi;
            }
        }
        int out = v[0];
#pragma unroll
        for (int a = 1; a < 16; ++a) out = r == a ? v[a] : out;
        return out;
    }
    int out = -1;
    for (int a = 0; a < cnt; ++a) {
        const int va = mem[a].x;
        int less = 0;
        for (int b = 0; b < cnt; ++b) less += mem[b].x < va ? 1 : 0;
        if (less == r) out = va;
    }
    return out;
}

__device__ __forceinline__ float oct_min(float v) {
    v = fminf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true)));
    v = fminf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, true)));
    v = fminf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x141, 0xF, 0xF, true)));
    return v;
}

__device__ __forceinline__ float oct_max(float v) {
    v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true)));
    v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, true)));
    v = fmaxf(v, __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x141, 0xF, 0xF, true)));
    return v;
}

// One thread per slot: gather the atom's wrapped position and parameters into
// cluster order (coalesced writes), clear the slot's force accumulators, and
// reduce each cluster's bounding box over its 8 lanes.  Slot counts are
// multiples of 8, so a cluster's lanes are all active or all gone.
__global__ void __launch_bounds__(256) k_gather_bbox(DevPtrs P, int setBegin) {
    const int c = setBegin + (int)blockIdx.y * P.setStride;
    const SetDesc& D = P.desc[c];
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    // FP64 side sums: after an evaluation that used them (k_pairlist then resets
    // the flag), the marked slots are cleared and so is their mask word (the 64
    // slots of a word are one wave's: set slot ranges start on multiples of 64)
    if (P.rt[c].big && s < D.slotCap) {
        const long long k = D.slotOff + s;
        const unsigned long long m = P.sdMask[k >> 6];
        if ((m >> (k & 63)) & 1ull) {
            double* d = P.sd + 3 * k;
            d[0] = 0.0; d[1] = 0.0; d[2] = 0.0;
        }
        if ((k & 63) == 0 && m != 0ull) P.sdMask[k >> 6] = 0ull;
    }
    if (s >= P.rt[c].nslots) return;
    const int k = D.slotOff + s;
    // this slot's atom: rank s - (cell start) of its cell in index order
    const int2 me = P.sslot[k];
    int i = -1;
    if (me.y >= 0) {
        const int base = P.colStart[D.colOff + me.y / D.nz] + P.cellStart[D.cellOff + me.y];
        i = cell_rank_atom(P.sslot + D.slotOff + base, P.cellCount[D.cellOff + me.y], s - base);
    }
    P.sidx[k] = i;
    // dummy slots: zero parameters, and positions far outside the box and at
    // least 4 nm from each other (x = 1e6 + 4 (s mod 2^20), y = 1e6 + 4 (s >> 20),
    // exact in FP32), so no pair with a dummy passes a cutoff test unless x/y
    // are minimum-imaged (sets without explicit image shifts test validity)
    float4 p = make_float4(1.0e6f + 4.0f * (float)(s & 0xFFFFF), 1.0e6f + 4.0f * (float)(s >> 20), 0.f, 0.f);
    float2 q = make_float2(0.f, 0.f);
    double4 qd = make_double4(0.0, 0.0, 0.0, 0.0);
    if (i >= 0) {
        p = P.bpos[D.setAtomOff + i];
        q = P.ljp[D.atomOff + i];
        qd = P.prm64[D.atomOff + i];
        P.slotOf[D.setAtomOff + i] = s;
    }
    P.spos[k] = p;
    P.sprm[k] = q;
    P.sprm64[k] = qd;
    P.sf[k] = 0; P.sf[P.sfn + k] = 0; P.sf[2 * P.sfn + k] = 0;
    const float inf = __builtin_huge_valf();
    const float lox = oct_min(i >= 0 ? p.x : inf), hix = oct_max(i >= 0 ? p.x : -inf);
    const float loy = oct_min(i >= 0 ? p.y : inf), hiy = oct_max(i >= 0 ? p.y : -inf);
    const float loz = oct_min(i >= 0 ? p.z : inf), hiz = oct_max(i >= 0 ? p.z : -inf);
    if ((s & 7) == 0) {
        // every cluster holds at least one real atom (padding < 8 per column)
        const int cl = D.clusterOff + (s >> 3);
        P.bbc[cl] = make_float4(0.5f * (lox + hix), 0.5f * (loy + hiy), 0.5f * (loz + hiz), 0.f);
        P.bbh[cl] = make_float4(0.5f * (hix - lox), 0.5f * (hiy - loy), 0.5f * (hiz - loz), 0.f);
    }
}

// ---------------------------------------------------------------- pair list


// One wave64 per group.  Entry = J (20 bits) | imask (8 bits) << 20 |
// x-shift code << 28 | y-shift code << 30 (code 1: +L, 2: -L).
// Candidate columns (up to 64 per pass, one per lane) are enumerated first; a
// wave prefix sum then flattens their (column, cluster) candidates so that
// all 64 lanes test clusters, whatever the column occupancy.
template <int W, bool TRI>
__global__ void __launch_bounds__(64 * W) k_pairlist(DevPtrs P, int setBegin) {
    const int c = setBegin + (int)blockIdx.y * P.setStride;
    const SetDesc D = P.desc[c];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int g = blockIdx.x * W + wave;
    if (g == 0 && lane == 0) P.rt[c].big = 0;  // the side sums were cleared by k_gather_bbox
    if (g >= P.rt[c].ngroups) return;
    const int4 gi = P.ginfo[D.groupOff + g];
    const int g0 = __builtin_amdgcn_readfirstlane(gi.x), ncl = __builtin_amdgcn_readfirstlane(gi.y);
    const int col = __builtin_amdgcn_readfirstlane(gi.z);
    const int gx = col / D.ny, gy = col % D.ny;
    // the group's cluster boxes in LDS (read per test: registers for all 8 would
    // hold the wave to 5 per SIMD, and the kernel waits on memory)
    __shared__ float4 sIc[W][kGroupClusters], sIh[W][kGroupClusters];
    float4 myc = make_float4(0.f, 0.f, 0.f, 0.f), myh = myc;
    if (lane < ncl) { myc = P.bbc[D.clusterOff + g0 + lane]; myh = P.bbh[D.clusterOff + g0 + lane]; }
    if (lane < kGroupClusters) { sIc[wave][lane] = myc; sIh[wave][lane] = myh; }
    wave_sync();
    uint32_t* out = P.entries + D.entryOff + (long long)g * D.entryCap;
    int nEnt = 0;
    const int ncx = D.fullX ? D.nx : 2 * D.Rx + 1, ncy = D.fullY ? D.ny : 2 * D.Ry + 1;
    const int ncand = ncx * ncy;
    // z-window of the group: a j-atom within the list cutoff of any i-atom has
    // z in [zlo - rcl, zhi + rcl] (periodic), so only clusters holding an atom
    // in the z-bins of that window (one bin of slack on each side for rounding)
    // can pair; they form one or two (wrapped) slot ranges per column.
    float zlo = sIc[wave][0].z - sIh[wave][0].z, zhi = sIc[wave][0].z + sIh[wave][0].z;
    for (int q = 1; q < ncl; ++q) {
        const float cz = sIc[wave][q].z, hz = sIh[wave][q].z;
        zlo = fminf(zlo, cz - hz);
        zhi = fmaxf(zhi, cz + hz);
    }
    const float rcl = sqrtf(D.rcList2);
    const int blo = (int)floorf((zlo - rcl) * D.invZe) - 1, bhi = (int)floorf((zhi + rcl) * D.invZe) + 1;
    const bool zall = bhi - blo + 1 >= D.nz;  // the window covers the column
    const int bl = zall ? 0 : ((blo % D.nz) + D.nz) % D.nz;
    const int bh = zall ? D.nz - 1 : ((bhi % D.nz) + D.nz) % D.nz;
    const bool zwrap = !zall && bl > bh;       // two ranges: [bl, nz) and [0, bh]
    // triclinic: every column is a candidate; a column whose x-y rectangle is beyond the
    // list cutoff of the group's x-y box under every image is dropped before its
    // clusters are tested.  A z-window inside [0, L2) admits only c-images n_c = 0.
    float gcx = 0.f, gcy = 0.f, ghx = 0.f, ghy = 0.f;
    const int ncSpan = (zwrap || zall) ? 1 : 0;
    if (TRI && D.tri) {
        float xlo = sIc[wave][0].x - sIh[wave][0].x, xhi = sIc[wave][0].x + sIh[wave][0].x;
        float ylo = sIc[wave][0].y - sIh[wave][0].y, yhi = sIc[wave][0].y + sIh[wave][0].y;
        for (int q = 1; q < ncl; ++q) {
            xlo = fminf(xlo, sIc[wave][q].x - sIh[wave][q].x);
            xhi = fmaxf(xhi, sIc[wave][q].x + sIh[wave][q].x);
            ylo = fminf(ylo, sIc[wave][q].y - sIh[wave][q].y);
            yhi = fmaxf(yhi, sIc[wave][q].y + sIh[wave][q].y);
        }
        gcx = 0.5f * (xlo + xhi); gcy = 0.5f * (ylo + yhi);
        ghx = 0.5f * (xhi - xlo); ghy = 0.5f * (yhi - ylo);
    }
    for (int cb = 0; cb < ncand; cb += 64) {
        // this lane's candidate column: up to two cluster ranges (a, b), shift codes
        int cnt_l = 0, c0_l = 0, codes_l = 0, cntb_l = 0, c0b_l = 0;
        const int ci = cb + lane;
        if (ci < ncand) {
            const int a = ci / ncy, b = ci % ncy;
            int jx = a, jy = b, sxc = 0, syc = 0;
            float gapx = 0.f, gapy = 0.f;
            if (!D.fullX) {
                const int ox = a - D.Rx;
                jx = gx + ox;
                if (jx < 0) { jx += D.nx; sxc = 2; } else if (jx >= D.nx) { jx -= D.nx; sxc = 1; }
                gapx = fmaxf(0.f, (float)(abs(ox) - 1)) * D.ce[0];
            }
            if (!D.fullY) {
                const int oy = b - D.Ry;
                jy = gy + oy;
                if (jy < 0) { jy += D.ny; syc = 2; } else if (jy >= D.ny) { jy -= D.ny; syc = 1; }
                gapy = fmaxf(0.f, (float)(abs(oy) - 1)) * D.ce[1];
            }
            bool colNear = gapx * gapx + gapy * gapy < D.rcList2;
            // triclinic, window wrapped at one end: the bins of the window inside [0, L_z) pair at
            // n_c = 0 only, the bins past the end (taken from the other end of the column) at
            // n_c = +-1 only; each range is kept when the column is near under its own images
            bool keepDirect = true, keepWrapped = true;
            if (TRI && D.tri) {
                // the column's rectangle, 1e-3 nm wider on each side than its edges (binning rounding)
                const float hx = ghx + 0.5f * D.ce[0] + 1e-3f, hy = ghy + 0.5f * D.ce[1] + 1e-3f;
                const float cdx = ((float)jx + 0.5f) * D.ce[0] - gcx, cdy = ((float)jy + 0.5f) * D.ce[1] - gcy;
                if (zwrap) {
                    keepDirect = tri_rect_near(D, cdx, cdy, hx, hy, 0, rcl, D.rcList2);
                    keepWrapped = tri_rect_near(D, cdx, cdy, hx, hy, 1, rcl, D.rcList2, true);
                    colNear = keepDirect || keepWrapped;
                } else {
                    colNear = tri_rect_near(D, cdx, cdy, hx, hy, ncSpan, rcl, D.rcList2);
                }
            }
            if (colNear) {
                const int colJ = jx * D.ny + jy;
                const int cs = P.colStart[D.colOff + colJ], cnt = P.colCount[D.colOff + colJ];
                const int* cst = P.cellStart + D.cellOff + (size_t)colJ * D.nz;
                const int* ccn = P.cellCount + D.cellOff + (size_t)colJ * D.nz;
                // slots [s0, s1) of bins [b0, b1] -> clusters [s0 >> 3, ceil(s1 / 8)), at least g0
                auto range = [&](int b0, int b1, int& c0, int& n) {
                    const int s0 = cs + cst[b0], s1 = cs + cst[b1] + ccn[b1];
                    const int lo = max(s0 >> 3, g0), hi = (s1 + 7) >> 3;
                    c0 = lo;
                    n = max(0, hi - lo);
                };
                if (cnt > 0) {
                    if (zall) {
                        const int cl0 = max(cs >> 3, g0), cl1 = (cs + ((cnt + 7) & ~7)) >> 3;
                        c0_l = cl0;
                        cnt_l = max(0, cl1 - cl0);
                    } else if (!zwrap) {
                        range(bl, bh, c0_l, cnt_l);
                    } else {
                        range(bl, D.nz - 1, c0_l, cnt_l);
                        range(0, bh, c0b_l, cntb_l);
                        // the window runs past the top (bhi >= nz: [bl, nz) is its direct part) or
                        // below 0 ([0, bh] is); a range whose images are not near is dropped
                        const bool topDirect = bhi >= D.nz;
                        const bool keepTop = topDirect ? keepDirect : keepWrapped;
                        const bool keepBottom = topDirect ? keepWrapped : keepDirect;
                        if (!keepTop) cnt_l = 0;
                        if (!keepBottom) cntb_l = 0;
                        // a cluster straddling the empty bins between the two windows
                        // would fall in both ranges: the top range starts after the bottom one
                        const int endb = c0b_l + cntb_l;
                        if (cntb_l > 0 && cnt_l > 0 && c0_l < endb) {
                            cnt_l = max(0, c0_l + cnt_l - endb);
                            c0_l = endb;
                        }
                    }
                    codes_l = sxc | (syc << 2);
                }
            }
        }
        cnt_l += cntb_l;  // flattened: range a, then range b
        const int incl = wave_incl_scan(cnt_l);
        const int excl = incl - cnt_l;
        const int total = __shfl(incl, 63, 64);
        for (int f0 = 0; f0 < total; f0 += 64) {
            const int f = f0 + lane;
            // the candidate column holding flat index f: the last lane with excl <= f
            int lo = 0;
            for (int step = 32; step > 0; step >>= 1) {
                const int probe = min(lo + step, 63);
                const int ex = __shfl(excl, probe, 64);
                if (lo + step <= 63 && ex <= f) lo = probe;
            }
            // every shuffle with all lanes active: a lane that a branch masks off
            // does not provide its value to the others
            const int off = f - __shfl(excl, lo, 64);
            const int na = __shfl(cnt_l - cntb_l, lo, 64);
            const int c0a = __shfl(c0_l, lo, 64), c0b = __shfl(c0b_l, lo, 64);
            const int J = off < na ? c0a + off : c0b + (off - na);
            const int codes = __shfl(codes_l, lo, 64);
            uint32_t mask = 0;
            if (f < total) {
                const int sxc = codes & 3, syc = codes >> 2;
                const float sx = D.shiftOK ? (sxc == 1 ? D.L[0] : (sxc == 2 ? -D.L[0] : 0.f)) : 0.f;
                const float sy = D.shiftOK ? (syc == 1 ? D.L[1] : (syc == 2 ? -D.L[1] : 0.f)) : 0.f;
                const float4 jc = P.bbc[D.clusterOff + J], jh = P.bbh[D.clusterOff + J];
                if (TRI && D.tri) {
                    // triclinic: every lattice image the boxes can reach (one loop copy, not unrolled)
                    for (int q = 0; q < ncl && g0 + q <= J; ++q) {
                        const float4 ic = sIc[wave][q], ih = sIh[wave][q];
                        if (tri_boxes_near(D, jc.x - ic.x, jc.y - ic.y, jc.z - ic.z, ih.x + jh.x, ih.y + jh.y,
                                           ih.z + jh.z, rcl, D.rcList2))
                            mask |= 1u << q;
                    }
                } else {
#pragma unroll
                    for (int q = 0; q < kGroupClusters; ++q) {
                        if (q < ncl && g0 + q <= J) {
                            const float4 ic = sIc[wave][q], ih = sIh[wave][q];
                            float dx = jc.x + sx - ic.x, dy = jc.y + sy - ic.y, dz = jc.z - ic.z;
                            if (!D.shiftOK) { dx = min_image(dx, D.L[0], D.invL[0]); dy = min_image(dy, D.L[1], D.invL[1]); }
                            dz = min_image(dz, D.L[2], D.invL[2]);
                            const float ax = fmaxf(0.f, fabsf(dx) - ih.x - jh.x);
                            const float ay = fmaxf(0.f, fabsf(dy) - ih.y - jh.y);
                            const float az = fmaxf(0.f, fabsf(dz) - ih.z - jh.z);
                            if (ax * ax + ay * ay + az * az < D.rcList2) mask |= 1u << q;
                        }
                    }
                }
            }
            const uint64_t bal = __ballot(mask != 0);
            const int pos = nEnt + __popcll(bal & lanemask_lt());
            if (mask != 0 && pos < D.entryCap) {
                const uint32_t sh = D.shiftOK ? ((uint32_t)(codes & 3) << 28) | ((uint32_t)(codes >> 2) << 30) : 0u;
                out[pos] = (uint32_t)J | (mask << 20) | sh;
            }
            nEnt += __popcll(bal);
        }
    }
    if (lane == 0) {
        const int kept = min(nEnt, D.entryCap);
        P.gcount[D.groupOff + g] = kept;
        atomicMax(&P.rt[c].maxEntries, nEnt);
        if (nEnt > D.entryCap) atomicOr(&P.rt[c].overflow, 1);
    }
}

// ---------------------------------------------------------------- pair kernel
// The tile kernel's group order, longest lists first, per set: one block per set
// setBegin + blockIdx.x setStride files its groups by list length (bucket
// min(63, entries >> kLptShift)) with LDS counting and writes them from the
// longest bucket down to lptItems[groupOff ...].  The tile kernel takes rank r as
// the (r / nsets)-th longest group of set r mod nsets, so the launch starts with
// every set's longest lists.  The order within a bucket follows LDS atomics,
// which no result depends on (each group's sums are its own).
__global__ void __launch_bounds__(512) k_order(DevPtrs P, int setBegin) {
    constexpr int NT = 512, NI = 4;  // groups held per thread (more take a second sweep)
    __shared__ int hist[kLptBuckets];
    const int t = threadIdx.x;
    const int c = setBegin + (int)blockIdx.x * P.setStride;
    const int gOff = P.desc[c].groupOff, ng = P.rt[c].ngroups;
    if (t < kLptBuckets) hist[t] = 0;
    __syncthreads();
    for (int g0 = 0; g0 < ng; g0 += NT * NI) {
        int bk[NI];
#pragma unroll
        for (int u = 0; u < NI; ++u) {
            const int g = g0 + u * NT + t;
            bk[u] = g < ng ? min(kLptBuckets - 1, P.gcount[gOff + g] >> kLptShift) : -1;
        }
#pragma unroll
        for (int u = 0; u < NI; ++u)
            if (bk[u] >= 0) atomicAdd(hist + bk[u], 1);
    }
    __syncthreads();
    if (t == 0) {
        int acc = 0;
        for (int b = kLptBuckets - 1; b >= 0; --b) {
            const int n = hist[b];
            hist[b] = acc;
            acc += n;
        }
    }
    __syncthreads();
    for (int g0 = 0; g0 < ng; g0 += NT * NI) {
        int bk[NI];
#pragma unroll
        for (int u = 0; u < NI; ++u) {
            const int g = g0 + u * NT + t;
            bk[u] = g < ng ? min(kLptBuckets - 1, P.gcount[gOff + g] >> kLptShift) : -1;
        }
#pragma unroll
        for (int u = 0; u < NI; ++u)
            if (bk[u] >= 0) P.lptItems[gOff + atomicAdd(hist + bk[u], 1)] = g0 + u * NT + t;
    }
}


// Decode an entry and load this lane's j-atom (lane = 8 jj + ii).
struct JAtom {
    float4 p;
    float2 prm;
    bool valid;
    int slot;
};

__device__ __forceinline__ JAtom load_j(const DevPtrs& P, const SetDesc& D, uint32_t ent, int jj) {
    JAtom a;
    const int J = ent & 0xFFFFF;
    a.slot = D.slotOff + J * kClusterSize + jj;
    a.p = P.spos[a.slot];
    a.prm = P.sprm[a.slot];
    a.valid = P.sidx[a.slot] >= 0;
    if (D.shiftOK) {
        const uint32_t sxc = (ent >> 28) & 3u, syc = (ent >> 30) & 3u;
        a.p.x += sxc == 1u ? D.L[0] : (sxc == 2u ? -D.L[0] : 0.f);
        a.p.y += syc == 1u ? D.L[1] : (syc == 2u ? -D.L[1] : 0.f);
    }
    return a;
}

// Flush one partial force vector to a slot.  The common case (all three
// components below fixMax) takes the branch-free fixed-point path.
__device__ __forceinline__ void flush_force(const DevPtrs& P, int c, float fixMax, int slot, float fx, float fy, float fz) {
    if (fmaxf(fabsf(fx), fmaxf(fabsf(fy), fabsf(fz))) < fixMax) {  // <= 2^27: |f scale| < 2^51
        long long* f = P.sf + slot;
        atomic_add_fixed(f, __double_as_longlong((double)fx * kForceScale + 6755399441055744.0) - 0x4338000000000000LL);
        atomic_add_fixed(f + P.sfn, __double_as_longlong((double)fy * kForceScale + 6755399441055744.0) - 0x4338000000000000LL);
        atomic_add_fixed(f + 2 * P.sfn, __double_as_longlong((double)fz * kForceScale + 6755399441055744.0) - 0x4338000000000000LL);
    } else {
        add_force(P, c, fixMax, slot, fx, fy, fz);
    }
}

// b is one of the far excepted partners of atom a (system-local indices).
__device__ __forceinline__ bool far_excluded(const DevPtrs& P, const SetDesc& D, int a, int b) {
    const int* st = P.xfarStart + D.atomOff + a;
    const int k1 = st[1];
    for (int k = st[0]; k < k1; ++k)
        if (P.xfar[k] == b) return true;
    return false;
}

// The pair (a, b), a != b, of D's system is an exception; ma = a's near mask.
__device__ __forceinline__ bool excluded_pair(const DevPtrs& P, const SetDesc& D, int a, int b, unsigned long long ma) {
    const uint32_t dd = (uint32_t)(b - a + 32);
    if (dd - 1u < 63u) return (ma >> dd) & 1ull;
    return ((ma >> 32) & 1ull) && far_excluded(P, D, a, b);
}


// Fixed-order sum of a set's energy slots: one per group, or estride per group
// (8: one per i-cluster wave of the filtered kernel; 2 or 4: one per part of a
// group whose list the tile kernel splits between waves).
__device__ __forceinline__ double group_energy(const DevPtrs& P, const SetDesc& D, int g, int estride) {
    if (estride == 1) return P.genergy[D.groupOff + g];
    const double* e = P.genergy + (long long)estride * (D.groupOff + g);
    if (estride == 8) return ((e[0] + e[1]) + (e[2] + e[3])) + ((e[4] + e[5]) + (e[6] + e[7]));
    double s = e[0];
    for (int k = 1; k < estride; ++k) s += e[k];
    return s;
}


// Add one force component (component k of slot `slot`) in 64-bit fixed point;
// zero sums (j-atoms without an in-cutoff partner, dummy slots) issue no atomic.
// Values at or beyond the set's fixMax (clashing pairs) take a second, rare
// path: the FP64 side accumulator.
__device__ __forceinline__ void flush_component_u(const DevPtrs& P, int c, float fixMax, int slot, int k, float f) {
    const bool small = fabsf(f) < fixMax;
    long long* dst = P.sf + (long long)k * P.sfn + slot;
    const long long v = small ? __double_as_longlong((double)f * kForceScale + 6755399441055744.0) - 0x4338000000000000LL : 0ll;
    if (v != 0) atomic_add_fixed(dst, v);
    if (!small) {
        int sl = slot;
        asm volatile("" : "+v"(sl));  // keeps the rare path's address math from being hoisted
        atomicAdd(P.sd + 3ll * sl + k, (double)f);
        mark_side(P, c, sl);
    }
}

// One wave64 per group of up to 8 i-clusters; lane = 8 jj + ii pairs i-atom ii
// of each interacting i-cluster with j-atom jj of the list entry's j-cluster.
// F64: scalar pair function in FP64 (exact r^2, default); otherwise FP32.
// Coordinates, force vectors and their per-lane accumulators are FP32 either
// way; forces are flushed to 64-bit fixed point, energies summed in FP64.
// List entries are fetched 64 at a time (one per lane) and broadcast with
// readlane, so the cluster mask and the diagonal test are scalar branches.
// At most 80 VGPRs (6 waves per SIMD): the FP64 forces variant with its two
// j-atom buffers then spills a few per-batch values outside the entry loop;
// on a 22 x 23k batch that measured 5 % faster than its unconstrained 88
// VGPRs (5 waves) and 14 % faster than 72 VGPRs (7 waves, which spill inside
// the entry loop).
//
// BWD (FP64 scalars, no forces): the backward pass on the same tiles -- per
// in-cutoff pair, the derivatives of its energy with respect to the stored
// per-atom parameters q' = q sqrt(k_e), h = sigma/2, e' = 2 sqrt(eps) of both
// atoms, from the exact r^2:
//   dE/dq'_i = q'_j phi,  phi = 1/r + k_rf r^2 - c_rf
//   dE/dh    = e'_i e'_j (12 (s/r)^12 - 6 (s/r)^6) / s = e'_i e'_j (12 s6 - 6) s2^2 u
//              with s = h_i + h_j, u = s / r^2, s2 = s u, s6 = s2^3 (0 at s = 0)
//   dE/de'_i = e'_j (s6^2 - s6)
// summed per lane in FP32 where the forward sums forces, weighted by wset[c]
// (dW/dU of the set) and added to FP64 per-slot accumulators (gq, gsig, geps).
// With wset null (an evaluation that also differentiates) the weights are 1,
// and the pair energies and counts are recorded as by the forward modes.
//
// EXC: the batch has exceptions or a triclinic box.  With exceptions (P.xmask)
// every in-cutoff pair is tested against the i-atom's near mask (and, for far
// partners, its list) and excepted pairs are dropped before the pair function;
// sets with a triclinic box take OpenMM's triclinic image (the instantiations
// without EXC carry neither).
//
// split: each group's list is divided into `split` consecutive parts, one per
// wave (blockIdx.x = group block * split + part): shorter waves, so the launch
// ends with less of a tail when the group lists are long.  Each part adds its
// own i-force sums and writes its own energy slot (split slots per group).
template <int W, bool FORCES, bool F64, bool BWD, bool EXC>
__global__ void __launch_bounds__(64 * W, 6) k_pairs(DevPtrs P, int setBegin, const double* wset, int split) {
    static_assert(!BWD || (F64 && !FORCES), "the backward mode is FP64 and computes no forces");
    constexpr bool JSUMS = FORCES || BWD;  // per-atom j sums flushed every kJFlush entries
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jj = lane >> 3, ii = lane & 7;
    int c, g, part;
    bool active;
    if (P.lptItems && P.lptMode == kOrderSetMajor) {
        // each set's groups longest list first, the sets one after the other (blockIdx.y)
        const int f = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * W + wave));
        const int pos = f / split;
        part = f - pos * split;
        c = setBegin + (int)blockIdx.y * P.setStride;
        active = pos < P.rt[c].ngroups;
        g = active ? __builtin_amdgcn_readfirstlane(P.lptItems[P.desc[c].groupOff + pos]) : 0;
    } else if (P.lptItems && P.lptMode != kOrderLaunch) {
        // longest lists first over windows of sets: this wave's rank f in launch order falls in
        // window f / (K gx) of K sets (K = every set, or kOrderWindowSets; gx waves per set), where
        // it is part (rank in the window) % split of item (rank in the window) / split: the
        // (item / K')-th longest group of the window's set item mod K' (K' its sets; k_order)
        const int gx = (int)gridDim.x * W;
        const int K = P.lptMode == kOrderAllSets ? (int)gridDim.y : kOrderWindowSets;
        const int f = __builtin_amdgcn_readfirstlane((int)((blockIdx.y * gridDim.x + blockIdx.x) * W + wave));
        const int win = f / (K * gx), within = f - win * (K * gx);
        const int kw = min(K, (int)gridDim.y - win * K);
        const int item = within / split;
        part = within - item * split;
        const int pos = item / kw;
        c = setBegin + (win * K + (item - pos * kw)) * P.setStride;
        active = within < kw * gx && pos < P.rt[c].ngroups;
        g = active ? __builtin_amdgcn_readfirstlane(P.lptItems[P.desc[c].groupOff + pos]) : 0;
    } else {
        c = setBegin + (int)blockIdx.y * P.setStride;
        part = blockIdx.x % split;
        g = (blockIdx.x / split) * W + wave;
        active = g < P.rt[c].ngroups;
    }
    const SetDesc D = P.desc[c];  // wave-uniform: lives in SGPRs for the whole kernel
    __shared__ float4 sPos[W][64];
    __shared__ float2 sPrm[W][F64 ? 1 : 64];
    __shared__ double4 sPrm64[W][F64 ? 64 : 1];
    __shared__ int sAi[W][EXC ? 64 : 1];                   // i-atom indices (system-local, -1: dummy)
    __shared__ unsigned long long sXm[W][EXC ? 64 : 1];    // their near exception masks
    int g0 = 0, ncl = 0;
    float zlo = 3.0e38f, zhi = -3.0e38f;  // z-range of the group's real i-atoms
    uint32_t ivm = 0;       // bit t: i-atom (t, ii) is a real atom
    uint64_t ivLanes = 0;   // bit 8 t + ii: the same, over the group's 64 slots
    if (active) {
        const int4 gi = P.ginfo[D.groupOff + g];
        g0 = __builtin_amdgcn_readfirstlane(gi.x);
        ncl = __builtin_amdgcn_readfirstlane(gi.y);
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        float2 q = make_float2(0.f, 0.f);
        double4 qd = make_double4(0.0, 0.0, 0.0, 0.0);
        bool v = false;
        int ai = -1;
        unsigned long long xm = 0ull;
        if (lane < ncl * kClusterSize) {
            const int s = D.slotOff + g0 * kClusterSize + lane;
            p = P.spos[s];
            if (F64) qd = P.sprm64[s]; else q = P.sprm[s];
            ai = P.sidx[s];
            v = ai >= 0;
            if (EXC && v && P.xmask) xm = P.xmask[D.atomOff + ai];  // no mask: no pair is excepted
        }
        if (EXC) {
            sAi[wave][lane] = ai;
            sXm[wave][lane] = xm;
        }
        sPos[wave][lane] = p;
        if (F64) sPrm64[wave][lane] = qd; else sPrm[wave][lane] = q;
        zlo = v ? p.z : zlo;
        zhi = v ? p.z : zhi;
        for (int o = 32; o > 0; o >>= 1) {
            zlo = fminf(zlo, __shfl_xor(zlo, o, 64));
            zhi = fmaxf(zhi, __shfl_xor(zhi, o, 64));
        }
        const uint64_t vb = __ballot(v);
        ivLanes = vb;
#pragma unroll
        for (int t = 0; t < kGroupClusters; ++t) ivm |= (uint32_t)((vb >> (t * kClusterSize + ii)) & 1ull) << t;
    }
    __syncthreads();
    if (!active) return;
    // wave-uniform from here on: keep the group's values in scalar registers
    g0 = __builtin_amdgcn_readfirstlane(g0);
    ncl = __builtin_amdgcn_readfirstlane(ncl);
    ivLanes = (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)ivLanes) |
              (uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(ivLanes >> 32)) << 32;
    const uint64_t upperLanes = __builtin_amdgcn_ballot_w64(ii < jj);  // i < j inside a diagonal cluster

    const float Lz = D.L[2], invLz = D.invL[2];
    const float Lx = D.L[0], invLx = D.invL[0], Ly = D.L[1], invLy = D.invL[1];
    const bool shiftOK = D.shiftOK;
    const double rc2 = D.rc2, krf64 = D.krf64, crf64 = D.crf64;
    const float krf = D.krf, crf = D.crf, fixMax = D.fixMax;
    const double wgt = BWD ? (wset ? wset[c] : 1.0) : 0.0;  // no weights: dU/dp of each set
    float fi[kGroupClusters][3];
#pragma unroll
    for (int t = 0; t < kGroupClusters; ++t) fi[t][0] = fi[t][1] = fi[t][2] = 0.f;
    double eE = 0.0;  // this lane's sum of whole pair energies
    uint32_t npair = 0;  // in-cutoff pairs evaluated by this wave (wave-uniform)
    const int nall = __builtin_amdgcn_readfirstlane(P.gcount[D.groupOff + g]);
    const int eb = (int)(((long long)part * nall) / split), ne = (int)(((long long)(part + 1) * nall) / split) - eb;
    const uint32_t* E = P.entries + D.entryOff + (long long)g * D.entryCap + eb;
    // This lane's j-atom of one list entry.  Two of these are in flight: entry
    // k + 1's loads are issued before entry k is evaluated, into the other
    // buffer, so the wait for them falls after entry k's work (and not behind
    // entry k's force atomics, which a single rotating buffer would wait for).
    struct JBuf {
        float4 p;
        double4 d;
        float2 q;
        int idx;
    };
    auto load_jbuf = [&](uint32_t ent) {
        JBuf b;
        const int js = D.slotOff + (int)(ent & 0xFFFFFu) * kClusterSize + jj;
        b.p = P.spos[js];
        b.d = make_double4(0.0, 0.0, 0.0, 0.0);
        b.q = make_float2(0.f, 0.f);
        if (F64) b.d = P.sprm64[js]; else b.q = P.sprm[js];
        // with explicit x/y image shifts, dummy slots sit far outside the box
        // (k_gather_bbox), so the FP64 path needs no validity test there
        b.idx = (F64 && shiftOK && !EXC) ? 0 : P.sidx[js];
        return b;
    };
    // lanes of `in` whose pair (i-atom (t, ii), j-atom aj) is an exception
    auto excepted = [&](int t, int aj, uint64_t in) __attribute__((always_inline)) {
        const int ai = sAi[wave][t * kClusterSize + ii];
        const unsigned long long mi = sXm[wave][t * kClusterSize + ii];
        const uint32_t dd = (uint32_t)(aj - ai + 32);
        const bool near = dd - 1u < 63u;
        bool ex = near && ((mi >> (dd & 63u)) & 1ull);
        const uint64_t farLanes = __builtin_amdgcn_ballot_w64(!near && ((mi >> 32) & 1ull)) & in;
        if (farLanes != 0ull && __builtin_amdgcn_inverse_ballot_w64(farLanes)) ex = far_excluded(P, D, ai, aj);
        return __builtin_amdgcn_ballot_w64(ex) & in;
    };
    // j-force sums of the last kJFlush entries, flushed together: one dense round
    // per component, lane 8 e + a adding atom a of entry e (8 consecutive slots,
    // one cache line, per 8 lanes)
    constexpr int kJFlush = 8;
    static_assert(kJFlush * kClusterSize == 64, "one flush round covers the wave");
    __shared__ float sJF[W][3][kJFlush * kClusterSize];  // [component][entry * 8 + jj]
    __shared__ int sJS[W][kJFlush];                       // the entry's first j-slot
    // the component each lane keeps after oct_sum3 (x: ii = 0, 7; y: 1, 6; z: 2, 5)
    const bool selX = ii == 0 || ii == 7, selY = ii == 1 || ii == 6;
    int ecount = 0;
    // cnt: entries held (kJFlush inside the entry loop, fewer at the end)
    auto flush_j = [&](int cnt) __attribute__((always_inline)) {
        wave_sync();  // the octets' writes are visible to every lane of the wave
        if (cnt == kJFlush || (lane >> 3) < cnt) {
            const int slot = sJS[wave][lane >> 3] + (lane & 7);
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const float f = sJF[wave][k][lane];
                if (BWD) {
                    double* acc = k == 0 ? P.gq : (k == 1 ? P.gsig : P.geps);
                    if (f != 0.f) atomicAdd(acc + slot, wgt * (double)f);
                } else {
                    flush_component_u(P, c, fixMax, slot, k, f);
                }
            }
        }
        wave_sync();  // read before the next entries overwrite them
    };
    // zdir: every i-atom lies at least rc inside [0, L_z), so the in-cutoff image of any j-atom
    // (wrapped into [0, L_z)) is the unshifted one in z: a j-atom within rc of an i-atom has
    // z in [z_i - rc, z_i + rc], inside [0, L_z), where no other z-image of it lies; pairs
    // outside the cutoff stay outside at the unshifted distance.  Such groups skip the z image
    // (and, in a triclinic box, the c step of the image rule: n_c = 0 for every in-cutoff pair).
    auto do_entry = [&](uint32_t ent, const JBuf& b, auto zdir) __attribute__((always_inline)) {
        const int J = ent & 0xFFFFF;
        const uint32_t imask = (ent >> 20) & 0xFFu;
        const int js = D.slotOff + J * kClusterSize + jj;
        float4 pj = b.p;
        const bool jv = b.idx >= 0;
        const float2 qj = b.q;
        const double4 jd = b.d;
        if (shiftOK) {
            const uint32_t sxc = (ent >> 28) & 3u, syc = (ent >> 30) & 3u;
            pj.x += sxc == 1u ? D.L[0] : (sxc == 2u ? -D.L[0] : 0.f);
            pj.y += syc == 1u ? D.L[1] : (syc == 2u ? -D.L[1] : 0.f);
        }
        const int tdiag = J - g0;  // the i-cluster equal to J, if any (its lanes count only i < j)
        // lanes with a real j-atom, as a wave mask (FP64 path with explicit shifts:
        // all of them, dummies sit far away)
        const uint64_t jvLanes = (F64 && shiftOK) ? ~0ull : __builtin_amdgcn_ballot_w64(jv);
        float fjx = 0.f, fjy = 0.f, fjz = 0.f;
        // i-atom positions read one step ahead (whether or not that step runs),
        // so the LDS latency falls behind the current step's work
        float4 piN = sPos[wave][ii];
#pragma unroll
        for (int t = 0; t < kGroupClusters; ++t) {
            const float4 pi = piN;
            if (t + 1 < kGroupClusters) piN = sPos[wave][(t + 1) * kClusterSize + ii];
            if (imask & (1u << t)) {
                float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                if (!shiftOK) {
                    if (EXC && D.tri) {
                        // z-direct groups: the in-cutoff image has n_c = 0
                        if constexpr (decltype(zdir)::value) tri_image_ab(D, dx, dy);
                        else tri_image_xy(D, dx, dy, dz);
                    } else {
                        dx = min_image(dx, Lx, invLx);
                        dy = min_image(dy, Ly, invLy);
                    }
                }
                if constexpr (!decltype(zdir)::value) dz = min_image(dz, Lz, invLz);
                float fs;
                if (F64) {
                    // The lane masks stay in scalar registers (ballot / inverse ballot).
                    // lanes whose i-atom (t, ii) is real: byte t of ivLanes repeated over the 8 jj
                    uint64_t iv = (ivLanes >> (t * kClusterSize)) & 0xFFull;
                    iv |= iv << 8; iv |= iv << 16; iv |= iv << 32;
                    const uint64_t okLanes = (t == tdiag ? upperLanes : ~0ull) & jvLanes & iv;
                    // exact: the squares of 2^-20-grid differences fit in 53 bits; decides the
                    // in-cutoff lanes and the early-out of a step with none
                    const double r2 = fma((double)dx, (double)dx, fma((double)dy, (double)dy, (double)dz * (double)dz));
                    uint64_t inLanes = __builtin_amdgcn_ballot_w64(r2 < rc2) & okLanes;
                    if (inLanes == 0ull) continue;  // no pair of this 8x8 step is inside
                    const double4 id = sPrm64[wave][t * kClusterSize + ii];
                    if (EXC && P.xmask) inLanes &= ~excepted(t, b.idx, inLanes);
                    npair += (uint32_t)__popcll(inLanes);
                    // lanes outside the cutoff are masked off (exec), not zeroed with selects
                    fs = 0.f;
                    if (__builtin_amdgcn_inverse_ballot_w64(inLanes)) {
                        if (BWD) {
                            double y = __builtin_amdgcn_rsq(r2);
                            {
                                const double tt = r2 * y;
                                const double ee = fma(-tt, y, 1.0);
                                y = fma(0.5 * y, ee, y);
                            }
                            const double u = (id.x + jd.x) * (y * y);
                            const double s2 = (id.x + jd.x) * u;
                            const double s4 = s2 * s2, s6 = s4 * s2;
                            const double lj = fma(s6, s6, -s6);
                            const double phi = fma(krf64, r2, y) - crf64;
                            const float dh = (float)(id.y * jd.y * fma(12.0, s6, -6.0) * s4 * u);
                            eE += fma(id.y * jd.y, lj, id.z * jd.z * phi);  // the pair energy
                            // (q', h, e') sums: i side in fi[t], j side in fjx, fjy, fjz
                            fi[t][0] += (float)(jd.z * phi); fjx += (float)(id.z * phi);
                            fi[t][1] += dh; fjy += dh;
                            fi[t][2] += (float)(jd.y * lj); fjz += (float)(id.y * lj);
                        } else {
                            const PairEF64 pe = pair_ef64(true, r2, id.x + jd.x, id.y * jd.y, id.z * jd.z, krf64, crf64);
                            eE += pe.e;
                            fs = (float)pe.fs;
                        }
                    }
                } else {
                    uint64_t iv = (ivLanes >> (t * kClusterSize)) & 0xFFull;
                    iv |= iv << 8; iv |= iv << 16; iv |= iv << 32;
                    const uint64_t okLanes = (t == tdiag ? upperLanes : ~0ull) & jvLanes & iv;
                    const float2 qi = sPrm[wave][t * kClusterSize + ii];
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    uint64_t inLanes = __builtin_amdgcn_ballot_w64(in_cutoff(D, dx, dy, dz, r2)) & okLanes;
                    if (EXC && P.xmask && inLanes != 0ull) inLanes &= ~excepted(t, b.idx, inLanes);
                    if (inLanes == 0ull) continue;
                    npair += (uint32_t)__popcll(inLanes);
                    const bool in = __builtin_amdgcn_inverse_ballot_w64(inLanes);
                    const PairE pe = pair_term(in, r2, pi.w * pj.w, qi.x + qj.x, qi.y * qj.y, krf);
                    // the whole pair energy in FP32 (its terms carry FP32 errors anyway), summed in FP64
                    eE += (double)(pe.elj + pe.qr + fmaf(krf, pe.qr2, -crf * pe.qq));
                    fs = pe.fs;
                }
                if (FORCES) {
                    fi[t][0] = fmaf(fs, dx, fi[t][0]); fi[t][1] = fmaf(fs, dy, fi[t][1]); fi[t][2] = fmaf(fs, dz, fi[t][2]);
                    fjx = fmaf(-fs, dx, fjx); fjy = fmaf(-fs, dy, fjy); fjz = fmaf(-fs, dz, fjz);
                }
            }
        }
        if (JSUMS) {
            // lanes ii = 0, 1, 2 of each jj octet keep the x, y, z sums (zero for dummy
            // j-slots, whose accumulators nothing reads)
            const float fk = oct_sum3(fjx, fjy, fjz, selX, selY);
            const int pe = ecount & (kJFlush - 1);
            if (ii < 3) sJF[wave][ii][pe * kClusterSize + jj] = fk;
            if (lane == 0) sJS[wave][pe] = js - jj;
            if (pe == kJFlush - 1) flush_j(kJFlush);
        }
        ++ecount;
    };
    // Entries are taken in pairs without conditions, so that the loads in flight
    // between a buffer's load and its use are the same on every path.  Past the
    // list end the batch holds entry 0 (cluster 0, empty i-mask): no pair step
    // runs and its j sums are zero; the load after a batch's last pair wraps to
    // lane 0 and is not used.
    auto entries = [&](auto zdir) __attribute__((always_inline)) {
        for (int e0 = 0; e0 < ne; e0 += 64) {
            const uint32_t ebuf = (e0 + lane < ne) ? E[e0 + lane] : 0u;
            const int nb = min(64, ne - e0);
            JBuf bufA = load_jbuf(__builtin_amdgcn_readlane(ebuf, 0)), bufB;
            for (int k = 0; k < nb; k += 2) {
                bufB = load_jbuf(__builtin_amdgcn_readlane(ebuf, k + 1));
                do_entry(__builtin_amdgcn_readlane(ebuf, k), bufA, zdir);
                bufA = load_jbuf(__builtin_amdgcn_readlane(ebuf, (k + 2) & 63));
                do_entry(__builtin_amdgcn_readlane(ebuf, k + 1), bufB, zdir);
            }
        }
    };
    // the i-atoms' z-range at least the cutoff (+ the list margin) inside the box in z; the
    // FP32 D.rc carries its rounding margin.  Without EXC only sets with explicit x/y shifts
    // take the z-direct loop (that loop then carries no x/y image code: fewer registers)
    const float rcz = D.rc * (1.0f + 1e-5f) + 1e-4f;
    const bool zDirect = (EXC || shiftOK) &&
                         __builtin_amdgcn_readfirstlane((int)(zlo - rcz >= 0.f && zhi + rcz < Lz)) != 0;
    if (zDirect) entries(std::true_type{}); else entries(std::false_type{});
    if (JSUMS && (ecount & (kJFlush - 1)) != 0) flush_j(ecount & (kJFlush - 1));
    if (JSUMS) {
#pragma unroll
        for (int t = 0; t < kGroupClusters; ++t) {
            if (t < ncl) {
                const float fx = stride8_sum(fi[t][0]), fy = stride8_sum(fi[t][1]), fz = stride8_sum(fi[t][2]);
                const int is = D.slotOff + (g0 + t) * kClusterSize + ii;
                if (jj == 0 && ((ivm >> t) & 1u)) {
                    if (BWD) {
                        if (fx != 0.f) atomicAdd(P.gq + is, wgt * (double)fx);
                        if (fy != 0.f) atomicAdd(P.gsig + is, wgt * (double)fy);
                        if (fz != 0.f) atomicAdd(P.geps + is, wgt * (double)fz);
                    } else {
                        flush_force(P, c, fixMax, is, fx, fy, fz);
                    }
                }
            }
        }
    }
    if (BWD && wset) return;  // a backward pass after the evaluation: energies and counts are the evaluation's
    double eacc = eE;
    eacc = wave_sum_d(eacc);
    if (lane == 0) {
        P.genergy[(long long)(D.groupOff + g) * split + part] = eacc;
        atomicAdd(&P.rt[c].npairs, (unsigned long long)npair);
    }
}

// ------------------------------------------------- atom-level j lists

// For every i-cluster t of a group, the j-atoms of the group's list entries
// (j-clusters in the entry's mask for t) that lie within the list cutoff of
// t's bounding box, appended in (entry, atom) order to t's list in 8-atom
// blocks.  Entry = j-slot within the set (23 bits) | x-shift code << 28 |
// y-shift code << 30; padding entries are kNoAtom.
constexpr uint32_t kNoAtom = 0xFFFFFFFFu;

__global__ void __launch_bounds__(64) k_alist(DevPtrs P, int setBegin) {
    // one wave per (group, i-cluster t): lane = 8 e + jj tests atom jj of list
    // entry e of the current 8-entry window against t's bounding box
    const int c = setBegin + (int)blockIdx.y * P.setStride;
    const SetDesc D = P.desc[c];
    const int g = blockIdx.x >> 3, t = blockIdx.x & 7, lane = threadIdx.x;
    if (g >= P.rt[c].ngroups) return;
    const int4 gi = P.ginfo[D.groupOff + g];
    const int g0 = __builtin_amdgcn_readfirstlane(gi.x), ncl = __builtin_amdgcn_readfirstlane(gi.y);
    if (t >= ncl) return;
    const int e = lane >> 3, jj = lane & 7;
    const float4 bc = P.bbc[D.clusterOff + g0 + t], bh = P.bbh[D.clusterOff + g0 + t];
    uint32_t* out = P.alist + D.alistOff + (long long)(g0 + t) * D.alistCap;
    int cnt = 0;
    const int ne = P.gcount[D.groupOff + g];
    const uint32_t* E = P.entries + D.entryOff + (long long)g * D.entryCap;
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    for (int e0 = 0; e0 < ne; e0 += 64) {
        const uint32_t ebuf = (e0 + lane < ne) ? E[e0 + lane] : 0u;
        const int nb = min(64, ne - e0);
        for (int w = 0; w < nb; w += 8) {
            const uint32_t ent = (uint32_t)__shfl((int)ebuf, w + e, 64);
            const bool has = w + e < nb && ((ent >> (20 + t)) & 1u);
            const int js = (int)(ent & 0xFFFFFu) * kClusterSize + jj;  // slot within the set
            bool keep = false;
            if (has && P.sidx[D.slotOff + js] >= 0) {
                float4 pj = P.spos[D.slotOff + js];
                if (D.shiftOK) {
                    const uint32_t sxc = (ent >> 28) & 3u, syc = (ent >> 30) & 3u;
                    pj.x += sxc == 1u ? D.L[0] : (sxc == 2u ? -D.L[0] : 0.f);
                    pj.y += syc == 1u ? D.L[1] : (syc == 2u ? -D.L[1] : 0.f);
                }
                float dx = pj.x - bc.x, dy = pj.y - bc.y, dz = pj.z - bc.z;
                if (!D.shiftOK) { dx = min_image(dx, D.L[0], D.invL[0]); dy = min_image(dy, D.L[1], D.invL[1]); }
                dz = min_image(dz, D.L[2], D.invL[2]);
                const float ax = fmaxf(0.f, fabsf(dx) - bh.x), ay = fmaxf(0.f, fabsf(dy) - bh.y);
                const float az = fmaxf(0.f, fabsf(dz) - bh.z);
                keep = ax * ax + ay * ay + az * az < D.rcList2;
            }
            const uint64_t m = __ballot(keep);
            const int pos = cnt + __popcll(m & below);
            if (keep) {
                if (pos < D.alistCap) out[pos] = (uint32_t)js | (ent & 0xF0000000u);
                else atomicOr(&P.rt[c].overflow, 2);
            }
            cnt += __popcll(m);
        }
    }
    // pad to whole 8-atom blocks
    const int padded = min((cnt + 7) & ~7, D.alistCap & ~7);
    if (lane < 8 && cnt + lane < padded) out[cnt + lane] = kNoAtom;
    if (lane == 0) {
        P.acount[D.clusterOff + g0 + t] = padded;
        atomicMax(&P.rt[c].maxAtoms, cnt);
    }
}

// One wave64 per group; for each i-cluster t of the group its atom list is
// walked in 8-atom blocks: lane = 8 jj + ii pairs i-atom (t, ii) with j-atom jj
// of the block (gathered per lane).  Pair math as the tile kernel's FP64 path;
// a pair counts when its j-slot is above its i-slot (half shell; inside the
// diagonal cluster i < j).
template <int W, bool FORCES>
__global__ void __launch_bounds__(64 * W) k_pairs_a(DevPtrs P, int setBegin) {
    const int c = setBegin + (int)blockIdx.y * P.setStride;
    const SetDesc D = P.desc[c];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jj = lane >> 3, ii = lane & 7;
    const int g = blockIdx.x * W + wave;
    const bool active = g < P.rt[c].ngroups;
    __shared__ float4 sPos[W][64];
    __shared__ double4 sPrm64[W][64];
    int g0 = 0, ncl = 0;
    uint32_t ivm = 0;
    if (active) {
        const int4 gi = P.ginfo[D.groupOff + g];
        g0 = __builtin_amdgcn_readfirstlane(gi.x);
        ncl = __builtin_amdgcn_readfirstlane(gi.y);
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        double4 qd = make_double4(0.0, 0.0, 0.0, 0.0);
        bool v = false;
        if (lane < ncl * kClusterSize) {
            const int s = D.slotOff + g0 * kClusterSize + lane;
            p = P.spos[s];
            qd = P.sprm64[s];
            v = P.sidx[s] >= 0;
        }
        sPos[wave][lane] = p;
        sPrm64[wave][lane] = qd;
        const uint64_t vb = __ballot(v);
#pragma unroll
        for (int t = 0; t < kGroupClusters; ++t) ivm |= (uint32_t)((vb >> (t * kClusterSize + ii)) & 1ull) << t;
    }
    __syncthreads();
    if (!active) return;
    const float Lz = D.L[2], invLz = D.invL[2];
    const float Lx = D.L[0], invLx = D.invL[0], Ly = D.L[1], invLy = D.invL[1];
    const bool shiftOK = D.shiftOK;
    const double rc2 = D.rc2, krf64 = D.krf64;
    const float rc2hi = D.rc2hi;
    double eE = 0.0, eQQ = 0.0, eQR2 = 0.0;
    uint32_t npair = 0;
    for (int t = 0; t < ncl; ++t) {
        const bool iv = (ivm >> t) & 1u;
        const float4 pi = sPos[wave][t * kClusterSize + ii];
        const double4 id = sPrm64[wave][t * kClusterSize + ii];
        const int islot = (g0 + t) * kClusterSize + ii;
        float fix = 0.f, fiy = 0.f, fiz = 0.f;
        const int na = __builtin_amdgcn_readfirstlane(P.acount[D.clusterOff + g0 + t]);
        const uint32_t* A = P.alist + D.alistOff + (long long)(g0 + t) * D.alistCap;
        for (int b0 = 0; b0 < na; b0 += 64) {
            const uint32_t abuf = (b0 + lane < na) ? A[b0 + lane] : kNoAtom;
            const int nblk = min(64, na - b0) >> 3;
            // software pipeline: the j-atom of block b + 1 is loaded while block b computes
            uint32_t entN = (uint32_t)__shfl((int)abuf, jj, 64);
            int jsN = entN != kNoAtom ? (int)(entN & 0x7FFFFFu) : 0;
            float4 pjN = P.spos[D.slotOff + jsN];
            double4 jdN = P.sprm64[D.slotOff + jsN];
            for (int b = 0; b < nblk; ++b) {
                const uint32_t ent = entN;
                const bool jv = ent != kNoAtom;
                const int js = jsN;
                float4 pj = pjN;
                const double4 jd = jdN;
                if (b + 1 < nblk) {
                    entN = (uint32_t)__shfl((int)abuf, (b + 1) * 8 + jj, 64);
                    jsN = entN != kNoAtom ? (int)(entN & 0x7FFFFFu) : 0;
                    pjN = P.spos[D.slotOff + jsN];
                    jdN = P.sprm64[D.slotOff + jsN];
                }
                if (shiftOK) {
                    const uint32_t sxc = (ent >> 28) & 3u, syc = (ent >> 30) & 3u;
                    pj.x += sxc == 1u ? D.L[0] : (sxc == 2u ? -D.L[0] : 0.f);
                    pj.y += syc == 1u ? D.L[1] : (syc == 2u ? -D.L[1] : 0.f);
                }
                float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                if (!shiftOK) { dx = min_image(dx, Lx, invLx); dy = min_image(dy, Ly, invLy); }
                dz = min_image(dz, Lz, invLz);
                const float r2f = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                const bool cand = r2f < rc2hi && iv && jv && js > islot;
                if (__builtin_amdgcn_ballot_w64(cand) == 0ull) continue;
                const double r2 = fma((double)dx, (double)dx, fma((double)dy, (double)dy, (double)dz * (double)dz));
                const bool in = cand && r2 < rc2;
                npair += (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(in));
                const PairE64 pe = pair_term64(in, r2, id.x + jd.x, id.y * jd.y, id.z * jd.z, krf64);
                eE += pe.e; eQQ += pe.qq; eQR2 += pe.qq * r2;
                if (FORCES) {
                    const float fs = (float)pe.fs;
                    const float fx = fs * dx, fy = fs * dy, fz = fs * dz;
                    fix += fx; fiy += fy; fiz += fz;
                    const float fjx = oct_sum(-fx), fjy = oct_sum(-fy), fjz = oct_sum(-fz);
                    if (ii == 0 && jv && (fjx != 0.f || fjy != 0.f || fjz != 0.f))
                        flush_force(P, c, D.fixMax, D.slotOff + js, fjx, fjy, fjz);
                }
            }
        }
        if (FORCES) {
            const float fx = stride8_sum(fix), fy = stride8_sum(fiy), fz = stride8_sum(fiz);
            if (jj == 0 && iv) flush_force(P, c, D.fixMax, D.slotOff + islot, fx, fy, fz);
        }
    }
    double eacc = eE + D.krf64 * eQR2 - D.crf64 * eQQ;
    eacc = wave_sum_d(eacc);
    if (lane == 0) {
        P.genergy[D.groupOff + g] = eacc;
        atomicAdd(&P.rt[c].npairs, (unsigned long long)npair);
    }
}

// ------------------------------------------------- compacted pair kernel

// Order LDS accesses of the lanes of one wave (LDS executes a wave's
// operations in issue order; this keeps the compiler from moving them).

// FP64 -> 64-bit fixed point (round to nearest) for |f| < 2^27 kJ/mol/nm.
__device__ __forceinline__ unsigned long long fixed64(double f) {
    const double t = fma(f, kForceScale, 6755399441055744.0);  // 1.5 * 2^52
    return (unsigned long long)(__double_as_longlong(t) - 0x4338000000000000LL);
}

__device__ __forceinline__ void lds_add_u64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

__device__ __forceinline__ void lds_sub_u64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_fetch_sub(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

// One wave64 per group, two passes per batch of JB list entries.
//  test:  lane = 8 jj + ii tests i-atom (t, ii) against j-atom jj of each
//         entry's j-cluster for every i-cluster t in the entry's mask, with an
//         FP32 r^2 and a margin above rc^2 (it never drops a pair inside the
//         cutoff), and appends the passing pairs as (i, j) codes to a ring in
//         LDS (ballot + mbcnt: order = entry, t, lane);
//  dense: whenever 64 pairs are pending each lane takes one queued pair,
//         decides r < rc on the exact FP64 r^2 and evaluates the FP64 pair
//         function, so the expensive part runs on in-cutoff pairs instead of
//         whole 8x8 tiles.
// The j-atoms of a batch are staged in one half of a two-half LDS buffer, so
// a batch's pairs can still be pending while the next batch loads: chunks stay
// full across batch boundaries.
// Forces: per pair FP64 components -> 64-bit fixed point -> integer LDS
// atomics into per-atom accumulators of the group (i) and of the batches (j),
// flushed to the global fixed-point accumulators with integer atomics.  The j
// accumulators exist in NCOPY copies (copy = lane % NCOPY, rows in different
// banks): a chunk holds few distinct j-atoms, and same-address LDS atomics
// serialise.  Integer sums do not depend on the order of the atomics, so
// results are bit-reproducible.  Components >= 2^27 kJ/mol/nm take FP64
// global atomics.
// STEP: drain after every test step (the queue then never holds more than 127
// pairs and needs 128 slots instead of 1024; the test loop runs over the set
// bits of the mask and reads the i-positions from LDS).
template <int W, int JB, int NCOPY, bool STEP, bool FORCES>
__global__ void __launch_bounds__(64 * W) k_pairs_q(DevPtrs P, int setBegin) {
    constexpr int NH = JB * kClusterSize;  // j-atoms per batch (one half of the j buffer)
    constexpr int NJ = 2 * NH;
    constexpr int JS = NJ + 1;             // odd row stride of the j accumulators
    constexpr int QCAP = STEP ? 128 : 1024;  // pending pairs < 64 + 64 (STEP) or 64 + 8 * 64
    const int c = setBegin + (int)blockIdx.y * P.setStride;
    const SetDesc D = P.desc[c];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int jj = lane >> 3, ii = lane & 7;
    const int g = blockIdx.x * W + wave;
    if (g >= P.rt[c].ngroups) return;  // waves synchronise only within themselves

    __shared__ float sIX[W][64], sIY[W][64], sIZ[W][64];
    __shared__ double sIH[W][64], sIE[W][64], sIQ[W][64];  // sigma/2, 2 sqrt(eps), q sqrt(k_e)
    __shared__ float sJX[W][NJ], sJY[W][NJ], sJZ[W][NJ], sJV[W][NJ];  // V = 1: real atom
    __shared__ double sJH[W][NJ], sJE[W][NJ], sJQ[W][NJ];
    __shared__ int sJSlot[W][NJ];
    __shared__ unsigned long long sIF[W][3][FORCES ? 64 : 1];
    __shared__ unsigned long long sJF[W][FORCES ? NCOPY * 3 * JS : 1];
    __shared__ unsigned short sQ[W][QCAP];

    const int4 gi = P.ginfo[D.groupOff + g];
    const int g0 = __builtin_amdgcn_readfirstlane(gi.x), ncl = __builtin_amdgcn_readfirstlane(gi.y);
    const int sBase = D.slotOff + g0 * kClusterSize;
    bool iv = false;
    {
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        double4 q = make_double4(0.0, 0.0, 0.0, 0.0);
        if (lane < ncl * kClusterSize) {
            p = P.spos[sBase + lane];
            q = P.sprm64[sBase + lane];
            iv = P.sidx[sBase + lane] >= 0;
        }
        sIX[wave][lane] = p.x; sIY[wave][lane] = p.y; sIZ[wave][lane] = p.z;
        sIH[wave][lane] = q.x; sIE[wave][lane] = q.y; sIQ[wave][lane] = q.z;
        if (FORCES) sIF[wave][0][lane] = sIF[wave][1][lane] = sIF[wave][2][lane] = 0ull;
    }
    const uint64_t ivb = __ballot(iv);
    uint32_t ivm = 0;  // bit t: i-atom (t, ii) is a real atom
#pragma unroll
    for (int t = 0; t < kGroupClusters; ++t) ivm |= (uint32_t)((ivb >> (t * kClusterSize + ii)) & 1ull) << t;
    wave_sync();
    float pix[STEP ? 1 : kGroupClusters], piy[STEP ? 1 : kGroupClusters], piz[STEP ? 1 : kGroupClusters];
    if (!STEP) {
#pragma unroll
        for (int t = 0; t < kGroupClusters; ++t) {
            pix[t] = sIX[wave][t * kClusterSize + ii];
            piy[t] = sIY[wave][t * kClusterSize + ii];
            piz[t] = sIZ[wave][t * kClusterSize + ii];
        }
    }

    const float Lx = D.L[0], invLx = D.invL[0], Ly = D.L[1], invLy = D.invL[1], Lz = D.L[2], invLz = D.invL[2];
    const bool shiftOK = D.shiftOK;
    const float rc2hi = D.rc2hi;
    const double rc2 = D.rc2, krf64 = D.krf64;
    const double fixMax = D.fixMax;
    double eE = 0.0, eQQ = 0.0, eQR2 = 0.0;
    uint32_t npair = 0;
    const int ne = __builtin_amdgcn_readfirstlane(P.gcount[D.groupOff + g]);
    const uint32_t* E = P.entries + D.entryOff + (long long)g * D.entryCap;
    int qh = 0, qn = 0;  // ring head and tail (wave-uniform, never reset)

    // Evaluate the n (<= 64) pairs at the ring head.
    auto dense = [&](int n) {
        wave_sync();
        const bool act = lane < n;
        const int code = act ? (int)sQ[wave][(qh + lane) & (QCAP - 1)] : 0;
        const int i = code & 63, j = code >> 6;
        float dx = sIX[wave][i] - sJX[wave][j], dy = sIY[wave][i] - sJY[wave][j], dz = sIZ[wave][i] - sJZ[wave][j];
        if (!shiftOK) { dx = min_image(dx, Lx, invLx); dy = min_image(dy, Ly, invLy); }
        dz = min_image(dz, Lz, invLz);
        const double dxd = dx, dyd = dy, dzd = dz;
        // exact: the squares of 2^-20-grid differences fit in 53 bits
        const double r2 = fma(dxd, dxd, fma(dyd, dyd, dzd * dzd));
        const bool in = act && r2 < rc2;
        npair += (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(in));
        const PairE64 pe = pair_term64(in, r2, sIH[wave][i] + sJH[wave][j], sIE[wave][i] * sJE[wave][j],
                                       sIQ[wave][i] * sJQ[wave][j], krf64);
        eE += pe.e; eQQ += pe.qq; eQR2 += pe.qq * r2;
        if (FORCES && in) {
            const double fx = pe.fs * dxd, fy = pe.fs * dyd, fz = pe.fs * dzd;
            if (fmax(fabs(fx), fmax(fabs(fy), fabs(fz))) < fixMax) {  // <= 2^27
                const unsigned long long vx = fixed64(fx), vy = fixed64(fy), vz = fixed64(fz);
                lds_add_u64(&sIF[wave][0][i], vx);
                lds_add_u64(&sIF[wave][1][i], vy);
                lds_add_u64(&sIF[wave][2][i], vz);
                unsigned long long* jf = &sJF[wave][(lane % NCOPY) * 3 * JS + j];
                lds_sub_u64(jf, vx);
                lds_sub_u64(jf + JS, vy);
                lds_sub_u64(jf + 2 * JS, vz);
            } else {
                double* di = P.sd + 3ll * (sBase + i);
                double* dj = P.sd + 3ll * sJSlot[wave][j];
                atomicAdd(di, fx); atomicAdd(di + 1, fy); atomicAdd(di + 2, fz);
                atomicAdd(dj, -fx); atomicAdd(dj + 1, -fy); atomicAdd(dj + 2, -fz);
                mark_side(P, c, sBase + i);
                mark_side(P, c, sJSlot[wave][j]);
            }
        }
        qh += n;
    };
    // Sum the copies of the j-forces of the batch in half h into the global accumulators.
    auto flush_half = [&](int h) {
        if (!FORCES) return;
#pragma unroll
        for (int r = 0; r < (NH + 63) / 64; ++r) {
            const int x = h * NH + r * 64 + lane;
            if (r * 64 + lane < NH && sJV[wave][x] != 0.f) {
                unsigned long long v[3] = {0ull, 0ull, 0ull};
#pragma unroll
                for (int cp = 0; cp < NCOPY; ++cp)
#pragma unroll
                    for (int k = 0; k < 3; ++k) v[k] += sJF[wave][(cp * 3 + k) * JS + x];
                if (v[0] | v[1] | v[2]) {
                    unsigned long long* f = reinterpret_cast<unsigned long long*>(P.sf) + sJSlot[wave][x];
                    atomicAdd(f, v[0]);
                    atomicAdd(f + P.sfn, v[1]);
                    atomicAdd(f + 2 * P.sfn, v[2]);
                }
            }
        }
    };

    int bstart0 = 0, bstart1 = 0;  // ring position where the batch in half 0 / 1 started
    int bi = 0;
    for (int b0 = 0; b0 < ne; b0 += JB, ++bi) {
        const int h = bi & 1;
        const int nb = min(JB, ne - b0);
        if (bi >= 2) {
            // half h still holds batch bi - 2, whose pairs end where batch bi - 1 began
            if (qh < (h ? bstart0 : bstart1)) dense(qn - qh);
            wave_sync();
            flush_half(h);
        }
        wave_sync();
        uint32_t entReg = 0u;  // lane l holds entry l >> 3 of the batch
        {
            const int k = jj;
            const int x = h * NH + lane;
            if (lane < NH) {
                uint32_t ent = 0u;
                float4 pj = make_float4(0.f, 0.f, 0.f, 0.f);
                double4 prm = make_double4(0.0, 0.0, 0.0, 0.0);
                int js = 0;
                bool jv = false;
                if (k < nb) {
                    ent = E[b0 + k];
                    js = D.slotOff + (int)(ent & 0xFFFFFu) * kClusterSize + ii;
                    pj = P.spos[js];
                    prm = P.sprm64[js];
                    jv = P.sidx[js] >= 0;
                    if (shiftOK) {
                        const uint32_t sxc = (ent >> 28) & 3u, syc = (ent >> 30) & 3u;
                        pj.x += sxc == 1u ? Lx : (sxc == 2u ? -Lx : 0.f);
                        pj.y += syc == 1u ? Ly : (syc == 2u ? -Ly : 0.f);
                    }
                }
                entReg = ent;
                sJX[wave][x] = pj.x; sJY[wave][x] = pj.y; sJZ[wave][x] = pj.z; sJV[wave][x] = jv ? 1.f : 0.f;
                sJH[wave][x] = prm.x; sJE[wave][x] = prm.y; sJQ[wave][x] = prm.z;
                sJSlot[wave][x] = js;
                if (FORCES) {
#pragma unroll
                    for (int r = 0; r < NCOPY * 3; ++r) sJF[wave][r * JS + x] = 0ull;
                }
            }
        }
        if (h) bstart1 = qn; else bstart0 = qn;
        wave_sync();
        for (int k = 0; k < nb; ++k) {
            const uint32_t ent = __builtin_amdgcn_readlane(entReg, k * 8);
            const int td = (int)(ent & 0xFFFFFu) - g0;  // the i-cluster equal to J, if any
            const uint32_t imask = (ent >> 20) & 0xFFu;
            const int jx = h * NH + k * 8 + jj;
            const float pjx = sJX[wave][jx], pjy = sJY[wave][jx], pjz = sJZ[wave][jx];
            uint32_t ok = sJV[wave][jx] != 0.f ? ivm : 0u;
            if (td >= 0 && td < kGroupClusters && ii >= jj) ok &= ~(1u << td);
            const unsigned short code0 = (unsigned short)(ii | (jx << 6));
            if (STEP) {
                uint32_t live = imask;
                while (live) {
                    const int t = __builtin_ctz(live);
                    live &= live - 1u;
                    const int ia = t * kClusterSize + ii;
                    float dx = sIX[wave][ia] - pjx, dy = sIY[wave][ia] - pjy, dz = sIZ[wave][ia] - pjz;
                    if (!shiftOK) { dx = min_image(dx, Lx, invLx); dy = min_image(dy, Ly, invLy); }
                    dz = min_image(dz, Lz, invLz);
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    const bool in = r2 < rc2hi && ((ok >> t) & 1u);
                    const uint64_t m = __ballot(in);
                    if (m == 0ull) continue;
                    if (in) {
                        const int pos = qn + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        sQ[wave][pos & (QCAP - 1)] = (unsigned short)(code0 + t * 8);
                    }
                    qn += __popcll(m);
                    if (qn - qh >= 64) dense(64);
                }
            } else {
#pragma unroll
                for (int t = 0; t < kGroupClusters; ++t) {
                    if (!(imask & (1u << t))) continue;
                    float dx = pix[t] - pjx, dy = piy[t] - pjy, dz = piz[t] - pjz;
                    if (!shiftOK) { dx = min_image(dx, Lx, invLx); dy = min_image(dy, Ly, invLy); }
                    dz = min_image(dz, Lz, invLz);
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    const bool in = r2 < rc2hi && ((ok >> t) & 1u);
                    const uint64_t m = __ballot(in);
                    if (m == 0ull) continue;
                    if (in) {
                        const int pos = qn + (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                                             __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
                        sQ[wave][pos & (QCAP - 1)] = (unsigned short)(code0 + t * 8);
                    }
                    qn += __popcll(m);
                }
                while (qn - qh >= 64) dense(64);
            }
        }
    }
    if (qn > qh) dense(qn - qh);
    wave_sync();
    if (bi >= 1) flush_half((bi - 1) & 1);
    if (bi >= 2) flush_half(bi & 1);
    if (FORCES && iv) {
        const unsigned long long vx = sIF[wave][0][lane], vy = sIF[wave][1][lane], vz = sIF[wave][2][lane];
        if (vx | vy | vz) {
            unsigned long long* f = reinterpret_cast<unsigned long long*>(P.sf) + (sBase + lane);
            atomicAdd(f, vx);
            atomicAdd(f + P.sfn, vy);
            atomicAdd(f + 2 * P.sfn, vz);
        }
    }
    double eacc = eE + D.krf64 * eQR2 - D.crf64 * eQQ;
    eacc = wave_sum_d(eacc);
    if (lane == 0) {
        P.genergy[D.groupOff + g] = eacc;
        atomicAdd(&P.rt[c].npairs, (unsigned long long)npair);
    }
}

// Parity/statistics: the same traversal, emitting in-cutoff pairs as original
// atom indices (i < j) and counting evaluated candidates.
__global__ void __launch_bounds__(64) k_pairs_dump(DevPtrs P, int c, int2* out, long long cap,
                                                   unsigned long long* counters) {
    const SetDesc& D = P.desc[c];
    const int lane = threadIdx.x & 63, jj = lane >> 3, ii = lane & 7;
    const int g = blockIdx.x;
    if (g >= P.rt[c].ngroups) return;
    const int4 gi = P.ginfo[D.groupOff + g];
    const int g0 = gi.x, ncl = gi.y;
    const int ne = P.gcount[D.groupOff + g];
    const uint32_t* E = P.entries + D.entryOff + (long long)g * D.entryCap;
    unsigned long long nin = 0, nev = 0;
    for (int e = 0; e < ne; ++e) {
        const uint32_t ent = E[e];
        const int J = ent & 0xFFFFF;
        const uint32_t imask = (ent >> 20) & 0xFFu;
        const JAtom ja = load_j(P, D, ent, jj);
        const int jidx = P.sidx[ja.slot];
        for (int t = 0; t < ncl; ++t) {
            if (!(imask & (1u << t))) continue;
            nev += 1;
            const int si = D.slotOff + (g0 + t) * kClusterSize + ii;
            const float4 pi = P.spos[si];
            const int iidx = P.sidx[si];
            float dx = pi.x - ja.p.x, dy = pi.y - ja.p.y, dz = pi.z - ja.p.z;
            pair_image(D, dx, dy, dz);
            const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
            bool in = in_cutoff(D, dx, dy, dz, r2) && ja.valid && iidx >= 0;
            if (g0 + t == J) in = in && (ii < jj);
            if (in && P.xmask) in = !excluded_pair(P, D, iidx, jidx, P.xmask[D.atomOff + iidx]);
            if (in) {
                nin += 1;
                unsigned long long k = atomicAdd(counters, 1ull);
                if ((long long)k < cap) out[k] = make_int2(min(iidx, jidx), max(iidx, jidx));
            }
        }
    }
    atomicAdd(counters + 1, nev);  // per lane: candidate pairs evaluated
    atomicAdd(counters + 2, nin);  // per lane: in-cutoff pairs
}

// ------------------------------------------------- filtered atom queues
//
// One wave64 per i-cluster t of a group (grid: 8 waves per group).  The wave
// keeps the group's list entries whose mask has bit t, tests the 8 atoms of 8
// kept entries at a time (lane = 8 k + jj: atom jj of kept entry k) against
// t's bounding box grown by the list cutoff, and appends the atoms that pass
// to a queue in LDS: position relative to a grid point r0 at the box centre
// (the image of each atom nearest r0, exact on the 2^-20 grid), j-slot and
// FP64 parameters.  Every 8 queued atoms make one 8 x 8 pair step
// (lane = 8 jj + ii: queued atom jj with i-atom ii) with the exact FP64 r^2,
// the exact r < rc and the FP64 pair function; a pair counts when its j-slot
// is above its i-slot (each entry's j-cluster is at or above t).  The atom
// test keeps ~45 % of a step's lane pairs inside the cutoff at 23k atoms
// against ~21 % for whole 8 x 8 cluster tiles.
//
// Nearest image to r0 = pairwise minimum image for every pair inside the
// cutoff when rc + (box half-extent) < L/2 in each dimension: |x_j - r0| <
// rc + h <= L/2 for the in-cutoff image.  Clusters for which that fails (small
// boxes, sparse columns) accept every atom in that dimension and apply the
// pairwise minimum image in their steps.
constexpr int kQueueCap = 72;  // 7 left over from the last step + one test round of 64

constexpr int kFiltEnergy = 0, kFiltForces = 1;

#ifndef ATM_FILT_MINW
#define ATM_FILT_MINW 1
#endif
template <int W, int MODE>
__global__ void __launch_bounds__(64 * W, ATM_FILT_MINW) k_pairs_f(DevPtrs P, int setBegin) {
    constexpr bool FORCES = MODE == kFiltForces;
    const int c = setBegin + (int)blockIdx.y * P.setStride;
    const SetDesc D = P.desc[c];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int gw = blockIdx.x * W + wave;
    const int g = gw >> 3, t = gw & 7;
    if (g >= P.rt[c].ngroups) return;
    double* eOut = P.genergy + 8ll * (D.groupOff + g) + t;
    const int4 gi = P.ginfo[D.groupOff + g];
    const int g0 = __builtin_amdgcn_readfirstlane(gi.x), ncl = __builtin_amdgcn_readfirstlane(gi.y);
    if (t >= ncl) {
        if (lane == 0) *eOut = 0.0;
        return;
    }
    __shared__ uint32_t sSel[W][64];              // kept entries of the current batch (j-cluster)
    __shared__ float4 sQp[W][kQueueCap];          // queued atom: position - r0, j-slot (bits of .w)
    __shared__ double2 sQa[W][kQueueCap];         // (sigma/2, 2 sqrt(eps))
    __shared__ double sQq[W][kQueueCap];          // q sqrt(k_e)
    constexpr int kJFlush = 8;
    __shared__ float sJF[W][kJFlush][kClusterSize * 3];
    __shared__ int sJS[W][kJFlush][kClusterSize];

    const int jj = lane >> 3, ii = lane & 7;
    const int I = g0 + t;                         // this wave's i-cluster (in the set)
    const float4 bc = P.bbc[D.clusterOff + I], bh = P.bbh[D.clusterOff + I];
    // reference point: the box centre rounded to the 2^-20 grid (exact for |x| < 16)
    const float r0x = rintf(bc.x * 1048576.0f) * (1.0f / 1048576.0f);
    const float r0y = rintf(bc.y * 1048576.0f) * (1.0f / 1048576.0f);
    const float r0z = rintf(bc.z * 1048576.0f) * (1.0f / 1048576.0f);
    const float rcl = sqrtf(D.rcList2);
    const float Lx = D.L[0], Ly = D.L[1], Lz = D.L[2];
    const float iLx = D.invL[0], iLy = D.invL[1], iLz = D.invL[2];
    // dimensions where the nearest image to r0 may differ from the pairwise one
    const bool imgX = rcl + bh.x + 1e-5f >= 0.5f * Lx, imgY = rcl + bh.y + 1e-5f >= 0.5f * Ly;
    const bool imgZ = rcl + bh.z + 1e-5f >= 0.5f * Lz;
    const bool pairImg = __builtin_amdgcn_readfirstlane(imgX || imgY || imgZ ? 1 : 0) != 0;  // wave-uniform
    const float hx = imgX ? 1e30f : bh.x, hy = imgY ? 1e30f : bh.y, hz = imgZ ? 1e30f : bh.z;

    // this lane's i-atom (ii), relative to r0; a dummy i-slot is moved far away
    const int islot = D.slotOff + I * kClusterSize + ii;
    float4 pi = P.spos[islot];
    const double4 ip = P.sprm64[islot];
    const bool iReal = P.sidx[islot] >= 0;
    const float rix = iReal ? pi.x - r0x : 1.0e6f, riy = iReal ? pi.y - r0y : 1.0e6f, riz = iReal ? pi.z - r0z : 1.0e6f;
    const double rc2 = D.rc2, krf64 = D.krf64, crf64 = D.crf64;
    const float fixMax = D.fixMax;

    float fix = 0.f, fiy = 0.f, fiz = 0.f;
    double eE = 0.0;
    uint32_t npair = 0;
    int qn = 0;      // queued atoms (wave-uniform)
    int nflush = 0;  // pair steps staged for the j-force flush

    // j-force sums of the last kJFlush steps, flushed together
    auto flush_j = [&](int cnt) __attribute__((always_inline)) {
        wave_sync();
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            const int v = r * 64 + lane;
            const int pe = v / (kClusterSize * 3), rem = v - pe * (kClusterSize * 3);
            if (pe < cnt) flush_component_u(P, c, fixMax, sJS[wave][pe][rem / 3], rem % 3, sJF[wave][pe][rem]);
        }
        wave_sync();
    };

    // one 8 x 8 step on queued atoms [b, b + nj)
    auto step = [&](int b, int nj) __attribute__((always_inline)) {
        const float4 rj = sQp[wave][b + jj];
        const double2 aj = sQa[wave][b + jj];
        const double qj = sQq[wave][b + jj];
        const int jslot = __float_as_int(rj.w);
        float dx = rix - rj.x, dy = riy - rj.y, dz = riz - rj.z;
        if (pairImg) {
            dx = min_image(dx, Lx, iLx);
            dy = min_image(dy, Ly, iLy);
            dz = min_image(dz, Lz, iLz);
        }
        // exact: the squares of 2^-20-grid differences fit in 53 bits
        const double r2 = fma((double)dx, (double)dx, fma((double)dy, (double)dy, (double)dz * (double)dz));
        const bool in = iReal && jj < nj && jslot > islot && r2 < rc2;
        npair += (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(in));
        float fjx = 0.f, fjy = 0.f, fjz = 0.f;
        if (in) {
            const PairEF64 pe = pair_ef64(true, r2, ip.x + aj.x, ip.y * aj.y, ip.z * qj, krf64, crf64);
            eE += pe.e;
            if (FORCES) {
                const float fs = (float)pe.fs;
                fix = fmaf(fs, dx, fix); fiy = fmaf(fs, dy, fiy); fiz = fmaf(fs, dz, fiz);
                fjx = -fs * dx; fjy = -fs * dy; fjz = -fs * dz;
            }
        }
        if (FORCES) {
            fjx = oct_sum(fjx); fjy = oct_sum(fjy); fjz = oct_sum(fjz);
            const float fk = ii == 0 ? fjx : (ii == 1 ? fjy : fjz);
            if (ii < 3) sJF[wave][nflush][jj * 3 + ii] = jj < nj ? fk : 0.f;
            if (ii == 0) sJS[wave][nflush][jj] = jj < nj ? jslot : islot - ii;
            if (++nflush == kJFlush) {
                flush_j(kJFlush);
                nflush = 0;
            }
        }
    };

    const int ne = __builtin_amdgcn_readfirstlane(P.gcount[D.groupOff + g]);
    const uint32_t* E = P.entries + D.entryOff + (long long)g * D.entryCap;
    const uint64_t below = lanemask_lt();
    const int k = lane >> 3;
    for (int e0 = 0; e0 < ne; e0 += 64) {
        const uint32_t ent = (e0 + lane < ne) ? E[e0 + lane] : 0u;
        const uint64_t keepE = __builtin_amdgcn_ballot_w64((ent >> (20 + t)) & 1u);
        if (keepE & (1ull << lane)) sSel[wave][__popcll(keepE & below)] = ent & 0xFFFFFu;
        const int nsel = __popcll(keepE);
        wave_sync();
        // the test round's j atoms: atom (lane & 7) of kept entry r0 + k; the next
        // round's loads are issued before this round's pair steps run
        auto load_round = [&](int r0, int& js, float4& pj, double4& pa) __attribute__((always_inline)) {
            const int J = (int)sSel[wave][r0 + k < nsel ? r0 + k : 0];
            js = D.slotOff + J * kClusterSize + (lane & 7);
            pj = P.spos[js];
            pa = P.sprm64[js];
        };
        int jsN;
        float4 pjN;
        double4 paN;
        if (nsel > 0) load_round(0, jsN, pjN, paN);
        for (int r0 = 0; r0 < nsel; r0 += 8) {
            const bool ek = r0 + k < nsel;
            const int js = jsN;
            const float4 pj = pjN;
            const double4 pa = paN;
            if (r0 + 8 < nsel) load_round(r0 + 8, jsN, pjN, paN);
            float ax = pj.x - r0x, ay = pj.y - r0y, az = pj.z - r0z;
            ax = fmaf(-Lx, rintf(ax * iLx), ax);
            ay = fmaf(-Ly, rintf(ay * iLy), ay);
            az = fmaf(-Lz, rintf(az * iLz), az);
            const float gx = fmaxf(fabsf(ax) - hx, 0.f), gy = fmaxf(fabsf(ay) - hy, 0.f), gz = fmaxf(fabsf(az) - hz, 0.f);
            // dummy slots sit at x >= 1e6 (k_gather_bbox)
            const bool keep = ek && pj.x < 1.0e5f && fmaf(gx, gx, fmaf(gy, gy, gz * gz)) < D.rcList2;
            const uint64_t km = __builtin_amdgcn_ballot_w64(keep);
            if (keep) {
                const int q = qn + __popcll(km & below);
                sQp[wave][q] = make_float4(ax, ay, az, __int_as_float(js));
                sQa[wave][q] = make_double2(pa.x, pa.y);
                sQq[wave][q] = pa.z;
            }
            qn += __popcll(km);
            wave_sync();
            int b = 0;
            for (; b + 8 <= qn; b += 8) step(b, 8);
            // move the < 8 atoms left over to the front (sources >= 8 > destinations)
            const int left = qn - b;
            if (b > 0 && left > 0) {
                float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
                double2 a = make_double2(0.0, 0.0);
                double qq = 0.0;
                if (lane < left) { p = sQp[wave][b + lane]; a = sQa[wave][b + lane]; qq = sQq[wave][b + lane]; }
                wave_sync();
                if (lane < left) { sQp[wave][lane] = p; sQa[wave][lane] = a; sQq[wave][lane] = qq; }
            }
            qn = left;
            wave_sync();  // the steps' queue reads come before the next test round's writes
        }
        wave_sync();  // sSel is rewritten by the next batch
    }
    if (qn > 0) step(0, qn);
    if (FORCES) {
        if (nflush > 0) flush_j(nflush);
        fix = stride8_sum(fix); fiy = stride8_sum(fiy); fiz = stride8_sum(fiz);
        if (jj == 0 && iReal) flush_force(P, c, fixMax, islot, fix, fiy, fiz);
    }
    eE = wave_sum_d(eE);
    if (lane == 0) {
        *eOut = eE;
        atomicAdd(&P.rt[c].npairs, (unsigned long long)npair);
    }
}

// ---------------------------------------------------------------- exceptions

// Add a partial force in FP64 to a slot: fixed point below fixMax, FP64 side sums beyond.
__device__ __forceinline__ void add_force64(const DevPtrs& P, int c, double fixMax, long long slot, double fx,
                                            double fy, double fz) {
    long long* f = P.sf + slot;
    double* d = P.sd + 3 * slot;
    bool big = false;
    if (fabs(fx) < fixMax) atomic_add_fixed(f, __double2ll_rn(fx * kForceScale)); else { atomicAdd(d, fx); big = true; }
    if (fabs(fy) < fixMax) atomic_add_fixed(f + P.sfn, __double2ll_rn(fy * kForceScale)); else { atomicAdd(d + 1, fy); big = true; }
    if (fabs(fz) < fixMax) atomic_add_fixed(f + 2 * P.sfn, __double2ll_rn(fz * kForceScale)); else { atomicAdd(d + 2, fz); big = true; }
    if (big) mark_side(P, c, slot);
}

// One exception term at the plain separation (dx, dy, dz) = a - b: its energy
// k_e q'q / r + 4 eps' ((s'/r)^12 - (s'/r)^6) (q = (k_e q'q, sigma, 4 eps', 0)),
// and fs, the force on a over the separation (F_a = fs (dx, dy, dz)).
__device__ __forceinline__ double exception_term(const double4& q, double dx, double dy, double dz, double& fs) {
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double rinv = 1.0 / sqrt(r2);
    const double r2inv = rinv * rinv;
    const double s2 = q.y * q.y * r2inv, s6 = s2 * s2 * s2;
    const double elj = q.z * fma(s6, s6, -s6);
    const double ec = q.x * rinv;
    fs = fma(q.z, fma(12.0 * s6, s6, -6.0 * s6), ec) * r2inv;
    return elj + ec;
}

// One block per set: the exception terms k_e q'q / r + 4 eps' ((s'/r)^12 - (s'/r)^6)
// of its system at the plain distance of the set's coordinates (x, or x + d for
// state 1; no cutoff, no periodic image), as NonbondedForce computes its
// exceptions by default (exceptions do not use periodic boundary conditions).
// FP64 throughout; energy summed per thread in a fixed order, then a fixed tree.
__global__ void __launch_bounds__(256) k_exceptions(DevPtrs P, int setBegin, int forces) {
    const int c = setBegin + (int)blockIdx.x * P.setStride;
    const SetDesc& D = P.desc[c];
    __shared__ double red[256];
    const int n = D.excCount;
    const float4* X = P.posq + D.atomOff;
    const float4* Xd = P.displ + D.atomOff;
    const bool shifted = D.state == 1;
    const double fixMax = D.fixMax;
    double e = 0.0;
    for (int k = threadIdx.x; k < n; k += 256) {
        const int2 pr = P.xpair[D.excOff + k];
        const double4 q = P.xprm[D.excOff + k];
        const float4 a = X[pr.x], b = X[pr.y];
        double ax = a.x, ay = a.y, az = a.z, bx = b.x, by = b.y, bz = b.z;
        if (shifted) {
            const float4 da = Xd[pr.x], db = Xd[pr.y];
            ax += (double)da.x; ay += (double)da.y; az += (double)da.z;
            bx += (double)db.x; by += (double)db.y; bz += (double)db.z;
        }
        const double dx = ax - bx, dy = ay - by, dz = az - bz;
        double fs;
        e += exception_term(q, dx, dy, dz, fs);
        if (forces) {
            const long long si = D.slotOff + P.slotOf[D.setAtomOff + pr.x];
            const long long sj = D.slotOff + P.slotOf[D.setAtomOff + pr.y];
            add_force64(P, c, fixMax, si, fs * dx, fs * dy, fs * dz);
            add_force64(P, c, fixMax, sj, -fs * dx, -fs * dy, -fs * dz);
        }
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) P.xenergy[c] = red[0];
}

// ---------------------------------------------------------------- state difference
//
// When few atoms are displaced, x + d is not binned: every pair with a displaced
// atom is evaluated at x and at x + d from the state-0 cells (the other atoms do
// not move), U(x + d) = U(x) - E_d(x) + E_d(x + d) and F(x + d) = F(x) + dF.

// Displaced atom a of a system at state st (0: x, 1: x + d), wrapped as k_bin does.
__device__ __forceinline__ float4 displaced_position(const DevPtrs& P, const SetDesc& D, int a, int st) {
    float4 p = P.posq[D.atomOff + a];
    if (st) {
        const float4 d = P.displ[D.atomOff + a];
        p.x += d.x; p.y += d.y; p.z += d.z;  // exact on the 2^-20 grid
    }
    if (D.tri) {
        wrap_tri(D, p);
    } else {
        p.x = wrap_coord(p.x, D.L[0], D.invL[0]);
        p.y = wrap_coord(p.y, D.L[1], D.invL[1]);
        p.z = wrap_coord(p.z, D.L[2], D.invL[2]);
    }
    return p;
}

__device__ __forceinline__ bool same_displacement(const DevPtrs& P, const SetDesc& D, int a, int b) {
    const float4 da = P.displ[D.atomOff + a], db = P.displ[D.atomOff + b];
    return da.x == db.x && da.y == db.y && da.z == db.z;
}

// One wave per (displaced atom, state) of system sysBegin + blockIdx.y: its
// in-cutoff pairs with the atoms that are not displaced (from the state-0 cells
// within the cutoff of its position at that state) and with the displaced atoms
// after it in the list whose displacement differs (pairs of atoms displaced alike
// keep their separation, so they are the same at both states and are left to
// the state-0 evaluation).  Exact FP64 r^2 < rc^2 and the FP64 pair function of
// the state-0 tiles, exceptions dropped as there.  The wave's energy goes to
// lenergy[2 (ligOff + k) + st], its pair count to rt[state-0 set].lpairs[st];
// with FORCES, -F (state 0) or +F (state 1) goes to dF, the state-1 set's
// fixed-point sums indexed by atom.  DUMP: the pairs (lower index first) are
// written to out instead (state dumpState, no energy or force).  BWD (no forces):
// the derivatives of the pair energies with respect to the stored parameters
// q' = q sqrt(k_e), h = sigma/2, e' = 2 sqrt(eps) of both atoms (the tile kernel's
// backward formulas), -dE (state 0) or +dE (state 1), into the x + d set's
// gradient accumulators indexed by atom.
template <bool FORCES, bool EXC, bool DUMP, bool BWD = false>
__global__ void __launch_bounds__(256) k_displaced(DevPtrs P, int sysBegin, int dumpState, int2* out, long long cap,
                                                   unsigned long long* ctr) {
    static_assert(!(BWD && (FORCES || DUMP)), "the backward mode computes no forces");
    const int s = sysBegin + blockIdx.y;
    const int c0 = 2 * s, c1 = 2 * s + 1;
    const SetDesc D = P.desc[c0];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int w = blockIdx.x * 4 + wave;
    const int nL = D.ligCount;
    // kDisplacedParts waves per (displaced atom, state): part q takes rounds q, q + parts, ...
    // of the candidates (part 0 also the displaced partners)
    int st, k, part;
    if (DUMP) { st = dumpState; k = w / kDisplacedParts; part = w % kDisplacedParts; if (k >= nL) return; }
    else {
        if (w >= 2 * nL * kDisplacedParts) return;
        part = w % kDisplacedParts;
        st = (w / kDisplacedParts) & 1;
        k = w / (2 * kDisplacedParts);
    }
    const int a = P.ligAtoms[D.ligOff + k];
    const float4 p = displaced_position(P, D, a, st);
    const double4 pa = P.prm64[D.atomOff + a];
    const unsigned long long ma = EXC ? P.xmask[D.atomOff + a] : 0ull;
    const double fixMax = P.desc[c1].fixMax;
    const long long dOff = P.desc[c1].slotOff;  // dF of atom b at dOff + b
    const double sgn = st ? 1.0 : -1.0;
    double eacc = 0.0, fax = 0.0, fay = 0.0, faz = 0.0;
    double gaq = 0.0, gah = 0.0, gae = 0.0;  // BWD: dE/d(q', h, e') of atom a
    uint32_t npair = 0;
    // one candidate: b at pb with FP64 parameters pbq; nb: slot or atom for dF
    auto pair = [&](int b, float4 pb, double4 pbq) __attribute__((always_inline)) {
        float dx = p.x - pb.x, dy = p.y - pb.y, dz = p.z - pb.z;
        if (D.tri) {
            tri_image_xy(D, dx, dy, dz);  // OpenMM's triclinic image (z below)
        } else {
            dx = min_image(dx, D.L[0], D.invL[0]);
            dy = min_image(dy, D.L[1], D.invL[1]);
        }
        dz = min_image(dz, D.L[2], D.invL[2]);
        const double r2 = fma((double)dx, (double)dx, fma((double)dy, (double)dy, (double)dz * (double)dz));
        bool in = r2 < D.rc2;
        if (EXC && in) in = !excluded_pair(P, D, a, b, ma);
        if (!in) return;
        ++npair;
        if (DUMP) {
            const unsigned long long at = atomicAdd(ctr, 1ull);
            if ((long long)at < cap) out[at] = make_int2(min(a, b), max(a, b));
            return;
        }
        const PairEF64 pe = pair_ef64(true, r2, pa.x + pbq.x, pa.y * pbq.y, pa.z * pbq.z, D.krf64, D.crf64);
        eacc += pe.e;
        if (BWD) {
            // phi = 1/r + k_rf r^2 - c_rf; with s = h_a + h_b, u = s / r^2, s2 = s u, s6 = s2^3:
            // dE/dh = e'_a e'_b (12 s6 - 6) s2^2 u, dE/de'_a = e'_b (s6^2 - s6), dE/dq'_a = q'_b phi
            const double r2inv = 1.0 / r2;
            const double phi = fma(D.krf64, r2, sqrt(r2inv)) - D.crf64;
            const double sg = pa.x + pbq.x, u = sg * r2inv, s2 = sg * u, s6 = s2 * s2 * s2;
            const double dh = pa.y * pbq.y * fma(12.0, s6, -6.0) * s2 * s2 * u;
            const double lj = fma(s6, s6, -s6);
            gaq += pbq.z * phi; gah += dh; gae += pbq.y * lj;
            atomicAdd(P.gq + dOff + b, sgn * pa.z * phi);
            atomicAdd(P.gsig + dOff + b, sgn * dh);
            atomicAdd(P.geps + dOff + b, sgn * pa.y * lj);
        }
        if (FORCES) {
            const double fx = pe.fs * dx, fy = pe.fs * dy, fz = pe.fs * dz;
            fax += fx; fay += fy; faz += fz;
            add_force64(P, c1, fixMax, dOff + b, -sgn * fx, -sgn * fy, -sgn * fz);
        }
    };
    // The state-0 clusters within reach.  One column per lane (columns within Rx + 1,
    // Ry + 1: one more than the list search needs, FP32 binning may place an atom one
    // column over); a column whose rectangle is farther than rc (+ 1e-3 nm) in x-y is
    // dropped, the others take the clusters holding the z-bins within sqrt(rc^2 - rho^2)
    // of the atom (+ one bin each side): one or two cluster ranges (two when the window
    // wraps; the second starts after the first).  The ranges are flattened by a wave
    // prefix sum; rounds of 64 clusters test their boxes against the atom (lane per
    // cluster, binary search over the lanes' range starts in LDS), and the clusters that
    // pass are taken 8 at a time, lane 8 k + m on atom m of the k-th.  An atom that is
    // displaced carries 1 in sprm64.w (it is paired at its own state, below).
    __shared__ int sEx[4][64], sB0[4][64], sL0[4][64], sB1[4][64], sKeep[4][64];
    const int cx = (int)fminf(fmaxf(p.x * D.invCe[0], 0.f), (float)(D.nx - 1));
    const int cy = (int)fminf(fmaxf(p.y * D.invCe[1], 0.f), (float)(D.ny - 1));
    const int rx = min(D.Rx + 1, D.nx), ry = min(D.Ry + 1, D.ny);
    const int nxs = min(2 * rx + 1, D.nx), nys = min(2 * ry + 1, D.ny);
    const float lim = D.rc + 1e-3f, lim2 = lim * lim;
    for (int c0 = 0; c0 < nxs * nys; c0 += 64) {
        int b0 = 0, l0 = 0, b1 = 0, l1 = 0;
        const int l = c0 + lane;
        if (l < nxs * nys) {
            const int ix = l / nys, iy = l - (l / nys) * nys;
            int ccx = nxs == D.nx ? ix : cx - rx + ix;
            ccx += ccx < 0 ? D.nx : (ccx >= D.nx ? -D.nx : 0);
            int ccy = nys == D.ny ? iy : cy - ry + iy;
            ccy += ccy < 0 ? D.ny : (ccy >= D.ny ? -D.ny : 0);
            float rho2;
            if (D.tri) {
                // triclinic (every column enumerated): the column's rectangle within lim of the atom
                // under some lattice image (1e-3 nm of binning slack), z-window of the full lim
                const bool nearCol = tri_rect_near(D, ((float)ccx + 0.5f) * D.ce[0] - p.x,
                                                   ((float)ccy + 0.5f) * D.ce[1] - p.y, 0.5f * D.ce[0] + 1e-3f,
                                                   0.5f * D.ce[1] + 1e-3f, 1, lim, lim2);
                rho2 = nearCol ? 0.f : 2.f * lim2;
            } else {
                const float ex = fmaxf(fabsf(min_image(p.x - ((float)ccx + 0.5f) * D.ce[0], D.L[0], D.invL[0])) -
                                           0.5f * D.ce[0], 0.f);
                const float ey = fmaxf(fabsf(min_image(p.y - ((float)ccy + 0.5f) * D.ce[1], D.L[1], D.invL[1])) -
                                           0.5f * D.ce[1], 0.f);
                rho2 = ex * ex + ey * ey;
            }
            if (rho2 <= lim2) {
                const int col = ccx * D.ny + ccy;
                const int cs = P.colStart[D.colOff + col];
                const int* ccs = P.cellStart + D.cellOff + (long long)col * D.nz;
                const int* ccn = P.cellCount + D.cellOff + (long long)col * D.nz;
                const float hz = sqrtf(lim2 - rho2);
                int za0 = (int)floorf((p.z - hz) * D.invZe) - 1, za1 = (int)floorf((p.z + hz) * D.invZe) + 1;
                int zb0 = 1, zb1 = 0;
                if (za1 - za0 + 1 >= D.nz) { za0 = 0; za1 = D.nz - 1; }
                else if (za0 < 0) { zb0 = za0 + D.nz; zb1 = D.nz - 1; za0 = 0; }
                else if (za1 >= D.nz) { zb0 = 0; zb1 = za1 - D.nz; za1 = D.nz - 1; }
                // slot ranges -> cluster ranges (clusters never straddle columns)
                const int s0 = cs + ccs[za0], e0 = cs + ccs[za1] + ccn[za1];
                b0 = s0 >> 3;
                l0 = e0 > s0 ? ((e0 + 7) >> 3) - b0 : 0;
                if (zb0 <= zb1) {
                    const int s1 = cs + ccs[zb0], e1 = cs + ccs[zb1] + ccn[zb1];
                    b1 = s1 >> 3;
                    l1 = e1 > s1 ? ((e1 + 7) >> 3) - b1 : 0;
                    // the two ranges are disjoint in slots but may share the cluster where
                    // they meet: that cluster stays with the range that comes first
                    if (l0 > 0 && l1 > 0) {
                        if (b1 >= b0) { const int e = b0 + l0; if (b1 < e) { l1 -= e - b1; b1 = e; } }
                        else { const int e = b1 + l1; if (b0 < e) { l0 -= e - b0; b0 = e; } }
                        l0 = max(l0, 0); l1 = max(l1, 0);
                    }
                }
            }
        }
        const int tot = l0 + l1;
        const int incl = wave_incl_scan(tot);
        const int T = __shfl(incl, 63, 64);
        sEx[wave][lane] = incl - tot;
        sB0[wave][lane] = b0; sL0[wave][lane] = l0; sB1[wave][lane] = b1;
        wave_sync();
        for (int t0 = 0; t0 < T; t0 += 64) {
            // this lane's cluster and its box test
            const int t = t0 + lane;
            bool keep = false;
            int cl = 0;
            if (t < T) {
                int lo = 0;
#pragma unroll
                for (int step = 32; step > 0; step >>= 1)
                    if (sEx[wave][lo + step] <= t) lo += step;
                const int off = t - sEx[wave][lo];
                const int l0u = sL0[wave][lo];
                cl = off < l0u ? sB0[wave][lo] + off : sB1[wave][lo] + (off - l0u);
                const float4 bc = P.bbc[D.clusterOff + cl], bh = P.bbh[D.clusterOff + cl];
                if (D.tri) {
                    keep = tri_boxes_near(D, bc.x - p.x, bc.y - p.y, bc.z - p.z, bh.x, bh.y, bh.z, lim, lim2);
                } else {
                    const float dx = fmaxf(fabsf(min_image(p.x - bc.x, D.L[0], D.invL[0])) - bh.x, 0.f);
                    const float dy = fmaxf(fabsf(min_image(p.y - bc.y, D.L[1], D.invL[1])) - bh.y, 0.f);
                    const float dz = fmaxf(fabsf(min_image(p.z - bc.z, D.L[2], D.invL[2])) - bh.z, 0.f);
                    keep = dx * dx + dy * dy + dz * dz <= lim2;
                }
            }
            const uint64_t km = __builtin_amdgcn_ballot_w64(keep);
            const int nk = __popcll(km);
            if (keep) sKeep[wave][__popcll(km & lanemask_lt())] = cl;
            wave_sync();
            for (int k0 = 0; k0 < nk; k0 += 8) {
                const int kk = k0 + (lane >> 3);
                if (kk < nk) {
                    const long long sl = D.slotOff + 8ll * sKeep[wave][kk] + (lane & 7);
                    const int bj = P.sidx[sl];
                    const float4 pj = P.spos[sl];
                    const double4 qj = P.sprm64[sl];
                    if (bj >= 0 && qj.w == 0.0) pair(bj, pj, qj);
                }
            }
            wave_sync();  // the next round overwrites the kept clusters
        }
        wave_sync();  // the next chunk's columns overwrite the LDS ranges
    }
    // displaced partners after a in the list, displaced differently
    for (int k0 = k + 1 + 64 * part; k0 < nL; k0 += 64 * kDisplacedParts) {
        const int kk = k0 + lane;
        if (kk < nL) {
            const int b = P.ligAtoms[D.ligOff + kk];
            if (!same_displacement(P, D, a, b)) pair(b, displaced_position(P, D, b, st), P.prm64[D.atomOff + b]);
        }
    }
    const unsigned long long np = (unsigned long long)wave_sum_d((double)npair);
    if (DUMP) return;
    eacc = wave_sum_d(eacc);
    if (FORCES) {
        fax = wave_sum_d(fax); fay = wave_sum_d(fay); faz = wave_sum_d(faz);
    }
    if (BWD) {
        gaq = wave_sum_d(gaq); gah = wave_sum_d(gah); gae = wave_sum_d(gae);
        if (lane == 0) {
            atomicAdd(P.gq + dOff + a, sgn * gaq);
            atomicAdd(P.gsig + dOff + a, sgn * gah);
            atomicAdd(P.geps + dOff + a, sgn * gae);
        }
    }
    if (lane == 0) {
        P.lenergy[(2ll * (D.ligOff + k) + st) * kDisplacedParts + part] = eacc;
        atomicAdd(&P.rt[c0].lpairs[st], np);
        if (FORCES) add_force64(P, c1, fixMax, dOff + a, sgn * fax, sgn * fay, sgn * faz);
    }
}

// One block per system: the exception terms with a displaced atom whose partner
// is displaced differently, at x + d minus at x: energy into xenergy[state-1 set],
// with forces their difference into dF.
__global__ void __launch_bounds__(256) k_displaced_exceptions(DevPtrs P, int sysBegin, int forces) {
    const int s = sysBegin + blockIdx.x;
    const int c1 = 2 * s + 1;
    const SetDesc& D = P.desc[2 * s];
    __shared__ double red[256];
    const float4* X = P.posq + D.atomOff;
    const float4* Xd = P.displ + D.atomOff;
    const double fixMax = P.desc[c1].fixMax;
    const long long dOff = P.desc[c1].slotOff;
    double e = 0.0;
    for (int k = threadIdx.x; k < D.lxCount; k += 256) {
        const int x = P.lxidx[D.lxOff + k];
        const int2 pr = P.xpair[x];
        const double4 q = P.xprm[x];
        const float4 a = X[pr.x], b = X[pr.y], da = Xd[pr.x], db = Xd[pr.y];
        const double dx0 = (double)a.x - (double)b.x, dy0 = (double)a.y - (double)b.y, dz0 = (double)a.z - (double)b.z;
        const double dx1 = ((double)a.x + (double)da.x) - ((double)b.x + (double)db.x);
        const double dy1 = ((double)a.y + (double)da.y) - ((double)b.y + (double)db.y);
        const double dz1 = ((double)a.z + (double)da.z) - ((double)b.z + (double)db.z);
        double fs0, fs1;
        const double e0 = exception_term(q, dx0, dy0, dz0, fs0);
        const double e1 = exception_term(q, dx1, dy1, dz1, fs1);
        e += e1 - e0;
        if (forces) {
            add_force64(P, c1, fixMax, dOff + pr.x, fs1 * dx1 - fs0 * dx0, fs1 * dy1 - fs0 * dy0, fs1 * dz1 - fs0 * dz0);
            add_force64(P, c1, fixMax, dOff + pr.y, fs0 * dx0 - fs1 * dx1, fs0 * dy0 - fs1 * dy1, fs0 * dz0 - fs1 * dz1);
        }
    }
    red[threadIdx.x] = e;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) P.xenergy[c1] = red[0];
}

// ---------------------------------------------------------------- ATM combine

// One block of 256 threads per system: fixed-order FP64 sums of the energy
// slots of both states (estride: pair_energy_stride), then the scalar ATM combine.
__device__ __forceinline__ double block_sum_256(double v, double* red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// delta: state difference (only the state-0 set was paired): U2 = U1 - E_d(x) +
// E_d(x + d) + the displaced exception terms' difference, offsets per state; the
// state-1 pair count is derived the same way.
__global__ void __launch_bounds__(256) k_energy(DevPtrs P, int sysBegin, int estride, int delta) {
    const int s = sysBegin + blockIdx.x;
    __shared__ double red[256];
    double U[2];
    if (delta) {
        const int c0 = 2 * s, c1 = 2 * s + 1;
        const SetDesc& D = P.desc[c0];
        const int ng = P.rt[c0].ngroups;
        double acc = 0.0;
        for (int g = threadIdx.x; g < ng; g += 256) acc += group_energy(P, D, g, estride);
        const double e0 = block_sum_256(acc, red) + (D.excCount > 0 ? P.xenergy[c0] : 0.0);
        double dd = 0.0;
        for (int k = threadIdx.x; k < D.ligCount; k += 256) {
            const double* e = P.lenergy + 2ll * kDisplacedParts * (D.ligOff + k);
            double e0 = e[0], e1 = e[kDisplacedParts];
            for (int q = 1; q < kDisplacedParts; ++q) { e0 += e[q]; e1 += e[kDisplacedParts + q]; }
            dd += e1 - e0;
        }
        const double de = block_sum_256(dd, red) + (D.lxCount > 0 ? P.xenergy[c1] : 0.0);
        U[0] = e0 + (P.eoff ? P.eoff[c0] : 0.0);
        U[1] = (e0 + de) + (P.eoff ? P.eoff[c1] : 0.0);
        if (threadIdx.x == 0)
            P.rt[c1].npairs = P.rt[c0].npairs - P.rt[c0].lpairs[0] + P.rt[c0].lpairs[1];
    }
    for (int st = 0; st < 2 && !delta; ++st) {
        const int c = 2 * s + st;
        const SetDesc& D = P.desc[c];
        const int ng = P.rt[c].ngroups;
        double acc = 0.0;
        for (int g = threadIdx.x; g < ng; g += 256) acc += group_energy(P, D, g, estride);
        red[threadIdx.x] = acc;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        U[st] = red[0] + (D.excCount > 0 ? P.xenergy[c] : 0.0) + (P.eoff ? P.eoff[c] : 0.0);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        SysResult r;
        atm_combine(P.params[s], U[0], U[1], r);
        P.res[s] = r;
    }
}

// Fixed-order energy of one set into P.setEnergy[c] (reference-structured path).
__global__ void __launch_bounds__(256) k_set_energy(DevPtrs P, int c, int estride, int withOffsets) {
    __shared__ double red[256];
    const SetDesc& D = P.desc[c];
    const int ng = P.rt[c].ngroups;
    double acc = 0.0;
    for (int g = threadIdx.x; g < ng; g += 256) acc += group_energy(P, D, g, estride);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        P.setEnergy[c] = red[0] + (D.excCount > 0 ? P.xenergy[c] : 0.0) + (P.eoff && withOffsets ? P.eoff[c] : 0.0);
}

// F_sys += w1 F1 + w2 F2 with (w1, w2) = (1-sp, sp) for direction > 0 and
// (sp, 1-sp) otherwise (ReferenceATMMetaForceKernels.cpp:100-110); with
// `clear`, F_sys = w1 F1 + w2 F2 (no read of the old force).  hostMode 2
// adds F1 alone (state-1-only evaluation).  In host mode 1
// sp and direction come from kernel arguments (the reference's HybridForce
// receives sp as a scalar argument, CommonATMMetaForceKernels.cpp:199-201).
// The FP64 side sums of a slot are read only where its mask bit is set.
// delta (state difference): F_sys (+)= F1 + w2 dF, dF read by atom from the
// state-1 set's sums (F2 = F1 + dF, w1 + w2 = 1).
// Block order: the (system, atom block) list, gx blocks per system, is cut
// into 8 contiguous eighths and launch block b takes item b / 8 of eighth
// b mod 8.  Workgroups go to the 8 XCDs round-robin by launch index, so each
// XCD's L2 gathers the F1 lines (slot order) of its own eighth of the systems
// instead of every XCD fetching lines of every system.
__global__ void k_blend(DevPtrs P, int sysBegin, int nsys, int gx, int hostMode, double spHost, double dirHost,
                        int clear, int delta) {
    const int per = (gx * nsys + 7) >> 3;
    const int item = ((int)blockIdx.x & 7) * per + ((int)blockIdx.x >> 3);
    if (item >= gx * nsys) return;
    const int sys = item / gx;
    const int s = sysBegin + sys;
    const SetDesc& D0 = P.desc[2 * s];
    const SetDesc& D1 = P.desc[2 * s + 1];
    const int i = (item - sys * gx) * blockDim.x + threadIdx.x;
    const int N = D0.natoms;
    if (i >= N) return;
    const double sp = hostMode ? spHost : P.res[s].sp;
    const double dir = hostMode ? dirHost : P.params[s].direction;
    const double w2 = dir > 0.0 ? sp : 1.0 - sp;
    const double w1 = dir > 0.0 ? 1.0 - sp : sp;
    const long long a = D0.slotOff + P.slotOf[D0.setAtomOff + i];
    double* F = P.outF + 3ll * D0.atomOff;
    const double inv = 1.0 / kForceScale;
    const long long* fa = P.sf + a;
    double f1x = (double)fa[0] * inv, f1y = (double)fa[P.sfn] * inv, f1z = (double)fa[2 * P.sfn] * inv;
    if (P.rt[2 * s].big && ((P.sdMask[a >> 6] >> (a & 63)) & 1ull)) {
        const double* d = P.sd + 3 * a;
        f1x += d[0]; f1y += d[1]; f1z += d[2];
    }
    const double ox = clear ? 0.0 : F[i], oy = clear ? 0.0 : F[N + i], oz = clear ? 0.0 : F[2 * N + i];
    if (hostMode == 2) {  // state 1 only: F += F1 (the state-2 set was not evaluated)
        F[i] = ox + f1x;
        F[N + i] = oy + f1y;
        F[2 * N + i] = oz + f1z;
        return;
    }
    if (delta) {
        const long long b = D1.slotOff + i;
        const long long* fb = P.sf + b;
        double gx = (double)fb[0] * inv, gy = (double)fb[P.sfn] * inv, gz = (double)fb[2 * P.sfn] * inv;
        if (P.rt[2 * s + 1].big && ((P.sdMask[b >> 6] >> (b & 63)) & 1ull)) {
            const double* d = P.sd + 3 * b;
            gx += d[0]; gy += d[1]; gz += d[2];
        }
        F[i] = ox + (f1x + w2 * gx);
        F[N + i] = oy + (f1y + w2 * gy);
        F[2 * N + i] = oz + (f1z + w2 * gz);
        return;
    }
    const long long b = D1.slotOff + P.slotOf[D1.setAtomOff + i];
    const long long* fb = P.sf + b;
    double f2x = (double)fb[0] * inv, f2y = (double)fb[P.sfn] * inv, f2z = (double)fb[2 * P.sfn] * inv;
    if (P.rt[2 * s + 1].big && ((P.sdMask[b >> 6] >> (b & 63)) & 1ull)) {
        const double* d = P.sd + 3 * b;
        f2x += d[0]; f2y += d[1]; f2z += d[2];
    }
    F[i] = ox + (w1 * f1x + w2 * f2x);
    F[N + i] = oy + (w1 * f1y + w2 * f2y);
    F[2 * N + i] = oz + (w1 * f1z + w2 * f2z);
}

__global__ void k_fill_i64(long long* p, long long n) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) p[k] = 0;
}

// ---------------------------------------------------------------- launchers

static inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

hipError_t launch_zero_counts(const DevPtrs& P, int setBegin, int nsets, int maxCells, double* f, long long nf,
                              hipStream_t st) {
    const long long per = std::max<long long>(maxCells, cdiv(nf, std::max(nsets, 1)));
    unsigned blocks = (unsigned)std::min<long long>(std::max<long long>(cdiv(per, 256), 1), 1024);
    k_zero_counts<<<dim3(blocks, nsets), 256, 0, st>>>(P, setBegin, f, nf);
    return hipGetLastError();
}

hipError_t launch_build(const DevPtrs& P, int setBegin, int nsets, int maxAtoms, int maxCells, int maxSlots,
                        hipStream_t st) {
    dim3 ga(cdiv(maxAtoms, 256), nsets);
    k_bin<<<ga, 256, 0, st>>>(P, setBegin);
    k_scan<<<nsets, 1024, 0, st>>>(P, setBegin);
    k_scatter_idx<<<ga, 256, 0, st>>>(P, setBegin);
    k_gather_bbox<<<dim3(cdiv(maxSlots, 256), nsets), 256, 0, st>>>(P, setBegin);
    return hipGetLastError();
}

hipError_t launch_pairlist(const DevPtrs& P, int setBegin, int nsets, int maxGroups, int listWaves, hipStream_t st) {
#define ATM_PAIRLIST(TRI)                                                                             \
    switch (listWaves) {                                                                              \
        case 1: k_pairlist<1, TRI><<<dim3(cdiv(maxGroups, 1), nsets), 64, 0, st>>>(P, setBegin); break;  \
        case 2: k_pairlist<2, TRI><<<dim3(cdiv(maxGroups, 2), nsets), 128, 0, st>>>(P, setBegin); break; \
        default: k_pairlist<4, TRI><<<dim3(cdiv(maxGroups, 4), nsets), 256, 0, st>>>(P, setBegin); break; \
    }
    if (P.anyTri) { ATM_PAIRLIST(true) } else { ATM_PAIRLIST(false) }
#undef ATM_PAIRLIST
    return hipGetLastError();
}

hipError_t launch_alist(const DevPtrs& P, int setBegin, int nsets, int maxGroups, hipStream_t st) {
    k_alist<<<dim3(maxGroups * kGroupClusters, nsets), 64, 0, st>>>(P, setBegin);
    return hipGetLastError();
}

hipError_t launch_pairs(const DevPtrs& P, int setBegin, int nsets, int maxGroups, int waves, bool forces, bool fp64,
                        int variant, int split, hipStream_t st) {
    variant = effective_pair_variant(P, variant);
    if (fp64 && variant == kPairKernelFiltered) {
#define ATM_PAIRSF(WV)                                                                                          \
    if (forces) k_pairs_f<WV, kFiltForces><<<dim3(cdiv(8ll * maxGroups, WV), nsets), 64 * WV, 0, st>>>(P, setBegin); \
    else k_pairs_f<WV, kFiltEnergy><<<dim3(cdiv(8ll * maxGroups, WV), nsets), 64 * WV, 0, st>>>(P, setBegin);
        switch (waves) {
            case 1: ATM_PAIRSF(1) break;
            case 2: ATM_PAIRSF(2) break;
            default: ATM_PAIRSF(4) break;
        }
#undef ATM_PAIRSF
        return hipGetLastError();
    }
    if (fp64 && variant == kPairKernelAtomLists) {
#define ATM_PAIRSA(WV)                                                                                      \
    if (forces) k_pairs_a<WV, true><<<dim3(cdiv(maxGroups, WV), nsets), 64 * WV, 0, st>>>(P, setBegin);    \
    else k_pairs_a<WV, false><<<dim3(cdiv(maxGroups, WV), nsets), 64 * WV, 0, st>>>(P, setBegin);
        switch (waves) {
            case 1: ATM_PAIRSA(1) break;
            case 2: ATM_PAIRSA(2) break;
            default: ATM_PAIRSA(4) break;
        }
#undef ATM_PAIRSA
        return hipGetLastError();
    }
    if (fp64 && variant != 0) {
#define ATM_PAIRSQ(WV, JB, NC, SD)                                                                                     \
    if (forces) k_pairs_q<WV, JB, NC, SD, true><<<dim3(cdiv(maxGroups, WV), nsets), 64 * WV, 0, st>>>(P, setBegin);  \
    else k_pairs_q<WV, JB, NC, SD, false><<<dim3(cdiv(maxGroups, WV), nsets), 64 * WV, 0, st>>>(P, setBegin);
#define ATM_PAIRSQ_W(JB, NC, SD)                 \
    switch (waves) {                             \
        case 1: ATM_PAIRSQ(1, JB, NC, SD) break; \
        case 2: ATM_PAIRSQ(2, JB, NC, SD) break; \
        default: ATM_PAIRSQ(4, JB, NC, SD) break; \
    }
        switch (variant) {
            case 1: ATM_PAIRSQ_W(4, 2, true) break;
            case 2: ATM_PAIRSQ_W(4, 2, false) break;
            case 3: ATM_PAIRSQ_W(4, 4, true) break;
            default: ATM_PAIRSQ_W(8, 2, true) break;
        }
#undef ATM_PAIRSQ_W
#undef ATM_PAIRSQ
        return hipGetLastError();
    }
    if (split < 1 || split > 8) return hipErrorInvalidValue;
    const DevPtrs& Pk = P;
    if (P.lptItems && P.lptMode != kOrderLaunch) k_order<<<nsets, 512, 0, st>>>(P, setBegin);
#define ATM_PAIRS_X(WV, F64, X)                                                                                   \
    if (forces) k_pairs<WV, true, F64, false, X><<<dim3(cdiv(maxGroups, WV) * split, nsets), 64 * WV, 0, st>>>(   \
        Pk, setBegin, nullptr, split);                                                                            \
    else k_pairs<WV, false, F64, false, X><<<dim3(cdiv(maxGroups, WV) * split, nsets), 64 * WV, 0, st>>>(        \
        Pk, setBegin, nullptr, split);
#define ATM_PAIRS(WV, F64)                                \
    if (P.xmask || P.anyTri) { ATM_PAIRS_X(WV, F64, true) } else { ATM_PAIRS_X(WV, F64, false) }
#define ATM_PAIRS_W(F64)               \
    switch (waves) {                   \
        case 1: ATM_PAIRS(1, F64) break; \
        case 2: ATM_PAIRS(2, F64) break; \
        case 8: ATM_PAIRS(8, F64) break; \
        default: ATM_PAIRS(4, F64) break; \
    }
    if (fp64) { ATM_PAIRS_W(true) } else { ATM_PAIRS_W(false) }
#undef ATM_PAIRS_W
#undef ATM_PAIRS
#undef ATM_PAIRS_X
    return hipGetLastError();
}

hipError_t launch_energy(const DevPtrs& P, int sysBegin, int nsys, int estride, hipStream_t st, bool delta) {
    k_energy<<<nsys, 256, 0, st>>>(P, sysBegin, estride, delta ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_set_energy(const DevPtrs& P, int c, int estride, bool withOffsets, hipStream_t st) {
    k_set_energy<<<1, 256, 0, st>>>(P, c, estride, withOffsets ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_backward_pairs(const DevPtrs& P, int setBegin, int nsets, int maxGroups, const double* wset,
                                 int split, hipStream_t st) {
    if (split < 1 || split > 8) return hipErrorInvalidValue;
    const DevPtrs& Pk = P;
    if (P.lptItems && P.lptMode != kOrderLaunch) k_order<<<nsets, 512, 0, st>>>(P, setBegin);
    if (P.xmask || P.anyTri) k_pairs<1, false, true, true, true><<<dim3(maxGroups * split, nsets), 64, 0, st>>>(Pk, setBegin, wset, split);
    else k_pairs<1, false, true, true, false><<<dim3(maxGroups * split, nsets), 64, 0, st>>>(Pk, setBegin, wset, split);
    return hipGetLastError();
}

hipError_t launch_exceptions(const DevPtrs& P, int setBegin, int nsets, bool forces, hipStream_t st) {
    k_exceptions<<<nsets, 256, 0, st>>>(P, setBegin, forces ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_blend(const DevPtrs& P, int sysBegin, int nsys, int maxAtoms, int hostMode, double sp, double dir,
                        bool clear, hipStream_t st, bool delta) {
    if (nsys <= 0 || maxAtoms <= 0) return hipSuccess;
    const int gx = (int)cdiv(maxAtoms, 256);
    const unsigned blocks = 8u * cdiv((long long)gx * nsys, 8);
    k_blend<<<blocks, 256, 0, st>>>(P, sysBegin, nsys, gx, hostMode, sp, dir, clear ? 1 : 0, delta ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_displaced(const DevPtrs& P, int sysBegin, int nsys, int maxLig, bool forces, hipStream_t st,
                            bool grads) {
    if (maxLig <= 0 || nsys <= 0) return hipSuccess;
    const dim3 grid(cdiv(2ll * maxLig * kDisplacedParts, 4), nsys);
    const bool exc = P.xmask != nullptr;
    if (grads) {
        if (exc) k_displaced<false, true, false, true><<<grid, 256, 0, st>>>(P, sysBegin, 0, nullptr, 0, nullptr);
        else k_displaced<false, false, false, true><<<grid, 256, 0, st>>>(P, sysBegin, 0, nullptr, 0, nullptr);
    } else if (forces) {
        if (exc) k_displaced<true, true, false><<<grid, 256, 0, st>>>(P, sysBegin, 0, nullptr, 0, nullptr);
        else k_displaced<true, false, false><<<grid, 256, 0, st>>>(P, sysBegin, 0, nullptr, 0, nullptr);
    } else {
        if (exc) k_displaced<false, true, false><<<grid, 256, 0, st>>>(P, sysBegin, 0, nullptr, 0, nullptr);
        else k_displaced<false, false, false><<<grid, 256, 0, st>>>(P, sysBegin, 0, nullptr, 0, nullptr);
    }
    return hipGetLastError();
}

hipError_t launch_displaced_dump(const DevPtrs& P, int sys, int state, int nLig, int2* out, long long cap,
                                 unsigned long long* ctr, hipStream_t st) {
    if (nLig <= 0) return hipSuccess;
    const dim3 grid(cdiv((long long)nLig * kDisplacedParts, 4), 1);
    if (P.xmask) k_displaced<false, true, true><<<grid, 256, 0, st>>>(P, sys, state, out, cap, ctr);
    else k_displaced<false, false, true><<<grid, 256, 0, st>>>(P, sys, state, out, cap, ctr);
    return hipGetLastError();
}

hipError_t launch_displaced_exceptions(const DevPtrs& P, int sysBegin, int nsys, bool forces, hipStream_t st) {
    if (nsys <= 0) return hipSuccess;
    k_displaced_exceptions<<<nsys, 256, 0, st>>>(P, sysBegin, forces ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_fill_i64(long long* p, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    unsigned blocks = (unsigned)std::min<long long>(cdiv(n, 256), 4096);
    k_fill_i64<<<blocks, 256, 0, st>>>(p, n);
    return hipGetLastError();
}

// Zero n8 8-byte words (a fill kernel: safe inside captured graphs).
hipError_t launch_fill_zero(void* p, long long n8, hipStream_t st) {
    return launch_fill_i64(static_cast<long long*>(p), n8, st);
}

hipError_t launch_pairs_dump(const DevPtrs& P, int c, int maxGroups, int2* out, long long cap,
                             unsigned long long* counters, hipStream_t st) {
    k_pairs_dump<<<maxGroups, 64, 0, st>>>(P, c, out, cap, counters);
    return hipGetLastError();
}

}  // namespace atm
