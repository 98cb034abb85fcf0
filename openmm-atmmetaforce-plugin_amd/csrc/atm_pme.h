// Device-side layout of the PME variable-group evaluation and of the bonded
// terms (kernels in atm_pme.hip, host side in atm_pme_capi.cpp).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "atm_internal.h"

namespace atm {

constexpr int kPmeCellMax = 1024;  // atoms per cell the cell sort holds
constexpr int kPmeJBatch = 64;     // neighbour-cell atoms staged in LDS per batch (direct kernel)
constexpr int kPmeJSplit = 4;      // parts of a neighbour cell's atoms taken by separate waves
constexpr int kPmeSpreadSplit = 4; // parts of a cell's atoms spread by separate blocks (own LDS tiles)
constexpr int kPmeHalfShell = 14;  // the cell itself and its 13 forward neighbour cells
constexpr int kPmeBinLds = 4096;   // cell grids up to this size are binned through LDS histograms
constexpr int kPmeDftLdsMax = 600; // grid lines up to this length are transformed from LDS
constexpr int kPmeTileLdsMax = 49152;  // bytes of LDS for one cell's charge-grid tile

// Scalars of one evaluation (by value in the kernel arguments).
struct PmeDesc {
    int n;                       // atoms
    int m[3], ncell;             // cell grid (edge >= rc, at least 3 per axis)
    int cellCap, chunks;         // atoms per cell the direct kernel covers (chunks of 64)
    int K[3], ngrid;             // PME grid
    int twOff[3];                // offset of each axis' twiddle table
    int nex;                     // exception pairs
    int nslot, nslotDirect, nslotEx;  // energy slots per state (direct | exceptions | reciprocal)
    int tile[3], tileBytes;      // per-cell charge-grid tile (0 bytes: spread without tiles)
    double L[3], invL[3];
    double rc2, alpha, twoAlphaSqrtPi, pi2OverAlpha2;
    double recipEnergyScale;     // k_e / (2 pi V)
    double recipPhiScale;        // k_e / (pi V)
    double selfEnergy;           // -k_e alpha / sqrt(pi) sum q^2
    double dispersionEnergy;     // long-range LJ correction / V
};

struct PmePtrs {
    const double* x;        // [3n] positions (state 0)
    double* x2;             // [3n] x + d (state 1)
    const double* displ;    // [3n]
    int* cellOf;            // [n]
    int* rankOf;            // [n]
    int* cellCount;         // [ncell]
    int* cellStart;         // [ncell + 1]
    int* cellAtoms;         // [n]
    double4* wpos;          // [n] wrapped positions
    const double2* prm;     // [n] (q sqrt(k_e), q)
    const double2* prm2;    // [n] (sigma/2, 2 sqrt(eps))
    const int4* exkey;      // [n] (exclusion group, rank in group, bitmap row offset, 0)
    double4* cpos;          // [n] cell order: wrapped x, y, z, q sqrt(k_e)
    double2* cprm;          // [n] cell order: sigma/2, 2 sqrt(eps)
    int4* cex;              // [n] cell order: group, rank, row offset, atom
    const uint32_t* exbits; // exclusion bitmaps
    const int2* expairs;    // [nex]
    const double4* exprm;   // [nex] (k_e qq', sigma', 4 eps', k_e q_i q_j)
    long long* qgrid;       // [ngrid] fixed-point charge grid
    double2* g0;            // [ngrid]
    double2* g1;            // [ngrid]
    const double2* tw;      // twiddles of the three axes
    const double* bmod;     // B-spline moduli of the three axes, concatenated
    unsigned long long* F;  // [2][3][n] fixed-point forces per state (SoA)
    double* eslot;          // [2][nslot]
    double* parts;          // [2][5]
    SysResult* result;
    const AtmParamsDev* params;  // the nine ATM parameters (uploaded before each evaluation)
    const double* eoff;          // [2] energies of other variable-group forces added to U1, U2 (uploaded with params)
    int* overflow;
};

struct BondedPtrs {
    const double* x;          // [3n]
    unsigned long long* F;    // [3n] fixed point (AoS)
    const int2* bonds;
    const double2* bprm;      // r0, k
    const int4* angles;       // i, j (vertex), k, -
    const double2* aprm;      // theta0, k
    const int4* torsions;
    const double4* tprm;      // periodicity, phase, k, -
    int nb, na, nt;
};

hipError_t launch_pme_fill(unsigned long long* p, long long n, hipStream_t st);
hipError_t launch_pme_state(const PmeDesc& D, const PmePtrs& P, int state, hipStream_t st);
hipError_t launch_pme_finish(const PmeDesc& D, const PmePtrs& P, bool single, double* forcesOut, hipStream_t st);
hipError_t launch_bonded(const BondedPtrs& B, double* eslot, int nblocks, hipStream_t st);

}  // namespace atm
