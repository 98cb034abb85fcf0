// gfx950 kernel of the flat-bottom restraints evaluated beside the ATM force.
//
// One block per centroid restraint: FP64 tree sums of the weighted member
// positions of both groups, the flat-bottom energy of the centroid distance,
// and the force on every member (its weight share of -dE/dc).  Position
// restraints: one thread each.  Forces are summed in 64-bit fixed point
// (2^-24 kJ/mol/nm) with integer atomics, so an atom shared by several
// restraints gets the same sum whatever the order; energies go to one slot
// per restraint and are summed on the host in a fixed order.
//
// The formulas restate the energy expressions that
// python/ATMMetaForceUtils.py builds as OpenMM custom forces:
//   centroid (:113-114): (kfcm/2) step(d12 - tolcm) (d12 - tolcm)^2,
//                        d12 = |c1 - offset - c2|;
//   position (:733-736): 0.5 fc select(step(dist - tol), (dist - tol)^2, 0),
//                        dist = periodicdistance(x, x0) or |x - x0|.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "atm_kernels.h"

namespace atm {

namespace {

constexpr double kScale = 16777216.0;  // 2^24 fixed-point units per kJ/mol/nm

__device__ __forceinline__ void add_fixed(unsigned long long* p, double f) {
    atomicAdd(p, (unsigned long long)__double2ll_rn(f * kScale));
}

// Block-wide FP64 sum (fixed order) of four values; result valid in thread 0.
template <int NT>
__device__ void block_sum4(double v[4], double (*red)[NT]) {
    const int t = threadIdx.x;
    for (int k = 0; k < 4; ++k) red[k][t] = v[k];
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1) {
        if (t < o)
            for (int k = 0; k < 4; ++k) red[k][t] += red[k][t + o];
        __syncthreads();
    }
    for (int k = 0; k < 4; ++k) v[k] = red[k][0];
    __syncthreads();
}

}  // namespace

__global__ void __launch_bounds__(256) k_restraints(RestraintPtrs R) {
    constexpr int NT = 256;
    __shared__ double red[4][NT];
    __shared__ double sh[8];
    const int b = blockIdx.x, t = threadIdx.x;
    if (b < R.ncent) {
        const atm_centroid_restraint C = R.cent[b];
        double c[2][4];
        const int first[2] = {C.group1_first, C.group2_first}, count[2] = {C.group1_count, C.group2_count};
        for (int g = 0; g < 2; ++g) {
            double v[4] = {0.0, 0.0, 0.0, 0.0};
            for (int m = t; m < count[g]; m += NT) {
                const int a = R.members[first[g] + m];
                const double w = R.weights[first[g] + m];
                v[0] += w * R.pos[3 * a];
                v[1] += w * R.pos[3 * a + 1];
                v[2] += w * R.pos[3 * a + 2];
                v[3] += w;
            }
            block_sum4<NT>(v, red);
            for (int k = 0; k < 4; ++k) c[g][k] = v[k];
        }
        if (t == 0) {
            const double dx = c[0][0] / c[0][3] - C.offset[0] - c[1][0] / c[1][3];
            const double dy = c[0][1] / c[0][3] - C.offset[1] - c[1][1] / c[1][3];
            const double dz = c[0][2] / c[0][3] - C.offset[2] - c[1][2] / c[1][3];
            const double d = sqrt(dx * dx + dy * dy + dz * dz);
            double e = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;  // g = dE/dc1
            if (d > C.tol) {
                const double s = d - C.tol;
                e = 0.5 * C.kf * s * s;
                const double k = C.kf * s / d;
                gx = k * dx; gy = k * dy; gz = k * dz;
            }
            R.energy[b] = e;
            sh[0] = gx; sh[1] = gy; sh[2] = gz;
            sh[3] = c[0][3]; sh[4] = c[1][3];
        }
        __syncthreads();
        const double gx = sh[0], gy = sh[1], gz = sh[2];
        for (int g = 0; g < 2; ++g) {
            const double sgn = g == 0 ? -1.0 : 1.0, W = sh[3 + g];
            for (int m = t; m < count[g]; m += NT) {
                const int a = R.members[first[g] + m];
                const double f = sgn * R.weights[first[g] + m] / W;
                if (gx != 0.0 || gy != 0.0 || gz != 0.0) {
                    add_fixed(R.fsum + 3 * a, f * gx);
                    add_fixed(R.fsum + 3 * a + 1, f * gy);
                    add_fixed(R.fsum + 3 * a + 2, f * gz);
                }
            }
        }
        return;
    }
    const int r = (b - R.ncent) * NT + t;
    if (r >= R.npos) return;
    const atm_position_restraint Q = R.posr[r];
    const int a = Q.atom;
    double d3[3];
    for (int k = 0; k < 3; ++k) {
        double x = R.pos[3 * a + k] - Q.ref[k];
        if (Q.periodic) x -= R.box[k] * rint(x / R.box[k]);
        d3[k] = x;
    }
    const double d = sqrt(d3[0] * d3[0] + d3[1] * d3[1] + d3[2] * d3[2]);
    double e = 0.0;
    if (d > Q.tol) {
        const double s = d - Q.tol;
        e = 0.5 * Q.kf * s * s;
        const double k = -Q.kf * s / d;
        for (int q = 0; q < 3; ++q) add_fixed(R.fsum + 3 * a + q, k * d3[q]);
    }
    R.energy[R.ncent + r] = e;
}

hipError_t launch_restraints(const RestraintPtrs& R, hipStream_t st) {
    const int blocks = R.ncent + (R.npos + 255) / 256;
    if (blocks == 0) return hipSuccess;
    k_restraints<<<blocks, 256, 0, st>>>(R);
    return hipGetLastError();
}

}  // namespace atm
