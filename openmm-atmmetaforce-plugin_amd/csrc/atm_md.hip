// gfx950 kernels of the batched Langevin dynamics: every system of a context
// integrated together on the ATM force W of its current lambda-state.
//
// Splitting (the LangevinMiddle / BAOAB scheme of OpenMM's
// LangevinMiddleIntegrator), per step with F = F(x_n) from the previous
// evaluation:
//   v += dt F / m;  x += dt/2 v;  v = a v + sqrt(kT (1 - a^2) / m) R;  x += dt/2 v
// with a = exp(-friction dt) and R ~ N(0, 1) per component; the forces at the
// new positions are then evaluated for the next step.  Positions and
// velocities are FP64; the FP32 copy the pair kernels read is rewritten from
// them every step.  Normal deviates come from Philox4x32-10 (Salmon, Moraes,
// Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3", SC11) keyed by
// the seed and counted by (atom, step), so a trajectory is reproducible and a
// step can be replayed bit for bit.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "atm_kernels.h"

namespace atm {

namespace {

__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
    // Philox round constants (published with the generator)
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)M0 * c[0], p1 = (uint64_t)M1 * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += W0; k1 += W1;
    }
}

// Three N(0, 1) deviates for counter (atom, step) under `seed` (Box-Muller on
// uniforms in (0, 1)).
__device__ __forceinline__ void normals3(uint64_t seed, uint32_t atom, uint64_t step, double n[3]) {
    uint32_t c[4] = {atom, (uint32_t)step, (uint32_t)(step >> 32), 0x41544dU};
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double s = 2.3283064365386963e-10;  // 2^-32
    double u[4];
    for (int k = 0; k < 4; ++k) u[k] = ((double)c[k] + 0.5) * s;
    const double r0 = sqrt(-2.0 * log(u[0])), r1 = sqrt(-2.0 * log(u[2]));
    const double twopi = 6.283185307179586;
    n[0] = r0 * cos(twopi * u[1]);
    n[1] = r0 * sin(twopi * u[1]);
    n[2] = r1 * cos(twopi * u[3]);
}

}  // namespace

// One thread per atom of system sysBegin + blockIdx.y.
__global__ void k_md_step(DevPtrs P, MDPtrs M, int sysBegin, double dt, double a, double kT, uint64_t seed,
                          uint64_t step) {
    const int s = sysBegin + blockIdx.y;
    const SetDesc& D = P.desc[2 * s];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int N = D.natoms;
    if (i >= N) return;
    const long long g = (long long)D.atomOff + i;
    const double im = M.invm[g];
    if (im == 0.0) return;  // massless particles do not move
    const double* F = P.outF + 3ll * D.atomOff;
    double x[3], v[3], n[3];
    normals3(seed, (uint32_t)g, step, n);
    const double sd = sqrt(kT * (1.0 - a * a) * im);
    for (int k = 0; k < 3; ++k) {
        x[k] = M.x[3 * g + k];
        v[k] = M.v[3 * g + k] + dt * F[(long long)k * N + i] * im;
        x[k] += 0.5 * dt * v[k];
        v[k] = a * v[k] + sd * n[k];
        x[k] += 0.5 * dt * v[k];
        M.x[3 * g + k] = x[k];
        M.v[3 * g + k] = v[k];
    }
    float4 p = M.posq[g];
    p.x = (float)x[0]; p.y = (float)x[1]; p.z = (float)x[2];
    M.posq[g] = p;
}

// Maxwell-Boltzmann velocities at kT (counter step = 2^64 - 1, never a step).
__global__ void k_md_init(DevPtrs P, MDPtrs M, long long total, double kT, uint64_t seed) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    double n[3];
    normals3(seed, (uint32_t)g, ~0ull, n);
    const double sd = sqrt(kT * M.invm[g]);
    for (int k = 0; k < 3; ++k) M.v[3 * g + k] = sd * n[k];
    float4 p = M.posq[g];
    p.x = (float)M.x[3 * g]; p.y = (float)M.x[3 * g + 1]; p.z = (float)M.x[3 * g + 2];
    M.posq[g] = p;
}

// Rewrite the FP32 positions from the FP64 ones (after restoring a snapshot).
__global__ void k_md_posq(DevPtrs P, MDPtrs M, long long total) {
    const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    float4 p = M.posq[g];
    p.x = (float)M.x[3 * g]; p.y = (float)M.x[3 * g + 1]; p.z = (float)M.x[3 * g + 2];
    M.posq[g] = p;
}

hipError_t launch_md_posq(const DevPtrs& P, const MDPtrs& M, long long total, hipStream_t st) {
    k_md_posq<<<(int)((total + 255) / 256), 256, 0, st>>>(P, M, total);
    return hipGetLastError();
}

hipError_t launch_md_step(const DevPtrs& P, const MDPtrs& M, int nsys, int maxAtoms, double dt, double a, double kT,
                          uint64_t seed, uint64_t step, hipStream_t st) {
    k_md_step<<<dim3((maxAtoms + 255) / 256, nsys), 256, 0, st>>>(P, M, 0, dt, a, kT, seed, step);
    return hipGetLastError();
}

hipError_t launch_md_init(const DevPtrs& P, const MDPtrs& M, long long total, double kT, uint64_t seed, hipStream_t st) {
    k_md_init<<<(int)((total + 255) / 256), 256, 0, st>>>(P, M, total, kT, seed);
    return hipGetLastError();
}

}  // namespace atm
