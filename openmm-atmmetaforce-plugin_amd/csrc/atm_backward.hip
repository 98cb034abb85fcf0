// Backward pass: dW/d(charge, sigma, epsilon) per atom.
//
// W depends on the per-atom parameters only through U1 and U2, with
// dW/dU_state given per set (w[c] = gW (1 - sp) for the state that supplies e0,
// gW sp for the other).  The pair terms' derivatives with respect to the
// stored parameters (q' = q sqrt(k_e), h = sigma/2, e' = 2 sqrt(eps)) are
// recomputed on the forward's lists by the cluster-tile pair kernel in its
// backward mode (launch_backward_pairs, atm_kernels.hip) and summed per
// slot; nothing per pair is stored between the passes.  The gather kernel here
// sums both states per atom and applies the chain rule to physical units:
// dq = sqrt(k_e) dE/dq', dsigma = dE/dh / 2, deps = dE/de' * 2 / e' (0 for an
// atom with eps = 0, where the derivative does not exist).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "atm_internal.h"
#include "atm_kernels.h"

namespace atm {

__global__ void k_fill_f64(double* p, long long n) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        p[k] = 0.0;
}

// wset null: the accumulators already hold weighted sums; otherwise they hold
// dU/dp of each set and wset[2 s], wset[2 s + 1] weight the two states here.
// delta (x + d evaluated as the difference from x, wset given): the x + d set's
// accumulators hold dU2/dp - dU1/dp by atom, so the weights are wa + wb on the x
// set's and wb on the difference.
__global__ void k_backward_gather(DevPtrs P, int sysBegin, double* dq, double* dsig, double* deps,
                                  const double* wset, int delta) {
    const int s = sysBegin + blockIdx.y;
    const SetDesc& D0 = P.desc[2 * s];
    const SetDesc& D1 = P.desc[2 * s + 1];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D0.natoms) return;
    const int a = D0.slotOff + P.slotOf[D0.setAtomOff + i];
    const int b = delta ? D1.slotOff + i : D1.slotOff + P.slotOf[D1.setAtomOff + i];
    const int k = D0.atomOff + i;
    double wa = wset ? wset[2 * s] : 1.0;
    const double wb = wset ? wset[2 * s + 1] : 1.0;
    if (delta) wa += wb;
    const double sqk = 11.787089759563214;  // sqrt(138.935456)
    const double ep = P.prm64[k].y;          // 2 sqrt(eps)
    if (dq) dq[k] = sqk * fma(wa, P.gq[a], wb * P.gq[b]);
    if (dsig) dsig[k] = 0.5 * fma(wa, P.gsig[a], wb * P.gsig[b]);
    // dE/deps = dE/de' de'/deps with de'/deps = 1/sqrt(eps) = 2/e': no derivative at eps = 0,
    // where 0 is returned (the convention of the FP64 oracle)
    if (deps) deps[k] = ep > 0.0 ? fma(wa, P.geps[a], wb * P.geps[b]) * 2.0 / ep : 0.0;
}

__global__ void k_set_positions(float4* posq, const float* xyz, long long n) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = posq[i];
    p.x = xyz[3 * i]; p.y = xyz[3 * i + 1]; p.z = xyz[3 * i + 2];
    posq[i] = p;
}

// Per-atom parameters of every atom of a context from device doubles, converted
// as atm_set_particles converts host ones: posq.w = q sqrt(k_e), ljp = (sigma/2,
// 2 sqrt(eps)) in FP32, prm64 = (sigma/2, 2 sqrt(eps), q sqrt(k_e), w) in FP64 (w, 1 for a
// displaced atom, is kept).
// k_check_particles counts the entries atm_set_particles would reject.
__global__ void k_check_particles(const double* q, const double* sig, const double* eps, long long n, int* bad) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double qi = q[i], si = sig[i], ei = eps[i];
    if (!(isfinite(qi) && isfinite(si) && isfinite(ei) && si >= 0.0 && ei >= 0.0)) atomicAdd(bad, 1);
}

__global__ void k_set_particles(float4* posq, float2* ljp, double4* prm64, const double* q, const double* sig,
                                const double* eps, long long n, double sqk) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double qi = q[i], si = sig[i], ei = eps[i];
    posq[i].w = (float)(qi * sqk);
    ljp[i] = make_float2((float)(0.5 * si), (float)(2.0 * sqrt(ei)));
    prm64[i] = make_double4(0.5 * si, 2.0 * sqrt(ei), qi * sqk, prm64[i].w);  // w: the displaced-atom mark
}

__global__ void k_set_marks(double4* prm64, const unsigned char* mark, long long n) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) prm64[i].w = (double)mark[i];
}

static inline unsigned cdivb(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

hipError_t launch_set_marks(double4* prm64, const unsigned char* mark, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_set_marks<<<cdivb(n, 256), 256, 0, st>>>(prm64, mark, n);
    return hipGetLastError();
}

hipError_t launch_check_particles(const double* q, const double* sig, const double* eps, long long n, int* bad,
                                  hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_check_particles<<<cdivb(n, 256), 256, 0, st>>>(q, sig, eps, n, bad);
    return hipGetLastError();
}

hipError_t launch_set_particles(const DevPtrs& P, const double* q, const double* sig, const double* eps, long long n,
                                double sqk, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_set_particles<<<cdivb(n, 256), 256, 0, st>>>(const_cast<float4*>(P.posq), const_cast<float2*>(P.ljp),
                                                   const_cast<double4*>(P.prm64), q, sig, eps, n, sqk);
    return hipGetLastError();
}

hipError_t launch_backward(const DevPtrs& P, long long totalSlots, int setBegin, int nsets, int maxGroups,
                           const double* wset, int split, hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<long long>(cdivb(totalSlots, 256), 4096);
    k_fill_f64<<<blocks, 256, 0, st>>>(P.gq, totalSlots);
    k_fill_f64<<<blocks, 256, 0, st>>>(P.gsig, totalSlots);
    k_fill_f64<<<blocks, 256, 0, st>>>(P.geps, totalSlots);
    return launch_backward_pairs(P, setBegin, nsets, maxGroups, wset, split, st);
}

hipError_t launch_backward_gather(const DevPtrs& P, int sysBegin, int nsys, int maxAtoms, double* dq, double* dsig,
                                  double* deps, const double* wset, hipStream_t st, bool delta) {
    k_backward_gather<<<dim3(cdivb(maxAtoms, 256), nsys), 256, 0, st>>>(P, sysBegin, dq, dsig, deps, wset,
                                                                      delta ? 1 : 0);
    return hipGetLastError();
}

hipError_t launch_set_positions(DevPtrs P, float4* posq, const float* xyz, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_set_positions<<<cdivb(n, 256), 256, 0, st>>>(posq, xyz, n);
    return hipGetLastError();
}

}  // namespace atm
