// Backward pass: dW/d(charge, sigma, epsilon) per atom.
//
// W depends on the per-atom parameters only through U1 and U2, with
// dW/dU_state given per set (w[c] = gW (1 - sp) for the state that supplies e0,
// gW sp for the other).  The kernel walks the same cluster-pair lists as the
// forward pair kernel, recomputes every in-cutoff pair term and accumulates,
// in slot order, the derivatives with respect to the stored parameters
//   q' = q sqrt(k_e):        dE/dq'_i  = q'_j (1/r + k_rf r^2 - c_rf)
//   h  = sigma/2:            dE/dh_i   = e'_i e'_j (12 (s/r)^12 - 6 (s/r)^6) / s
//   e' = 2 sqrt(eps):        dE/de'_i  = e'_j ((s/r)^12 - (s/r)^6)
// (s = h_i + h_j; dE/dh = 0 where s = 0).  The gather kernel sums both states
// and applies the chain rule to physical units: dq = sqrt(k_e) dE/dq',
// dsigma = dE/dh / 2, deps = dE/de' * 2 / e' (0 for an atom with eps = 0, where
// the derivative does not exist).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "atm_internal.h"
#include "atm_kernels.h"

namespace atm {

namespace {

__device__ __forceinline__ float bmin_image(float d, float L, float invL) { return fmaf(-L, rintf(d * invL), d); }

__device__ __forceinline__ bool b_in_cutoff(const SetDesc& D, float dx, float dy, float dz, float r2) {
    bool in = r2 < D.rc2lo;
    if (r2 >= D.rc2lo && r2 < D.rc2hi) {
        double r2d = (double)dx * dx + (double)dy * dy + (double)dz * dz;
        in = r2d < D.rc2;
    }
    return in;
}

__device__ __forceinline__ float oct_sum_b(float v) {
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0xB1, 0xF, 0xF, true));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x4E, 0xF, 0xF, true));
    v += __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x141, 0xF, 0xF, true));
    return v;
}

__device__ __forceinline__ float stride8_sum_b(float v) {
    v += __shfl_xor(v, 8, 64);
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
}

}  // namespace

__global__ void __launch_bounds__(64) k_backward(DevPtrs P, int setBegin, const double* wset) {
    const int c = setBegin + blockIdx.y;
    const SetDesc& D = P.desc[c];
    const int lane = threadIdx.x & 63, jj = lane >> 3, ii = lane & 7;
    const int g = blockIdx.x;
    if (g >= P.rt[c].ngroups) return;
    const double w = wset[c];
    const int4 gi = P.ginfo[D.groupOff + g];
    const int g0 = gi.x, ncl = gi.y;
    float gi_q[kGroupClusters], gi_h[kGroupClusters], gi_e[kGroupClusters];
#pragma unroll
    for (int t = 0; t < kGroupClusters; ++t) gi_q[t] = gi_h[t] = gi_e[t] = 0.f;
    const int ne = P.gcount[D.groupOff + g];
    const uint32_t* E = P.entries + D.entryOff + (long long)g * D.entryCap;
    for (int e = 0; e < ne; ++e) {
        const uint32_t ent = E[e];
        const int J = ent & 0xFFFFF;
        const uint32_t imask = (ent >> 20) & 0xFFu;
        const int js = D.slotOff + J * kClusterSize + jj;
        float4 pj = P.spos[js];
        const float2 qj = P.sprm[js];
        const bool jv = P.sidx[js] >= 0;
        if (D.shiftOK) {
            const uint32_t sxc = (ent >> 28) & 3u, syc = (ent >> 30) & 3u;
            pj.x += sxc == 1u ? D.L[0] : (sxc == 2u ? -D.L[0] : 0.f);
            pj.y += syc == 1u ? D.L[1] : (syc == 2u ? -D.L[1] : 0.f);
        }
        float gj_q = 0.f, gj_h = 0.f, gj_e = 0.f;
#pragma unroll
        for (int t = 0; t < kGroupClusters; ++t) {
            if (imask & (1u << t)) {
                const int is = D.slotOff + (g0 + t) * kClusterSize + ii;
                const float4 pi = P.spos[is];
                const float2 qi = P.sprm[is];
                const bool iv = P.sidx[is] >= 0;
                float dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
                if (!D.shiftOK) { dx = bmin_image(dx, D.L[0], D.invL[0]); dy = bmin_image(dy, D.L[1], D.invL[1]); }
                dz = bmin_image(dz, D.L[2], D.invL[2]);
                const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                bool in = b_in_cutoff(D, dx, dy, dz, r2) && jv && iv;
                if (g0 + t == J) in = in && (ii < jj);
                float rinv = __builtin_amdgcn_rsqf(r2);
                {  // Newton step on the residual (unbiased)
                    const float tt = r2 * rinv;
                    const float ee = fmaf(-tt, rinv, 1.0f);
                    rinv = fmaf(0.5f * rinv, ee, rinv);
                }
                const float r2inv = rinv * rinv;
                const float sig = qi.x + qj.x;
                const float s2 = sig * sig * r2inv;
                const float s6 = s2 * s2 * s2;
                const float phi = fmaf(D.krf, r2, rinv) - D.crf;   // RF potential factor
                const float lj = fmaf(s6, s6, -s6);                // s12 - s6
                const float dlj = s6 * fmaf(12.0f, s6, -6.0f);     // 12 s12 - 6 s6
                const float eps4 = qi.y * qj.y;
                // dE/dh = 4 eps (12 s^11/r^12 - 6 s^5/r^6): 0 at s = 0, where the quotient is 0/0
                const float dh = (in && sig > 0.f) ? eps4 * dlj / sig : 0.f;
                const float dqi = in ? pj.w * phi : 0.f, dqj = in ? pi.w * phi : 0.f;
                const float dei = in ? qj.y * lj : 0.f, dej = in ? qi.y * lj : 0.f;
                gi_q[t] += dqi; gi_h[t] += dh; gi_e[t] += dei;
                gj_q += dqj; gj_h += dh; gj_e += dej;
            }
        }
        gj_q = oct_sum_b(gj_q); gj_h = oct_sum_b(gj_h); gj_e = oct_sum_b(gj_e);
        if (ii == 0 && jv) {
            atomicAdd(P.gq + js, w * (double)gj_q);
            atomicAdd(P.gsig + js, w * (double)gj_h);
            atomicAdd(P.geps + js, w * (double)gj_e);
        }
    }
#pragma unroll
    for (int t = 0; t < kGroupClusters; ++t) {
        if (t < ncl) {
            const float a = stride8_sum_b(gi_q[t]), b = stride8_sum_b(gi_h[t]), d = stride8_sum_b(gi_e[t]);
            const int is = D.slotOff + (g0 + t) * kClusterSize + ii;
            if (jj == 0 && P.sidx[is] >= 0) {
                atomicAdd(P.gq + is, w * (double)a);
                atomicAdd(P.gsig + is, w * (double)b);
                atomicAdd(P.geps + is, w * (double)d);
            }
        }
    }
}

__global__ void k_fill_f64(double* p, long long n) {
    for (long long k = blockIdx.x * (long long)blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        p[k] = 0.0;
}

__global__ void k_backward_gather(DevPtrs P, int sysBegin, double* dq, double* dsig, double* deps) {
    const int s = sysBegin + blockIdx.y;
    const SetDesc& D0 = P.desc[2 * s];
    const SetDesc& D1 = P.desc[2 * s + 1];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D0.natoms) return;
    const int a = D0.slotOff + P.slotOf[D0.setAtomOff + i];
    const int b = D1.slotOff + P.slotOf[D1.setAtomOff + i];
    const int k = D0.atomOff + i;
    const double sqk = 11.787089759563214;  // sqrt(138.935456)
    const double ep = P.prm64[k].y;          // 2 sqrt(eps)
    dq[k] = sqk * (P.gq[a] + P.gq[b]);
    dsig[k] = 0.5 * (P.gsig[a] + P.gsig[b]);
    // dE/deps = dE/de' de'/deps with de'/deps = 1/sqrt(eps) = 2/e': no derivative at eps = 0,
    // where 0 is returned (the convention of the FP64 oracle)
    deps[k] = ep > 0.0 ? (P.geps[a] + P.geps[b]) * 2.0 / ep : 0.0;
}

__global__ void k_set_positions(float4* posq, const float* xyz, long long n) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = posq[i];
    p.x = xyz[3 * i]; p.y = xyz[3 * i + 1]; p.z = xyz[3 * i + 2];
    posq[i] = p;
}

static inline unsigned cdivb(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

hipError_t launch_backward(const DevPtrs& P, long long totalSlots, int setBegin, int nsets, int maxGroups,
                           const double* wset, hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<long long>(cdivb(totalSlots, 256), 4096);
    k_fill_f64<<<blocks, 256, 0, st>>>(P.gq, totalSlots);
    k_fill_f64<<<blocks, 256, 0, st>>>(P.gsig, totalSlots);
    k_fill_f64<<<blocks, 256, 0, st>>>(P.geps, totalSlots);
    k_backward<<<dim3(maxGroups, nsets), 64, 0, st>>>(P, setBegin, wset);
    return hipGetLastError();
}

hipError_t launch_backward_gather(const DevPtrs& P, int sysBegin, int nsys, int maxAtoms, double* dq, double* dsig,
                                  double* deps, hipStream_t st) {
    k_backward_gather<<<dim3(cdivb(maxAtoms, 256), nsys), 256, 0, st>>>(P, sysBegin, dq, dsig, deps);
    return hipGetLastError();
}

hipError_t launch_set_positions(DevPtrs P, float4* posq, const float* xyz, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_set_positions<<<cdivb(n, 256), 256, 0, st>>>(posq, xyz, n);
    return hipGetLastError();
}

}  // namespace atm
