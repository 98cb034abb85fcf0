// Scalar ATM combine shared by the device energy kernel and the host
// reference-structured path.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include "atm_internal.h"

namespace atm {

// Soft-core map of the perturbation energy (u, umax, acore, ubcore) -> (u_sc, du_sc/du),
// ReferenceATMMetaForceKernels.cpp:26-37.
__host__ __device__ inline double softcore_map(double u, double umax, double acore, double ubcore, double* fp) {
    if (u <= ubcore) { *fp = 1.0; return u; }
    const double g = (u - ubcore) / (acore * (umax - ubcore));
    const double zeta = 1.0 + 2.0 * g * (g + 1.0);
    const double za = pow(zeta, acore);
    *fp = 4.0 * (2.0 * g + 1.0) / zeta * za / ((1.0 + za) * (1.0 + za));
    return (umax - ubcore) * (za - 1.0) / (za + 1.0) + ubcore;
}

// W, u_sc, f', bfp, sp from U1, U2 (ReferenceATMMetaForceKernels.cpp:83-98).
// ln(1 + e^x) is evaluated in a form that does not overflow for large x.
__host__ __device__ inline void atm_combine(const AtmParamsDev& p, double U1, double U2, SysResult& r) {
    const bool fwd = p.direction > 0.0;
    const double u = fwd ? U2 - U1 : U1 - U2;
    const double e0 = fwd ? U1 : U2;
    double fp;
    const double usc = softcore_map(u, p.umax, p.acore, p.ubcore, &fp);
    const double x = -p.alpha * (usc - p.u0);
    const double inv_ee = x > 0.0 ? exp(-x) / (1.0 + exp(-x)) : 1.0 / (1.0 + exp(x));  // 1/(1+e^x)
    double ebias = 0.0;
    if (p.alpha > 0.0) {
        const double lnee = x > 0.0 ? x + log1p(exp(-x)) : log1p(exp(x));
        ebias = (p.lambda2 - p.lambda1) / p.alpha * lnee;
    }
    ebias += p.lambda2 * usc + p.w0;
    const double bfp = (p.lambda2 - p.lambda1) * inv_ee + p.lambda1;
    r.W = e0 + ebias;
    r.usc = usc;
    r.U1 = U1;
    r.U2 = U2;
    r.u = u;
    r.fp = fp;
    r.bfp = bfp;
    r.sp = bfp * fp;
}


inline void atm_combine_host(const AtmParamsDev& p, double U1, double U2, SysResult& r) { atm_combine(p, U1, U2, r); }

}  // namespace atm
