// Host side of the PME variable-group evaluation and of the bonded terms
// (atm_pme_* and atm_bonded_* in include/atm_mi355x.h).
//
// An ATM evaluation with a PME NonbondedForce as the variable group: the
// reference's end-to-end setup (python/tests/test_abfe.py:42-43, 86-96) puts
// the NonbondedForce of an AMBER system (PME, 1 nm cutoff) in the ATM force's
// variable group; ATMMetaForceImpl::calcForcesAndEnergy
// (openmmapi/src/ATMMetaForceImpl.cpp:90-128) evaluates that group at x and at
// x + d and combines them (platforms/reference/src/ReferenceATMMetaForceKernels.cpp:56-110).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "atm_capi_util.h"
#include "atm_math.h"
#include "atm_pme.h"

using namespace atm;

namespace {

constexpr double kPi = 3.14159265358979323846;

int smooth_size(int n) {
    for (;; ++n) {
        int m = n;
        for (int f : {2, 3, 5, 7})
            while (m % f == 0) m /= f;
        if (m == 1) return n;
    }
}

// M_n(t) of the cardinal B-spline (support [0, n]).
double bspline_value(double t, int n) {
    if (n == 2) return (t >= 0.0 && t <= 2.0) ? 1.0 - std::fabs(t - 1.0) : 0.0;
    return t / (n - 1) * bspline_value(t, n - 1) + (n - t) / (n - 1) * bspline_value(t - 1.0, n - 1);
}

// |sum_{k=0}^{3} M_5(k+1) e^{2 pi i m k / K}|^2, near-zero moduli replaced by
// the mean of their neighbours.
std::vector<double> bspline_moduli(int K) {
    double w[4];
    for (int k = 0; k < 4; ++k) w[k] = bspline_value(k + 1.0, 5);
    std::vector<double> mod(K);
    for (int m = 0; m < K; ++m) {
        double c = 0.0, s = 0.0;
        for (int k = 0; k < 4; ++k) {
            const double a = 2.0 * kPi * (double)((long long)m * k % K) / K;
            c += w[k] * std::cos(a);
            s += w[k] * std::sin(a);
        }
        mod[m] = c * c + s * s;
    }
    std::vector<double> out = mod;
    for (int m = 0; m < K; ++m)
        if (mod[m] < 1e-7) out[m] = 0.5 * (mod[(m - 1 + K) % K] + mod[(m + 1) % K]);
    return out;
}

struct PmeForce {
    int device = 0;
    hipStream_t stream = nullptr;
    int n = 0;
    double box[3] = {0, 0, 0};
    double cutoff = 0.0, tol = 5e-4, alpha = 0.0;
    int K[3] = {0, 0, 0};
    bool dispersion = true;
    std::vector<double> q, sig, eps, displ;
    std::vector<int> exI, exJ;
    std::vector<double> exQQ, exSig, exEps;
    atm_params params{};
    double eoff[2] = {0.0, 0.0};  // atm_pme_set_energy_offsets
    bool haveParticles = false, haveParams = false, setupDone = false;
    double dispCoef = 0.0, selfEnergy = 0.0;
    int cellCap = 0;
    int ncellAlloc = 0, nslotAlloc = 0;

    DevBuf dX, dX2, dDispl, dCellOf, dRankOf, dCellCount, dCellStart, dCellAtoms, dWpos, dPrm, dPrm2, dExkey;
    DevBuf dCpos, dCprm, dCex, dExbits, dExpairs, dExprm, dQgrid, dG0, dG1, dTw, dBmod, dF, dEslot, dParts;
    DevBuf dResult, dOverflow, dOut, dParams;
    PmeDesc D{};
    PmePtrs P{};
    std::vector<double> hOut;
    // captured evaluation per (state 1 only, forces) and the descriptors it was captured with
    hipGraphExec_t graph[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};
    PmeDesc graphD[2][2];
    PmePtrs graphP[2][2];

    void dropGraphs() {
        for (auto& row : graph)
            for (hipGraphExec_t& g : row) {
                if (g) (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
    }
    ~PmeForce() {
        dropGraphs();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

PmeForce& pme_of(void* h) {
    if (!h) fail(ATM_ERR_INVALID_ARGUMENT, "null PME force");
    return *static_cast<PmeForce*>(h);
}

void upload(hipStream_t st, DevBuf& b, const void* src, size_t bytes) {
    b.alloc(bytes);
    if (bytes) HIPCHK(hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, st));
}

// Static tables: per-atom parameters, exclusion groups and bitmaps, exception
// parameters, B-spline moduli and twiddles.
void setup(PmeForce& S) {
    const int n = S.n;
    const double ke = ATM_ONE_4PI_EPS0, sk = std::sqrt(ke);
    std::vector<double2> prm(n), prm2(n);
    double q2 = 0.0;
    for (int i = 0; i < n; ++i) {
        prm[i] = make_double2(S.q[i] * sk, S.q[i]);
        prm2[i] = make_double2(0.5 * S.sig[i], 2.0 * std::sqrt(S.eps[i]));
        q2 += S.q[i] * S.q[i];
    }
    S.selfEnergy = -ke * S.alpha / std::sqrt(kPi) * q2;
    // dispersion correction coefficient (energy = coefficient / V)
    S.dispCoef = 0.0;
    if (S.dispersion) {
        std::map<std::pair<double, double>, long long> classes;
        for (int i = 0; i < n; ++i) classes[{S.sig[i], S.eps[i]}]++;
        std::vector<std::pair<std::pair<double, double>, long long>> cl(classes.begin(), classes.end());
        double s1 = 0.0, s2 = 0.0;
        for (size_t a = 0; a < cl.size(); ++a)
            for (size_t b = a; b < cl.size(); ++b) {
                const double na = (double)cl[a].second, nb = (double)cl[b].second;
                const double cnt = a == b ? na * (na + 1.0) / 2.0 : na * nb;
                const double s = 0.5 * (cl[a].first.first + cl[b].first.first);
                const double e = std::sqrt(cl[a].first.second * cl[b].first.second);
                const double s6 = s * s * s * s * s * s;
                s1 += cnt * e * s6 * s6;
                s2 += cnt * e * s6;
            }
        const double N = (double)n, pairs = N * (N + 1.0) / 2.0;
        s1 /= pairs;
        s2 /= pairs;
        const double rc = S.cutoff;
        S.dispCoef = 8.0 * N * N * kPi * (s1 / (9.0 * std::pow(rc, 9)) - s2 / (3.0 * std::pow(rc, 3)));
    }
    // exclusion groups: connected components of the exception graph
    std::vector<int> parent(n);
    std::iota(parent.begin(), parent.end(), 0);
    std::function<int(int)> root = [&](int a) {
        while (parent[a] != a) a = parent[a] = parent[parent[a]];
        return a;
    };
    const int nex = (int)S.exI.size();
    for (int k = 0; k < nex; ++k) {
        int a = root(S.exI[k]), b = root(S.exJ[k]);
        if (a != b) parent[std::max(a, b)] = std::min(a, b);
    }
    std::vector<int> grp(n), rank(n), size(n, 0);
    for (int i = 0; i < n; ++i) {
        grp[i] = root(i);
        rank[i] = size[grp[i]]++;  // atoms visited in index order: rank = order in the group
    }
    long long words = 1;  // word 0: all zero (singletons)
    std::vector<long long> rowOff(n, 0);
    for (int i = 0; i < n; ++i) {
        const int sz = size[grp[i]];
        if (sz > 1) {
            rowOff[i] = words;
            words += (sz + 31) / 32;
        }
    }
    if (words > (1ll << 31) - 1) fail(ATM_ERR_INVALID_ARGUMENT, "exclusion bitmaps too large");
    std::vector<uint32_t> bits((size_t)words, 0u);
    std::vector<int2> expairs(std::max(nex, 1));
    std::vector<double4> exprm(std::max(nex, 1));
    for (int k = 0; k < nex; ++k) {
        const int i = S.exI[k], j = S.exJ[k];
        bits[rowOff[i] + (rank[j] >> 5)] |= 1u << (rank[j] & 31);
        bits[rowOff[j] + (rank[i] >> 5)] |= 1u << (rank[i] & 31);
        expairs[k] = make_int2(i, j);
        exprm[k] = make_double4(ke * S.exQQ[k], S.exSig[k], 4.0 * S.exEps[k], ke * S.q[i] * S.q[j]);
    }
    std::vector<int4> exkey(n);
    for (int i = 0; i < n; ++i) exkey[i] = make_int4(grp[i], rank[i], (int)rowOff[i], 0);
    // twiddles and moduli
    std::vector<double2> tw;
    std::vector<double> bmod;
    for (int d = 0; d < 3; ++d) {
        S.D.twOff[d] = (int)tw.size();
        for (int t = 0; t < S.K[d]; ++t) {
            const double a = 2.0 * kPi * t / S.K[d];
            tw.push_back(make_double2(std::cos(a), std::sin(a)));
        }
        const std::vector<double> m = bspline_moduli(S.K[d]);
        bmod.insert(bmod.end(), m.begin(), m.end());
    }
    const hipStream_t st = S.stream;
    upload(st, S.dPrm, prm.data(), sizeof(double2) * n);
    upload(st, S.dPrm2, prm2.data(), sizeof(double2) * n);
    upload(st, S.dExkey, exkey.data(), sizeof(int4) * n);
    upload(st, S.dExbits, bits.data(), sizeof(uint32_t) * bits.size());
    upload(st, S.dExpairs, expairs.data(), sizeof(int2) * expairs.size());
    upload(st, S.dExprm, exprm.data(), sizeof(double4) * exprm.size());
    upload(st, S.dTw, tw.data(), sizeof(double2) * tw.size());
    upload(st, S.dBmod, bmod.data(), sizeof(double) * bmod.size());
    upload(st, S.dDispl, S.displ.data(), sizeof(double) * 3 * n);
    HIPCHK(hipStreamSynchronize(st));
    const long long ng = (long long)S.K[0] * S.K[1] * S.K[2];
    S.dX.alloc(sizeof(double) * 3 * n);
    S.dX2.alloc(sizeof(double) * 3 * n);
    S.dCellOf.alloc(sizeof(int) * n);
    S.dRankOf.alloc(sizeof(int) * n);
    S.dCellAtoms.alloc(sizeof(int) * n);
    S.dWpos.alloc(sizeof(double4) * n);
    S.dCpos.alloc(sizeof(double4) * n);
    S.dCprm.alloc(sizeof(double2) * n);
    S.dCex.alloc(sizeof(int4) * n);
    S.dQgrid.alloc(sizeof(long long) * ng);
    S.dG0.alloc(sizeof(double2) * ng);
    S.dG1.alloc(sizeof(double2) * ng);
    S.dF.alloc(sizeof(unsigned long long) * 6 * n);
    S.dParts.alloc(sizeof(double) * 10);
    S.dResult.alloc(sizeof(SysResult));
    S.dParams.alloc(sizeof(AtmParamsDev) + 2 * sizeof(double));  // parameters, then the two energy offsets
    S.dOverflow.alloc(sizeof(int) * 2);
    S.dOut.alloc(sizeof(double) * 3 * n);
    S.hOut.assign(3 * (size_t)n, 0.0);
    S.D.n = n;
    S.D.nex = nex;
    S.D.ngrid = (int)ng;
    for (int d = 0; d < 3; ++d) S.D.K[d] = S.K[d];
    S.D.alpha = S.alpha;
    S.D.twoAlphaSqrtPi = 2.0 * S.alpha / std::sqrt(kPi);
    S.D.pi2OverAlpha2 = kPi * kPi / (S.alpha * S.alpha);
    S.D.rc2 = S.cutoff * S.cutoff;
    S.D.selfEnergy = S.selfEnergy;
    S.setupDone = true;
}

// Box-dependent quantities: cell grid, scales, slot layout (buffers grow as needed).
void set_box(PmeForce& S, const double box[3]) {
    int m[3];
    for (int d = 0; d < 3; ++d) {
        if (!(box[d] > 0.0) || !std::isfinite(box[d])) fail(ATM_ERR_INVALID_ARGUMENT, "box edges must be positive");
        if (box[d] < 2.0 * S.cutoff) fail(ATM_ERR_INVALID_ARGUMENT, "box edges must be at least twice the cutoff");
        m[d] = (int)std::floor(box[d] / S.cutoff);
        if (m[d] < 3) fail(ATM_ERR_INVALID_ARGUMENT, "box edges must hold three cutoff lengths");
        m[d] = std::min(m[d], 40);  // at most 64,000 cells (one grid dimension of the direct kernel)
        S.box[d] = box[d];
        S.D.L[d] = box[d];
        S.D.invL[d] = 1.0 / box[d];
        S.D.m[d] = m[d];
    }
    S.D.ncell = m[0] * m[1] * m[2];
    const double V = box[0] * box[1] * box[2];
    S.D.recipEnergyScale = ATM_ONE_4PI_EPS0 / (2.0 * kPi * V);
    S.D.recipPhiScale = ATM_ONE_4PI_EPS0 / (kPi * V);
    S.D.dispersionEnergy = S.dispCoef / V;
    const double mean = (double)S.n / S.D.ncell;
    int cap = S.cellCap > 0 ? S.cellCap : (int)std::ceil(1.5 * mean + 64.0);  // grown on overflow
    cap = std::min(kPmeCellMax, (cap + 63) / 64 * 64);
    S.D.cellCap = cap;
    S.D.chunks = cap / 64;
    S.D.nslotDirect = S.D.chunks * kPmeHalfShell * kPmeJSplit * S.D.ncell;
    S.D.nslotEx = std::max(1, (S.D.nex + 255) / 256);
    S.D.nslot = S.D.nslotDirect + S.D.nslotEx + (S.D.ngrid + 255) / 256;
    // charge-grid tile of one cell: grid indices [floor(c K/m) - 5, ... + ceil(K/m) + 7)
    long long tb = 8;
    for (int d = 0; d < 3; ++d) {
        S.D.tile[d] = (S.K[d] + m[d] - 1) / m[d] + 7;
        tb *= S.D.tile[d];
    }
    S.D.tileBytes = tb <= kPmeTileLdsMax ? (int)tb : 0;
    if (S.D.ncell > S.ncellAlloc) {
        S.dCellCount.alloc(sizeof(int) * (S.D.ncell + 2));
        S.dCellStart.alloc(sizeof(int) * (S.D.ncell + 2));
        S.ncellAlloc = S.D.ncell;
    }
    if (S.D.nslot > S.nslotAlloc) {
        S.dEslot.alloc(sizeof(double) * 2 * (size_t)S.D.nslot);
        S.nslotAlloc = S.D.nslot;
    }
    PmePtrs& P = S.P;
    P.x = S.dX.as<double>();
    P.x2 = S.dX2.as<double>();
    P.displ = S.dDispl.as<double>();
    P.cellOf = S.dCellOf.as<int>();
    P.rankOf = S.dRankOf.as<int>();
    P.cellCount = S.dCellCount.as<int>();
    P.cellStart = S.dCellStart.as<int>();
    P.cellAtoms = S.dCellAtoms.as<int>();
    P.wpos = S.dWpos.as<double4>();
    P.prm = S.dPrm.as<double2>();
    P.prm2 = S.dPrm2.as<double2>();
    P.exkey = S.dExkey.as<int4>();
    P.cpos = S.dCpos.as<double4>();
    P.cprm = S.dCprm.as<double2>();
    P.cex = S.dCex.as<int4>();
    P.exbits = S.dExbits.as<uint32_t>();
    P.expairs = S.dExpairs.as<int2>();
    P.exprm = S.dExprm.as<double4>();
    P.qgrid = S.dQgrid.as<long long>();
    P.g0 = S.dG0.as<double2>();
    P.g1 = S.dG1.as<double2>();
    P.tw = S.dTw.as<double2>();
    P.bmod = S.dBmod.as<double>();
    P.F = S.dF.as<unsigned long long>();
    P.eslot = S.dEslot.as<double>();
    P.parts = S.dParts.as<double>();
    P.result = S.dResult.as<SysResult>();
    P.params = S.dParams.as<AtmParamsDev>();
    P.eoff = reinterpret_cast<const double*>(S.dParams.as<char>() + sizeof(AtmParamsDev));
    P.overflow = S.dOverflow.as<int>();
}

// ------------------------------------------------------------------ bonded

struct BondedSet {
    int device = 0;
    hipStream_t stream = nullptr;
    int n = 0, nb = 0, na = 0, nt = 0, nblocks = 0;
    DevBuf dX, dF, dBonds, dBprm, dAngles, dAprm, dTors, dTprm, dE;
    std::vector<long long> hF;
    std::vector<double> hE;
    ~BondedSet() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

BondedSet& bonded_of(void* h) {
    if (!h) fail(ATM_ERR_INVALID_ARGUMENT, "null bonded set");
    return *static_cast<BondedSet*>(h);
}

void check_device(int32_t device) {
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) fail(ATM_ERR_INVALID_ARGUMENT, "no GPU with index " + std::to_string(device));
    HIPCHK(hipSetDevice(device));
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ PME variable group

int atm_pme_create(int32_t device, const atm_pme_desc* desc, void** out) {
    return guarded([&] {
        if (!out || !desc) fail(ATM_ERR_INVALID_ARGUMENT, "null argument");
        *out = nullptr;
        const atm_pme_desc& d = *desc;
        if (d.num_atoms <= 0 || d.num_atoms > (1 << 24)) fail(ATM_ERR_INVALID_ARGUMENT, "num_atoms out of range");
        if (!(d.cutoff > 0.0)) fail(ATM_ERR_INVALID_ARGUMENT, "cutoff must be positive");
        if (d.alpha == 0.0 && !(d.ewald_tolerance > 0.0 && d.ewald_tolerance < 0.5))
            fail(ATM_ERR_INVALID_ARGUMENT, "ewald_tolerance must be in (0, 0.5)");
        if (d.alpha < 0.0) fail(ATM_ERR_INVALID_ARGUMENT, "alpha must not be negative");
        check_device(device);
        std::unique_ptr<PmeForce> S(new PmeForce());
        S->device = device;
        S->n = d.num_atoms;
        S->cutoff = d.cutoff;
        S->tol = d.ewald_tolerance;
        S->dispersion = d.dispersion_correction != 0;
        S->alpha = d.alpha > 0.0 ? d.alpha : std::sqrt(-std::log(2.0 * d.ewald_tolerance)) / d.cutoff;
        for (int k = 0; k < 3; ++k) {
            if (!(d.box[k] > 0.0)) fail(ATM_ERR_INVALID_ARGUMENT, "box edges must be positive");
            S->box[k] = d.box[k];
            int g = d.grid[k];
            if (g <= 0) {
                if (!(d.ewald_tolerance > 0.0)) fail(ATM_ERR_INVALID_ARGUMENT, "grid needs ewald_tolerance");
                g = smooth_size(std::max(6, (int)std::ceil(2.0 * S->alpha * d.box[k] /
                                                         (3.0 * std::pow(d.ewald_tolerance, 0.2)))));
            }
            if (g < 5 || g > 1024) fail(ATM_ERR_INVALID_ARGUMENT, "PME grid dimension out of range [5, 1024]");
            S->K[k] = g;
        }
        S->q.assign(S->n, 0.0);
        S->sig.assign(S->n, 0.1);
        S->eps.assign(S->n, 0.0);
        S->displ.assign(3 * (size_t)S->n, 0.0);
        S->params = atm_params{0.5, 0.5, 0.0, 0.0, 0.0, 1e300, 1e300, 1.0, 1.0};
        HIPCHK(hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking));
        *out = S.release();
    });
}

int atm_pme_destroy(void* h) {
    return guarded([&] {
        if (!h) return;
        PmeForce* S = static_cast<PmeForce*>(h);
        (void)hipSetDevice(S->device);
        (void)hipStreamSynchronize(S->stream);
        delete S;
    });
}

int atm_pme_get_info(void* h, double* alpha, int32_t grid[3]) {
    return guarded([&] {
        PmeForce& S = pme_of(h);
        if (alpha) *alpha = S.alpha;
        if (grid)
            for (int k = 0; k < 3; ++k) grid[k] = S.K[k];
    });
}

int atm_pme_set_particles(void* h, const double* charges, const double* sigmas, const double* epsilons,
                          const double* displacements) {
    return guarded([&] {
        PmeForce& S = pme_of(h);
        if (!charges && !sigmas && !epsilons) {  // displacements only
            if (!displacements) fail(ATM_ERR_INVALID_ARGUMENT, "nothing to set");
            if (!S.haveParticles) fail(ATM_ERR_INVALID_ARGUMENT, "particles not set");
            S.displ.assign(displacements, displacements + 3 * (size_t)S.n);
            S.setupDone = false;
            return;
        }
        if (!charges || !sigmas || !epsilons) fail(ATM_ERR_INVALID_ARGUMENT, "null particle array");
        for (int i = 0; i < S.n; ++i) {
            if (!std::isfinite(charges[i]) || !(sigmas[i] >= 0.0) || !(epsilons[i] >= 0.0) ||
                !std::isfinite(sigmas[i]) || !std::isfinite(epsilons[i]))
                fail(ATM_ERR_INVALID_ARGUMENT, "invalid parameters of particle " + std::to_string(i));
        }
        S.q.assign(charges, charges + S.n);
        S.sig.assign(sigmas, sigmas + S.n);
        S.eps.assign(epsilons, epsilons + S.n);
        if (displacements) S.displ.assign(displacements, displacements + 3 * (size_t)S.n);
        S.haveParticles = true;
        S.setupDone = false;
    });
}

int atm_pme_set_exceptions(void* h, int32_t num, const int32_t* pairs, const double* charge_prod, const double* sigma,
                           const double* epsilon) {
    return guarded([&] {
        PmeForce& S = pme_of(h);
        if (num < 0 || (num > 0 && (!pairs || !charge_prod || !sigma || !epsilon)))
            fail(ATM_ERR_INVALID_ARGUMENT, "invalid exception arrays");
        std::vector<long long> seen;
        seen.reserve(num);
        for (int k = 0; k < num; ++k) {
            const int i = pairs[2 * k], j = pairs[2 * k + 1];
            if (i < 0 || j < 0 || i >= S.n || j >= S.n || i == j)
                fail(ATM_ERR_INVALID_ARGUMENT, "exception " + std::to_string(k) + ": invalid atom pair");
            seen.push_back((long long)std::min(i, j) * S.n + std::max(i, j));
        }
        std::sort(seen.begin(), seen.end());
        if (std::adjacent_find(seen.begin(), seen.end()) != seen.end())
            fail(ATM_ERR_INVALID_ARGUMENT, "duplicate exception pair");
        S.exI.resize(num);
        S.exJ.resize(num);
        for (int k = 0; k < num; ++k) {
            S.exI[k] = pairs[2 * k];
            S.exJ[k] = pairs[2 * k + 1];
        }
        S.exQQ.assign(charge_prod, charge_prod + num);
        S.exSig.assign(sigma, sigma + num);
        S.exEps.assign(epsilon, epsilon + num);
        S.setupDone = false;
    });
}

int atm_pme_set_params(void* h, const atm_params* p) {
    return guarded([&] {
        PmeForce& S = pme_of(h);
        if (!p) fail(ATM_ERR_INVALID_ARGUMENT, "null parameters");
        S.params = *p;
        S.haveParams = true;
    });
}

int atm_pme_set_energy_offsets(void* h, double state1, double state2) {
    return guarded([&] {
        PmeForce& S = pme_of(h);
        if (!std::isfinite(state1) || !std::isfinite(state2)) fail(ATM_ERR_INVALID_ARGUMENT, "energy offsets must be finite");
        S.eoff[0] = state1;
        S.eoff[1] = state2;
    });
}

int atm_pme_eval(void* h, const double* positions, const double box[3], int32_t flags, double* forces,
                 atm_result* result, double* parts) {
    return guarded([&] {
        PmeForce& S = pme_of(h);
        if (!positions) fail(ATM_ERR_INVALID_ARGUMENT, "null positions");
        if (flags & ~ATM_EVAL_STATE1_ONLY) fail(ATM_ERR_INVALID_ARGUMENT, "unknown flags");
        if (!S.haveParticles) fail(ATM_ERR_INVALID_ARGUMENT, "particles not set");
        for (size_t k = 0; k < 3 * (size_t)S.n; ++k)
            if (!std::isfinite(positions[k])) fail(ATM_ERR_INVALID_ARGUMENT, "non-finite position");
        HIPCHK(hipSetDevice(S.device));
        if (!S.setupDone) setup(S);
        const hipStream_t st = S.stream;
        HIPCHK(hipMemcpyAsync(S.dX.p, positions, sizeof(double) * 3 * S.n, hipMemcpyHostToDevice, st));
        struct {
            AtmParamsDev prm;
            double eoff[2];
        } up{{S.params.lambda1, S.params.lambda2, S.params.alpha, S.params.u0, S.params.w0, S.params.umax,
              S.params.ubcore, S.params.acore, S.params.direction},
             {S.eoff[0], S.eoff[1]}};
        static_assert(sizeof(up) == sizeof(AtmParamsDev) + 2 * sizeof(double), "parameters and offsets packed");
        const bool single = (flags & ATM_EVAL_STATE1_ONLY) != 0;
        HIPCHK(hipMemcpyAsync(S.dParams.p, &up, sizeof(up), hipMemcpyHostToDevice, st));
        for (int attempt = 0; attempt < 2; ++attempt) {
            set_box(S, box ? box : S.box);
            // one captured graph per (state 1 only, forces); captured again when the
            // descriptor or a buffer changed (box, cell capacity, parameters)
            hipGraphExec_t& g = S.graph[single][forces != nullptr];
            PmeDesc& gD = S.graphD[single][forces != nullptr];
            PmePtrs& gP = S.graphP[single][forces != nullptr];
            if (g && (std::memcmp(&gD, &S.D, sizeof(PmeDesc)) != 0 || std::memcmp(&gP, &S.P, sizeof(PmePtrs)) != 0)) {
                (void)hipGraphExecDestroy(g);
                g = nullptr;
            }
            if (!g) {
                hipGraph_t graph = nullptr;
                HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                hipError_t e = launch_pme_fill(reinterpret_cast<unsigned long long*>(S.dOverflow.p), 1, st);
                if (e == hipSuccess) e = launch_pme_state(S.D, S.P, 0, st);
                if (e == hipSuccess && !single) e = launch_pme_state(S.D, S.P, 1, st);
                if (e == hipSuccess) e = launch_pme_finish(S.D, S.P, single, forces ? S.dOut.as<double>() : nullptr, st);
                const hipError_t ee = hipStreamEndCapture(st, &graph);
                HIPCHK(e);
                HIPCHK(ee);
                const hipError_t ei = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HIPCHK(ei);
                gD = S.D;
                gP = S.P;
            }
            HIPCHK(hipGraphLaunch(g, st));
            int ovf[2] = {0, 0};
            HIPCHK(hipMemcpyAsync(ovf, S.dOverflow.p, sizeof(int) * 2, hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            if (!ovf[0]) break;
            if (S.D.cellCap >= kPmeCellMax)
                fail(ATM_ERR_CAPACITY, "a cell holds more than " + std::to_string(kPmeCellMax) + " atoms");
            S.cellCap = kPmeCellMax;
        }
        SysResult r;
        double hp[10];
        HIPCHK(hipMemcpyAsync(&r, S.dResult.p, sizeof(SysResult), hipMemcpyDeviceToHost, st));
        HIPCHK(hipMemcpyAsync(hp, S.dParts.p, sizeof(hp), hipMemcpyDeviceToHost, st));
        if (forces)
            HIPCHK(hipMemcpyAsync(S.hOut.data(), S.dOut.p, sizeof(double) * 3 * S.n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        if (forces)
            for (size_t k = 0; k < 3 * (size_t)S.n; ++k) forces[k] += S.hOut[k];
        if (parts) std::memcpy(parts, hp, sizeof(hp));
        if (result) {
            result->energy = r.W;
            result->perturbation_energy = r.usc;
            result->state1_energy = r.U1;
            result->state2_energy = r.U2;
            result->raw_perturbation = r.u;
            result->softcore_derivative = r.fp;
            result->softplus_derivative = r.bfp;
            result->hybrid_weight = r.sp;
        }
    });
}

// ------------------------------------------------------------------ bonded terms

int atm_bonded_create(int32_t device, void** out) {
    return guarded([&] {
        if (!out) fail(ATM_ERR_INVALID_ARGUMENT, "null output pointer");
        *out = nullptr;
        check_device(device);
        std::unique_ptr<BondedSet> B(new BondedSet());
        B->device = device;
        HIPCHK(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
        *out = B.release();
    });
}

int atm_bonded_destroy(void* h) {
    return guarded([&] {
        if (!h) return;
        BondedSet* B = static_cast<BondedSet*>(h);
        (void)hipSetDevice(B->device);
        (void)hipStreamSynchronize(B->stream);
        delete B;
    });
}

int atm_bonded_set(void* h, int32_t num_atoms, int32_t num_bonds, const int32_t* bond_atoms, const double* bond_params,
                   int32_t num_angles, const int32_t* angle_atoms, const double* angle_params, int32_t num_torsions,
                   const int32_t* torsion_atoms, const double* torsion_params) {
    return guarded([&] {
        BondedSet& B = bonded_of(h);
        if (num_atoms <= 0 || num_bonds < 0 || num_angles < 0 || num_torsions < 0)
            fail(ATM_ERR_INVALID_ARGUMENT, "invalid counts");
        if ((num_bonds && (!bond_atoms || !bond_params)) || (num_angles && (!angle_atoms || !angle_params)) ||
            (num_torsions && (!torsion_atoms || !torsion_params)))
            fail(ATM_ERR_INVALID_ARGUMENT, "null term array");
        auto chk = [&](int a) {
            if (a < 0 || a >= num_atoms) fail(ATM_ERR_INVALID_ARGUMENT, "term atom index out of range");
        };
        std::vector<int2> bonds(std::max(num_bonds, 1));
        std::vector<double2> bprm(std::max(num_bonds, 1));
        for (int k = 0; k < num_bonds; ++k) {
            chk(bond_atoms[2 * k]);
            chk(bond_atoms[2 * k + 1]);
            bonds[k] = make_int2(bond_atoms[2 * k], bond_atoms[2 * k + 1]);
            bprm[k] = make_double2(bond_params[2 * k], bond_params[2 * k + 1]);
        }
        std::vector<int4> ang(std::max(num_angles, 1));
        std::vector<double2> aprm(std::max(num_angles, 1));
        for (int k = 0; k < num_angles; ++k) {
            for (int c = 0; c < 3; ++c) chk(angle_atoms[3 * k + c]);
            ang[k] = make_int4(angle_atoms[3 * k], angle_atoms[3 * k + 1], angle_atoms[3 * k + 2], 0);
            aprm[k] = make_double2(angle_params[2 * k], angle_params[2 * k + 1]);
        }
        std::vector<int4> tor(std::max(num_torsions, 1));
        std::vector<double4> tprm(std::max(num_torsions, 1));
        for (int k = 0; k < num_torsions; ++k) {
            for (int c = 0; c < 4; ++c) chk(torsion_atoms[4 * k + c]);
            tor[k] = make_int4(torsion_atoms[4 * k], torsion_atoms[4 * k + 1], torsion_atoms[4 * k + 2],
                               torsion_atoms[4 * k + 3]);
            tprm[k] = make_double4(torsion_params[3 * k], torsion_params[3 * k + 1], torsion_params[3 * k + 2], 0.0);
        }
        HIPCHK(hipSetDevice(B.device));
        HIPCHK(hipStreamSynchronize(B.stream));
        B.n = num_atoms;
        B.nb = num_bonds;
        B.na = num_angles;
        B.nt = num_torsions;
        B.nblocks = (num_bonds + num_angles + num_torsions + 255) / 256;
        upload(B.stream, B.dBonds, bonds.data(), sizeof(int2) * bonds.size());
        upload(B.stream, B.dBprm, bprm.data(), sizeof(double2) * bprm.size());
        upload(B.stream, B.dAngles, ang.data(), sizeof(int4) * ang.size());
        upload(B.stream, B.dAprm, aprm.data(), sizeof(double2) * aprm.size());
        upload(B.stream, B.dTors, tor.data(), sizeof(int4) * tor.size());
        upload(B.stream, B.dTprm, tprm.data(), sizeof(double4) * tprm.size());
        B.dX.alloc(sizeof(double) * 3 * num_atoms);
        B.dF.alloc(sizeof(long long) * 3 * num_atoms);
        B.dE.alloc(sizeof(double) * 3 * std::max(B.nblocks, 1));
        B.hF.assign(3 * (size_t)num_atoms, 0);
        B.hE.assign(3 * (size_t)std::max(B.nblocks, 1), 0.0);
        HIPCHK(hipStreamSynchronize(B.stream));
    });
}

int atm_bonded_eval(void* h, const double* positions, double* forces, double energies[3]) {
    return guarded([&] {
        BondedSet& B = bonded_of(h);
        if (!positions || !energies) fail(ATM_ERR_INVALID_ARGUMENT, "null argument");
        energies[0] = energies[1] = energies[2] = 0.0;
        if (B.nblocks == 0) return;
        HIPCHK(hipSetDevice(B.device));
        const hipStream_t st = B.stream;
        HIPCHK(hipMemcpyAsync(B.dX.p, positions, sizeof(double) * 3 * B.n, hipMemcpyHostToDevice, st));
        HIPCHK(launch_pme_fill(B.dF.as<unsigned long long>(), 3ll * B.n, st));
        BondedPtrs P;
        P.x = B.dX.as<double>();
        P.F = B.dF.as<unsigned long long>();
        P.bonds = B.dBonds.as<int2>();
        P.bprm = B.dBprm.as<double2>();
        P.angles = B.dAngles.as<int4>();
        P.aprm = B.dAprm.as<double2>();
        P.torsions = B.dTors.as<int4>();
        P.tprm = B.dTprm.as<double4>();
        P.nb = B.nb;
        P.na = B.na;
        P.nt = B.nt;
        HIPCHK(launch_bonded(P, B.dE.as<double>(), B.nblocks, st));
        HIPCHK(hipMemcpyAsync(B.hE.data(), B.dE.p, sizeof(double) * 3 * B.nblocks, hipMemcpyDeviceToHost, st));
        if (forces)
            HIPCHK(hipMemcpyAsync(B.hF.data(), B.dF.p, sizeof(long long) * 3 * B.n, hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        for (int b = 0; b < B.nblocks; ++b)
            for (int k = 0; k < 3; ++k) energies[k] += B.hE[3 * b + k];
        if (forces)
            for (size_t k = 0; k < 3 * (size_t)B.n; ++k) forces[k] += (double)B.hF[k] / ATM_FORCE_SCALE;
    });
}

}  // extern "C"
