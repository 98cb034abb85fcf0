// Host side of the C ABI declared in include/atm_mi355x.h.
//
// A context owns every device buffer of a batch of systems, one HIP stream and,
// per flag combination, one captured HIP graph of the batched evaluation.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <map>
#include <memory>
#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "atm_internal.h"
#include "atm_kernels.h"
#include "atm_math.h"
#include "atm_capi_util.h"
#include "atm_layout.h"
#include "../../include/atm_mi355x.h"

namespace atm {
// defined in atm_backward.hip
hipError_t launch_backward(const DevPtrs& P, long long totalSlots, int setBegin, int nsets, int maxGroups,
                           const double* wset, int split, hipStream_t st);
hipError_t launch_backward_gather(const DevPtrs& P, int sysBegin, int nsys, int maxAtoms, double* dq, double* dsig,
                                  double* deps, const double* wset, hipStream_t st, bool delta);
hipError_t launch_set_positions(DevPtrs P, float4* posq, const float* xyz, long long n, hipStream_t st);
hipError_t launch_check_particles(const double* q, const double* sig, const double* eps, long long n, int* bad,
                                  hipStream_t st);
hipError_t launch_set_particles(const DevPtrs& P, const double* q, const double* sig, const double* eps, long long n,
                                double sqk, hipStream_t st);
hipError_t launch_set_marks(double4* prm64, const unsigned char* mark, long long n, hipStream_t st);
}

namespace atm {
std::string& last_error() {
    static thread_local std::string msg;
    return msg;
}
}  // namespace atm

using namespace atm;

namespace {
struct Context {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    int nsys = 0;
    std::vector<atm_system_desc> sys;
    std::vector<std::array<double, 3>> tilt;  // per system (b_x, c_x, c_y): all zero = orthorhombic
    std::vector<int> atomOff;
    long long totalAtoms = 0;
    int maxAtoms = 0;
    std::vector<SetDesc> sets;
    std::vector<int> entryCapOverride;  // per set, 0 = automatic
    std::vector<int> alistCapOverride;  // per set, 0 = from the entry capacity
    int maxClusters = 0, maxGroups = 0, maxCells = 0, maxSlots = 0;
    long long totalCells = 0, totalCols = 0, totalSlots = 0, totalClusters = 0, totalGroups = 0, totalEntries = 0;
    long long totalAlist = 0;
    Tuning tuning;
    double tunedMs = 0.0;

    // device buffers
    DevBuf dDesc, dRt, dPosq, dDispl, dLjp, dPrm64, dParams, dRes, dOutF;
    DevBuf dBpos, dBcell, dBrank, dSlotOf, dCellCount, dCellStart, dColStart, dColCount;
    DevBuf dSpos, dSprm, dSprm64, dSidx, dSslot, dSf, dSd, dSdMask, dBbc, dBbh, dGinfo, dGcount, dGenergy, dEntries, dSetEnergy;
    DevBuf dAlist, dAcount;
    DevBuf dLptItems;  // the tile kernel's longest-first group order
    DevBuf dGq, dGsig, dGeps, dWset, dDq, dDsig, dDeps;
    DevBuf dFlag, dPosStage;  // validation counter; FP32 staging of host positions
    // exceptions (atm_set_exceptions): per system, near masks, far partner lists and
    // the terms with a nonzero charge product or epsilon; batch arrays on the device
    struct SysExceptions {
        std::vector<unsigned long long> mask;  // empty: none
        std::vector<std::vector<int>> far;
        std::vector<int2> pairs;
        std::vector<double4> prm;              // (k_e q'q, sigma, 4 epsilon, 0)
        int maxPerAtom = 0;
    };
    std::vector<SysExceptions> exc;
    bool hasExc = false;
    DevBuf dXmask, dXfarStart, dXfar, dXpair, dXprm, dXenergy;
    // batched Langevin dynamics (atm_md_*)
    DevBuf dMx, dMv, dMinvm, dMxSnap, dMvSnap;
    std::vector<double> masses;
    bool mdReady = false;
    uint64_t mdSeed = 0, mdStep = 0;
    DevPtrs P{};

    // host mirrors
    std::vector<float4> hPosq, hDispl;
    std::vector<float2> hLjp;
    std::vector<double4> hPrm64;
    std::vector<atm_params> params;
    std::vector<char> haveParticles;
    std::vector<atm_result> results;
    bool resultsValid = false;
    bool resultsOnHost = false;  // reference-structured path computed results on the host
    bool evaluated = false;
    bool paramsDirty = false;    // host params not yet on the device
    bool chargesOnDevice = false;  // the device charges were set from device arrays (host mirror stale)
    bool gradsReady = false;     // the last evaluation left dU/dp per set in gq, gsig, geps (ATM_EVAL_PARAM_GRADS)
    std::vector<double> eoff;    // [2 S] energies of other variable-group forces per set (atm_set_energy_offsets)
    DevBuf dEoff;
    bool eoffDirty = false;

    // state difference: displaced atoms (d != 0) per system and the exception terms whose
    // two atoms are displaced differently (batch index); see update_displaced
    std::vector<int> ligAtoms, lxidx;
    std::vector<unsigned char> ligMark;
    int maxLig = 0;
    bool deltaOK = false;        // every system displaces at most a quarter of its atoms
    bool lastDelta = false;      // the last batched evaluation used the state difference
    DevBuf dLig, dLigMark, dLenergy, dLxidx;

    std::map<int, hipGraphExec_t> graphs;
    // a second stream for work that runs beside the pair kernel (the displaced-atom kernels),
    // forked from and joined to the context stream (graph branches when captured)
    hipStream_t side = nullptr;
    hipEvent_t evFork = nullptr, evJoin = nullptr;

    ~Context() {
        dropGraphs();
        if (side) (void)hipStreamSynchronize(side);
        if (evFork) (void)hipEventDestroy(evFork);
        if (evJoin) (void)hipEventDestroy(evJoin);
        if (side) (void)hipStreamDestroy(side);
        if (own) (void)hipStreamDestroy(own);
    }
    void dropGraphs() {
        for (auto& kv : graphs) (void)hipGraphExecDestroy(kv.second);
        graphs.clear();
    }
};

void check_ctx(void* c) {
    if (!c) fail(ATM_ERR_INVALID_ARGUMENT, "null context");
}
Context& ctx_of(void* c) {
    check_ctx(c);
    return *static_cast<Context*>(c);
}
void check_sys(const Context& C, int s) {
    if (s < 0 || s >= C.nsys) fail(ATM_ERR_INVALID_ARGUMENT, "system index " + std::to_string(s) + " out of range");
}

// Every copy and clear is ordered on the context's stream: the stream is
// created non-blocking, so work on the legacy default stream (plain hipMemset,
// hipMemcpy) would not be ordered with the kernels.
void zero_on_stream(Context& C, void* p, size_t bytes) { HIPCHK(hipMemsetAsync(p, 0, bytes, C.stream)); }
void to_device(Context& C, void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, C.stream));
    HIPCHK(hipStreamSynchronize(C.stream));
}
void to_host(Context& C, void* dst, const void* src, size_t bytes) {
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, C.stream));
    HIPCHK(hipStreamSynchronize(C.stream));
}

}  // namespace

namespace atm {

void validate_desc(const atm_system_desc& d, int s) {
    const std::string pre = "system " + std::to_string(s) + ": ";
    if (d.num_atoms <= 0) fail(ATM_ERR_INVALID_ARGUMENT, pre + "num_atoms must be positive");
    if (d.num_atoms > (1 << 23)) fail(ATM_ERR_INVALID_ARGUMENT, pre + "more than 2^23 atoms");
    if (!(d.cutoff > 0.0)) fail(ATM_ERR_INVALID_ARGUMENT, pre + "cutoff must be positive");
    for (int k = 0; k < 3; ++k) {
        if (!(d.box[k] > 0.0) || !std::isfinite(d.box[k])) fail(ATM_ERR_INVALID_ARGUMENT, pre + "box edges must be positive");
        if (d.cutoff > 0.5 * d.box[k])
            fail(ATM_ERR_INVALID_ARGUMENT, pre + "the cutoff distance cannot be greater than half the periodic box size");
    }
    if (!(d.rf_dielectric > 0.0)) fail(ATM_ERR_INVALID_ARGUMENT, pre + "reaction-field dielectric must be positive");
}

// Geometry and capacities of one coordinate set.
SetDesc make_set(const atm_system_desc& d, const Tuning& tu, int s, int state, const double* tilt) {
    SetDesc D{};
    const bool tri = tilt && (tilt[0] != 0.0 || tilt[1] != 0.0 || tilt[2] != 0.0);
    const double L[3] = {d.box[0], d.box[1], d.box[2]};
    const double rc = d.cutoff;
    const double rho = d.num_atoms / (L[0] * L[1] * L[2]);
    const double ceT = tu.colScale * std::cbrt(kClusterSize / rho);
    const int nx = std::max(1, (int)std::floor(L[0] / ceT));
    const int ny = std::max(1, (int)std::floor(L[1] / ceT));
    const double cex = L[0] / nx, cey = L[1] / ny;
    const double zeT = std::min(cex, cey) / std::max(1, tu.zbinsPerColumn);
    const int nz = std::max(1, std::min((int)std::floor(L[2] / zeT), 4096));
    const double rcl = rc * (1.0 + 1e-5) + 1e-4;  // list cutoff: conservative w.r.t. FP32 box rounding
    int Rx = (int)std::ceil(rcl / cex), Ry = (int)std::ceil(rcl / cey);
    // triclinic: the c and b images move x and y by tilts, so every column is a candidate
    const int fullX = tri || (2 * Rx + 1 > nx), fullY = tri || (2 * Ry + 1 > ny);
    if (fullX) Rx = nx;
    if (fullY) Ry = ny;
    const int shiftOK = !fullX && !fullY && (Rx + 1) * cex < L[0] - rc && (Ry + 1) * cey < L[1] - rc;
    for (int k = 0; k < 3; ++k) {
        D.L[k] = (float)L[k];
        D.invL[k] = (float)(1.0 / L[k]);
    }
    D.ce[0] = (float)cex; D.ce[1] = (float)cey;
    D.invCe[0] = (float)(nx / L[0]); D.invCe[1] = (float)(ny / L[1]);
    D.ze = (float)(L[2] / nz); D.invZe = (float)(nz / L[2]);
    D.rc = (float)rc;
    D.rcList2 = (float)(rcl * rcl);
    D.rc2 = rc * rc;
    D.rc2lo = (float)(D.rc2 * (1.0 - 4e-6));
    D.rc2hi = (float)(D.rc2 * (1.0 + 4e-6));
    const double eps = d.rf_dielectric;
    D.krf = (float)((eps - 1.0) / ((2.0 * eps + 1.0) * rc * rc * rc));
    D.crf = (float)(3.0 * eps / ((2.0 * eps + 1.0) * rc));
    D.krf64 = (eps - 1.0) / ((2.0 * eps + 1.0) * rc * rc * rc);
    D.crf64 = 3.0 * eps / ((2.0 * eps + 1.0) * rc);
    D.nx = nx; D.ny = ny; D.nz = nz;
    D.Rx = Rx; D.Ry = Ry; D.fullX = fullX; D.fullY = fullY; D.shiftOK = shiftOK;
    D.tri = tri;
    for (int k = 0; k < 3; ++k) D.tilt[k] = tri ? (float)tilt[k] : 0.f;
    D.sys = s; D.state = state;
    D.natoms = d.num_atoms;
    const int ncols = nx * ny;
    D.slotCap = (d.num_atoms + 7 * ncols + 63) & ~63;  // set slot ranges start on multiples of 64
    D.clusterCap = D.slotCap / kClusterSize;
    D.groupCap = ncols + D.clusterCap / kGroupClusters + 1;
    const long long ncand = (long long)(fullX ? nx : 2 * Rx + 1) * (fullY ? ny : 2 * Ry + 1);
    const double m = (double)d.num_atoms / (ncols * (double)kClusterSize) + 1.0;
    long long cap = (long long)(ncand * (1.5 * m + 8.0));
    cap = std::max<long long>(64, (cap + 31) / 32 * 32);
    cap = std::min<long long>(cap, (long long)D.clusterCap + 32);
    D.entryCap = tu.listCapacity > 0 ? tu.listCapacity : (int)cap;
    return D;
}

// Largest partial force (kJ/mol/nm) that enters a set's 64-bit fixed-point sums.
// One slot receives at most: one flush per group that lists its cluster (tile
// kernel), up to 64 pair terms per group's j-batch sum (compacted kernel), one
// flush per i-cluster that lists it (atom-list kernel), the pair terms of its
// own group and extraTerms more (exception terms of one atom); each term is below fixMax * 2^24, so choosing
// fixMax <= 2^63 / (2^24 maxTerms) keeps every sum inside the int64 range.
float fixed_point_limit(const SetDesc& D, int extraTerms) {
    const double maxTerms = 64.0 * D.groupCap + (double)D.clusterCap + 8.0 * D.entryCap + 2.0 + extraTerms;
    return (float)std::min<double>(kFixedMax, std::ldexp(1.0, 39) / maxTerms);
}

}  // namespace atm

namespace {

// Batch arrays of the systems' exceptions (near masks of every atom, far partner
// lists as CSR, exception terms in system order; SetDesc::excOff follows the same
// order); all null when no system has any.
void upload_exceptions(Context& C) {
    DevPtrs& P = C.P;
    P.xmask = nullptr; P.xfarStart = nullptr; P.xfar = nullptr;
    P.xpair = nullptr; P.xprm = nullptr; P.xenergy = nullptr;
    C.hasExc = false;
    for (const auto& X : C.exc) C.hasExc |= !X.mask.empty();
    if (!C.hasExc) {
        C.dXmask.release(); C.dXfarStart.release(); C.dXfar.release();
        C.dXpair.release(); C.dXprm.release(); C.dXenergy.release();
        return;
    }
    std::vector<unsigned long long> mask(C.totalAtoms, 0ull);
    std::vector<int> farStart(C.totalAtoms + 1, 0), far;
    std::vector<int2> pairs;
    std::vector<double4> prm;
    for (int s = 0; s < C.nsys; ++s) {
        const auto& X = C.exc[s];
        const long long off = C.atomOff[s];
        for (int a = 0; a < C.sys[s].num_atoms; ++a) {
            farStart[off + a] = (int)far.size();
            if (X.mask.empty()) continue;
            mask[off + a] = X.mask[a];
            far.insert(far.end(), X.far[a].begin(), X.far[a].end());
        }
        pairs.insert(pairs.end(), X.pairs.begin(), X.pairs.end());
        prm.insert(prm.end(), X.prm.begin(), X.prm.end());
    }
    farStart[C.totalAtoms] = (int)far.size();
    if (far.empty()) far.push_back(-1);
    if (pairs.empty()) { pairs.push_back(make_int2(0, 0)); prm.push_back(make_double4(0, 0, 0, 0)); }
    C.dXmask.alloc(sizeof(unsigned long long) * mask.size());
    C.dXfarStart.alloc(sizeof(int) * farStart.size());
    C.dXfar.alloc(sizeof(int) * far.size());
    C.dXpair.alloc(sizeof(int2) * pairs.size());
    C.dXprm.alloc(sizeof(double4) * prm.size());
    C.dXenergy.alloc(sizeof(double) * 2 * C.nsys);
    // (DevBuf rounds small allocations up: copy the vectors' own sizes)
    to_device(C, C.dXmask.p, mask.data(), sizeof(unsigned long long) * mask.size());
    to_device(C, C.dXfarStart.p, farStart.data(), sizeof(int) * farStart.size());
    to_device(C, C.dXfar.p, far.data(), sizeof(int) * far.size());
    to_device(C, C.dXpair.p, pairs.data(), sizeof(int2) * pairs.size());
    to_device(C, C.dXprm.p, prm.data(), sizeof(double4) * prm.size());
    zero_on_stream(C, C.dXenergy.p, C.dXenergy.bytes);
    P.xmask = C.dXmask.as<unsigned long long>();
    P.xfarStart = C.dXfarStart.as<int>();
    P.xfar = C.dXfar.as<int>();
    P.xpair = C.dXpair.as<int2>();
    P.xprm = C.dXprm.as<double4>();
    P.xenergy = C.dXenergy.as<double>();
}

// Some set has exception terms to evaluate.
bool has_exception_terms(const Context& C) {
    for (const SetDesc& D : C.sets)
        if (D.excCount > 0) return true;
    return false;
}

void layout(Context& C) {
    C.sets.resize(2 * C.nsys);
    C.entryCapOverride.resize(2 * C.nsys, 0);
    C.alistCapOverride.resize(2 * C.nsys, 0);
    long long cellOff = 0, colOff = 0, slotOff = 0, clOff = 0, gOff = 0, eOff = 0, aOff = 0, setAtomOff = 0;
    long long xOff = 0;
    C.exc.resize(C.nsys);
    C.maxClusters = C.maxGroups = C.maxCells = C.maxSlots = 0;
    for (int s = 0; s < C.nsys; ++s) {
        const Context::SysExceptions& X = C.exc[s];
        for (int st = 0; st < 2; ++st) {
            const int c = 2 * s + st;
            SetDesc D = make_set(C.sys[s], C.tuning, s, st, C.tilt[s].data());
            if (C.entryCapOverride[c] > D.entryCap) D.entryCap = C.entryCapOverride[c];
            D.fixMax = fixed_point_limit(D, X.maxPerAtom);
            D.excOff = (int)xOff;
            D.excCount = (int)X.pairs.size();
            D.atomOff = C.atomOff[s];
            D.setAtomOff = (int)setAtomOff;
            D.cellOff = (int)cellOff;
            D.colOff = (int)colOff;
            D.slotOff = (int)slotOff;
            D.clusterOff = (int)clOff;
            D.groupOff = (int)gOff;
            D.entryOff = eOff;
            D.alistCap = std::max((D.entryCap + 7) / 8 * 8, C.alistCapOverride[c]);
            D.alistOff = aOff;
            if (D.clusterCap >= (1 << 20)) fail(ATM_ERR_UNSUPPORTED, "too many clusters in one system");
            setAtomOff += D.natoms;
            cellOff += (long long)D.nx * D.ny * D.nz;
            colOff += (long long)D.nx * D.ny;
            slotOff += D.slotCap;
            clOff += D.clusterCap;
            gOff += D.groupCap;
            eOff += (long long)D.groupCap * D.entryCap;
            aOff += (long long)D.clusterCap * D.alistCap;
            C.maxClusters = std::max(C.maxClusters, D.clusterCap);
            C.maxGroups = std::max(C.maxGroups, D.groupCap);
            C.maxCells = std::max(C.maxCells, D.nx * D.ny * D.nz);
            C.maxSlots = std::max(C.maxSlots, D.slotCap);
            C.sets[c] = D;
        }
        xOff += (long long)X.pairs.size();
    }
    if (xOff >= (1ll << 31)) fail(ATM_ERR_UNSUPPORTED, "too many exception terms in one batch");
    if (slotOff >= (1ll << 31) || cellOff >= (1ll << 31)) fail(ATM_ERR_UNSUPPORTED, "batch too large for 32-bit slot indices");
    C.totalCells = cellOff;
    C.totalCols = colOff;
    C.totalSlots = slotOff;
    C.totalClusters = clOff;
    C.totalGroups = gOff;
    C.totalEntries = eOff;
    C.totalAlist = aOff;
}

void update_displaced(Context& C);

void alloc_layout(Context& C) {
    C.dropGraphs();
    const long long S2 = 2ll * C.nsys;
    const long long setAtoms = 2 * C.totalAtoms;
    C.dDesc.alloc(sizeof(SetDesc) * S2);
    C.dRt.alloc(sizeof(SetRuntime) * S2);
    C.dBpos.alloc(sizeof(float4) * setAtoms);
    C.dBcell.alloc(sizeof(int) * setAtoms);
    C.dBrank.alloc(sizeof(int) * setAtoms);
    C.dSlotOf.alloc(sizeof(int) * setAtoms);
    C.dCellCount.alloc(sizeof(int) * C.totalCells);
    C.dCellStart.alloc(sizeof(int) * C.totalCells);
    C.dColStart.alloc(sizeof(int) * C.totalCols);
    C.dColCount.alloc(sizeof(int) * C.totalCols);
    C.dSpos.alloc(sizeof(float4) * C.totalSlots);
    C.dSprm.alloc(sizeof(float2) * C.totalSlots);
    C.dSprm64.alloc(sizeof(double4) * C.totalSlots);
    C.dSidx.alloc(sizeof(int) * C.totalSlots);
    C.dSslot.alloc(sizeof(int2) * C.totalSlots);
    C.dSf.alloc(sizeof(long long) * 3 * C.totalSlots);
    C.dSd.alloc(sizeof(double) * 3 * C.totalSlots);
    C.dSdMask.alloc(sizeof(unsigned long long) * (C.totalSlots / 64 + 1));
    zero_on_stream(C, C.dSd.p, sizeof(double) * 3 * C.totalSlots);  // side sums start clear
    C.dBbc.alloc(sizeof(float4) * C.totalClusters);
    C.dBbh.alloc(sizeof(float4) * C.totalClusters);
    C.dGinfo.alloc(sizeof(int4) * C.totalGroups);
    C.dGcount.alloc(sizeof(int) * C.totalGroups);
    C.dGenergy.alloc(sizeof(double) * 8 * C.totalGroups);
    C.dEntries.alloc(sizeof(uint32_t) * C.totalEntries);
    if (C.tuning.pairKernel == kPairKernelAtomLists) {  // only the atom-list pair kernel reads them
        C.dAlist.alloc(sizeof(uint32_t) * C.totalAlist);
        C.dAcount.alloc(sizeof(int) * C.totalClusters);
    } else {
        C.dAlist.release();
        C.dAcount.release();
    }
    C.dSetEnergy.alloc(sizeof(double) * S2);
    C.dLptItems.alloc(sizeof(int) * std::max<long long>(C.totalGroups, 1));
    C.dGq.release(); C.dGsig.release(); C.dGeps.release();
    C.gradsReady = false;
    zero_on_stream(C, C.dRt.p, sizeof(SetRuntime) * S2);
    zero_on_stream(C, C.dSdMask.p, C.dSdMask.bytes);
    zero_on_stream(C, C.dCellCount.p, sizeof(int) * C.totalCells);
    zero_on_stream(C, C.dGcount.p, sizeof(int) * C.totalGroups);
    to_device(C, C.dDesc.p, C.sets.data(), sizeof(SetDesc) * S2);

    DevPtrs& P = C.P;
    P.desc = C.dDesc.as<SetDesc>();
    P.rt = C.dRt.as<SetRuntime>();
    P.posq = C.dPosq.as<float4>();
    P.displ = C.dDispl.as<float4>();
    P.ljp = C.dLjp.as<float2>();
    P.prm64 = C.dPrm64.as<double4>();
    P.params = C.dParams.as<AtmParamsDev>();
    P.res = C.dRes.as<SysResult>();
    P.outF = C.dOutF.as<double>();
    P.bpos = C.dBpos.as<float4>();
    P.bcell = C.dBcell.as<int>();
    P.brank = C.dBrank.as<int>();
    P.slotOf = C.dSlotOf.as<int>();
    P.cellCount = C.dCellCount.as<int>();
    P.cellStart = C.dCellStart.as<int>();
    P.colStart = C.dColStart.as<int>();
    P.colCount = C.dColCount.as<int>();
    P.spos = C.dSpos.as<float4>();
    P.sprm = C.dSprm.as<float2>();
    P.sprm64 = C.dSprm64.as<double4>();
    P.sidx = C.dSidx.as<int>();
    P.sslot = C.dSslot.as<int2>();
    P.sf = C.dSf.as<long long>();
    P.sfn = C.totalSlots;
    P.sd = C.dSd.as<double>();
    P.sdMask = C.dSdMask.as<unsigned long long>();
    P.bbc = C.dBbc.as<float4>();
    P.bbh = C.dBbh.as<float4>();
    P.ginfo = C.dGinfo.as<int4>();
    P.gcount = C.dGcount.as<int>();
    P.genergy = C.dGenergy.as<double>();
    P.entries = C.dEntries.as<uint32_t>();
    P.alist = C.dAlist.as<uint32_t>();
    P.acount = C.dAcount.as<int>();
    P.setEnergy = C.dSetEnergy.as<double>();
    P.eoff = C.dEoff.as<double>();
    P.lptItems = C.dLptItems.as<int>();
    P.lptMode = C.tuning.pairOrder;
    P.anyTri = 0;
    for (const SetDesc& D : C.sets) P.anyTri |= D.tri;
    P.gq = P.gsig = P.geps = nullptr;
    upload_exceptions(C);
    C.dropGraphs();  // new pointers in P (the displaced-atom arrays may be reallocated below)
    update_displaced(C);
}

// The displaced atoms of every system and the exception terms between atoms displaced
// differently, into the set descriptors (both sets of a system) and device arrays.
// Called after every relayout and whenever displacements or exceptions change.
void update_displaced(Context& C) {
    const int oldMax = C.maxLig;
    const bool oldOK = C.deltaOK, oldLx = C.lxidx.empty();
    C.ligAtoms.clear();
    C.lxidx.clear();
    C.ligMark.assign(C.totalAtoms, 0);
    C.maxLig = 0;
    C.deltaOK = true;
    auto isDisplaced = [](const float4& d) { return d.x != 0.f || d.y != 0.f || d.z != 0.f; };
    for (int s = 0; s < C.nsys; ++s) {
        const int off = C.atomOff[s], n = C.sys[s].num_atoms;
        const int lb = (int)C.ligAtoms.size();
        for (int i = 0; i < n; ++i)
            if (isDisplaced(C.hDispl[off + i])) {
                C.ligAtoms.push_back(i);
                C.ligMark[off + i] = 1;
            }
        const int nl = (int)C.ligAtoms.size() - lb;
        C.maxLig = std::max(C.maxLig, nl);
        if (4ll * nl > n) C.deltaOK = false;
        const int xb = (int)C.lxidx.size();
        if (s < (int)C.exc.size()) {
            const auto& X = C.exc[s];
            for (size_t k = 0; k < X.pairs.size(); ++k) {
                const float4 da = C.hDispl[off + X.pairs[k].x], db = C.hDispl[off + X.pairs[k].y];
                if (da.x != db.x || da.y != db.y || da.z != db.z) C.lxidx.push_back(C.sets[2 * s].excOff + (int)k);
            }
        }
        for (int st = 0; st < 2; ++st) {
            SetDesc& D = C.sets[2 * s + st];
            D.ligOff = lb; D.ligCount = nl;
            D.lxOff = xb; D.lxCount = (int)C.lxidx.size() - xb;
        }
    }
    const size_t nl = std::max<size_t>(C.ligAtoms.size(), 1), nx = std::max<size_t>(C.lxidx.size(), 1);
    bool moved = false;  // a buffer was reallocated
    auto fit = [&](DevBuf& b, size_t bytes) {
        if (b.bytes < bytes) { b.alloc(bytes); moved = true; }
    };
    fit(C.dLig, sizeof(int) * nl);
    fit(C.dLenergy, sizeof(double) * 2 * kDisplacedParts * nl);
    fit(C.dLxidx, sizeof(int) * nx);
    fit(C.dLigMark, (size_t)C.totalAtoms);
    if (!C.ligAtoms.empty()) to_device(C, C.dLig.p, C.ligAtoms.data(), sizeof(int) * C.ligAtoms.size());
    if (!C.lxidx.empty()) to_device(C, C.dLxidx.p, C.lxidx.data(), sizeof(int) * C.lxidx.size());
    to_device(C, C.dLigMark.p, C.ligMark.data(), (size_t)C.totalAtoms);
    // the mark also rides in prm64.w (gathered into sprm64 with the parameters)
    bool markMoved = false;
    for (long long i = 0; i < C.totalAtoms; ++i)
        if (C.hPrm64[i].w != (double)C.ligMark[i]) {
            C.hPrm64[i].w = (double)C.ligMark[i];
            markMoved = true;
        }
    if (markMoved) {
        if (C.chargesOnDevice)
            // the device parameters are newer than the host mirror: rewrite w only
            HIPCHK(launch_set_marks(C.dPrm64.as<double4>(), C.dLigMark.as<unsigned char>(), C.totalAtoms, C.stream));
        else
            to_device(C, C.dPrm64.p, C.hPrm64.data(), sizeof(double4) * C.totalAtoms);
    }
    to_device(C, C.dDesc.p, C.sets.data(), sizeof(SetDesc) * C.sets.size());
    DevPtrs& P = C.P;
    P.ligAtoms = C.dLig.as<int>();
    P.ligMark = C.dLigMark.as<unsigned char>();
    P.lenergy = C.dLenergy.as<double>();
    P.lxidx = C.dLxidx.as<int>();
    // the graphs hold the pointers, the displaced-atom kernel's grid (maxLig) and whether the
    // displaced exception kernel runs
    if (moved || C.maxLig != oldMax || C.deltaOK != oldOK || C.lxidx.empty() != oldLx) C.dropGraphs();
}

// The batched evaluation with these flags uses the state difference.
bool use_delta(const Context& C, int flags) {
    (void)flags;
    return C.deltaOK && !C.tuning.fullState2;
}

// Pair kernel variant 5 (FP64 pairs only) reads the atom-level j lists.
bool uses_atom_lists(const Context& C, int flags) {
    return effective_pair_variant(C.P, C.tuning.pairKernel) == kPairKernelAtomLists && (flags & ATM_EVAL_FP32_PAIRS) == 0;
}

int energy_stride(const Context& C, int flags) {
    return pair_energy_stride(effective_pair_variant(C.P, C.tuning.pairKernel), (flags & ATM_EVAL_FP32_PAIRS) == 0,
                              C.tuning.pairSplit);
}

void require_ready(const Context& C) {
    for (int s = 0; s < C.nsys; ++s)
        if (!C.haveParticles[s])
            fail(ATM_ERR_STATE, "system " + std::to_string(s) + ": particles were never set (atm_set_particles)");
}

// Upload the ATM parameters of every system if any changed since the last upload.
static_assert(sizeof(atm_params) == sizeof(AtmParamsDev), "atm_params mirrors AtmParamsDev");

void flush_params(Context& C) {
    if (C.eoffDirty) {
        HIPCHK(hipMemcpyAsync(C.dEoff.p, C.eoff.data(), sizeof(double) * C.eoff.size(), hipMemcpyHostToDevice, C.stream));
        HIPCHK(hipStreamSynchronize(C.stream));
        C.eoffDirty = false;
    }
    if (!C.paramsDirty) return;
    HIPCHK(hipMemcpyAsync(C.dParams.p, C.params.data(), sizeof(AtmParamsDev) * C.nsys, hipMemcpyHostToDevice, C.stream));
    HIPCHK(hipStreamSynchronize(C.stream));
    C.paramsDirty = false;
}

// Per-slot gradient accumulators of the backward pass (allocated on first use;
// graphs captured before hold no pointer to them and are kept).
void ensure_grad_buffers(Context& C) {
    if (C.dGq.p) return;
    C.dGq.alloc(sizeof(double) * C.totalSlots);
    C.dGsig.alloc(sizeof(double) * C.totalSlots);
    C.dGeps.alloc(sizeof(double) * C.totalSlots);
    C.P.gq = C.dGq.as<double>();
    C.P.gsig = C.dGsig.as<double>();
    C.P.geps = C.dGeps.as<double>();
    C.dropGraphs();  // DevPtrs is captured by value: graphs must see the new pointers
}

// Batched evaluation of every set and system, kernel by kernel.  With `ev`
// (profiling), an event is recorded after each stage: ev[0] before the first
// kernel, then after build, pair list, pair kernel, exception terms and displaced
// atoms, energy and blend.
// With the state difference (use_delta) only the state-0 sets are built and paired
// (set stride 2), then the displaced-atom kernels evaluate what x + d changes.
void enqueue_batched(Context& C, int flags, hipEvent_t* ev = nullptr) {
    const hipStream_t st = C.stream;
    const bool delta = use_delta(C, flags);
    DevPtrs P = C.P;
    P.setStride = delta ? 2 : 1;
    P.lptMode = C.tuning.pairOrder;
    const int nsets = delta ? C.nsys : 2 * C.nsys;
    const bool forces = (flags & ATM_EVAL_FORCES) != 0;
    const bool grads = (flags & ATM_EVAL_PARAM_GRADS) != 0;
    if (ev) HIPCHK(hipEventRecord(ev[0], st));
    // with forces, the blend writes the cleared force itself
    const bool clearHere = (flags & ATM_EVAL_CLEAR_FORCES) && !forces;
    HIPCHK(launch_zero_counts(P, 0, nsets, C.maxCells, P.outF, clearHere ? 3 * C.totalAtoms : 0, st));
    HIPCHK(launch_build(P, 0, nsets, C.maxAtoms, C.maxCells, C.maxSlots, st));
    if (ev) HIPCHK(hipEventRecord(ev[1], st));
    // the displaced-atom kernels need only the x sets' cells and clusters: without stage
    // events they run on the side stream beside the pair list and pair kernel (not with
    // parameter gradients, whose accumulators the main stream clears first)
    const bool fork = delta && !ev && !grads;
    if (fork) {
        if (!C.side) {
            HIPCHK(hipStreamCreateWithFlags(&C.side, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&C.evFork, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&C.evJoin, hipEventDisableTiming));
        }
        HIPCHK(hipEventRecord(C.evFork, st));
        HIPCHK(hipStreamWaitEvent(C.side, C.evFork, 0));
        HIPCHK(launch_displaced(P, 0, C.nsys, C.maxLig, forces, C.side, grads));
        if (!C.lxidx.empty()) HIPCHK(launch_displaced_exceptions(P, 0, C.nsys, forces, C.side));
        HIPCHK(hipEventRecord(C.evJoin, C.side));
    }
    HIPCHK(launch_pairlist(P, 0, nsets, C.maxGroups, 4, st));
    if (uses_atom_lists(C, flags) && !grads) HIPCHK(launch_alist(P, 0, nsets, C.maxGroups, st));
    if (ev) HIPCHK(hipEventRecord(ev[2], st));
    if (grads) {
        // the backward mode of the tile kernel with unit weights: energies, pair counts
        // and dU/d(q', h, e') of every set in one pass over the pairs
        HIPCHK(launch_backward(P, C.totalSlots, 0, nsets, C.maxGroups, nullptr, C.tuning.pairSplit, st));
    } else {
        HIPCHK(launch_pairs(P, 0, nsets, C.maxGroups, C.tuning.pairWaves, forces, (flags & ATM_EVAL_FP32_PAIRS) == 0,
                            C.tuning.pairKernel, C.tuning.pairSplit, st));
    }
    if (ev) HIPCHK(hipEventRecord(ev[3], st));
    if (has_exception_terms(C)) HIPCHK(launch_exceptions(P, 0, nsets, forces, st));
    if (delta && !fork) {
        HIPCHK(launch_displaced(P, 0, C.nsys, C.maxLig, forces, st, grads));
        if (!C.lxidx.empty()) HIPCHK(launch_displaced_exceptions(P, 0, C.nsys, forces, st));
    }
    if (fork) HIPCHK(hipStreamWaitEvent(st, C.evJoin, 0));
    if (ev) HIPCHK(hipEventRecord(ev[4], st));
    HIPCHK(launch_energy(P, 0, C.nsys, grads ? C.tuning.pairSplit : energy_stride(C, flags), st, delta));
    if (ev) HIPCHK(hipEventRecord(ev[5], st));
    if (forces)
        HIPCHK(launch_blend(P, 0, C.nsys, C.maxAtoms, 0, 0.0, 0.0, (flags & ATM_EVAL_CLEAR_FORCES) != 0, st, delta));
    if (ev) HIPCHK(hipEventRecord(ev[6], st));
    C.lastDelta = delta;
}

void eval_batched(Context& C, int flags) {
    flush_params(C);
    if (flags & ATM_EVAL_PARAM_GRADS) ensure_grad_buffers(C);
    const int key = flags & (ATM_EVAL_FORCES | ATM_EVAL_CLEAR_FORCES | ATM_EVAL_FP32_PAIRS | ATM_EVAL_PARAM_GRADS);
    if (flags & ATM_EVAL_NO_GRAPH) {
        enqueue_batched(C, key);
        return;
    }
    // graphs are kept per flag set and per state mode (the mode follows the displacements)
    const int gkey = key | (use_delta(C, key) ? (1 << 24) : 0);
    C.lastDelta = use_delta(C, key);
    auto it = C.graphs.find(gkey);
    if (it == C.graphs.end()) {
        hipGraph_t g = nullptr;
        HIPCHK(hipStreamBeginCapture(C.stream, hipStreamCaptureModeThreadLocal));
        try {
            enqueue_batched(C, key);
        } catch (...) {
            hipGraph_t junk = nullptr;
            (void)hipStreamEndCapture(C.stream, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            throw;
        }
        HIPCHK(hipStreamEndCapture(C.stream, &g));
        hipGraphExec_t ex = nullptr;
        hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(e);
        it = C.graphs.emplace(gkey, ex).first;
    }
    HIPCHK(hipGraphLaunch(it->second, C.stream));
}

// Reference-structured path: per system, state 1 then state 2 evaluated as
// separate launch sequences with a blocking energy read-back each (the two
// inner-context evaluations of ATMMetaForceImpl.cpp:113,116), the soft-core
// and softplus on the host (ReferenceATMMetaForceKernels.cpp:83-98) and a
// hybrid-force launch with a host scalar (CommonATMMetaForceKernels.cpp:199-201).
void eval_plain(Context& C, int flags) {
    flush_params(C);
    const hipStream_t st = C.stream;
    const DevPtrs& P = C.P;
    const bool forces = (flags & ATM_EVAL_FORCES) != 0;
    for (int s = 0; s < C.nsys; ++s) {
        const SetDesc& D0 = C.sets[2 * s];
        if (flags & ATM_EVAL_CLEAR_FORCES)
            HIPCHK(launch_fill_zero(P.outF + 3ll * D0.atomOff, 3ll * D0.natoms, st));
        double U[2];
        for (int stt = 0; stt < 2; ++stt) {
            const int c = 2 * s + stt;
            const SetDesc& D = C.sets[c];
            HIPCHK(launch_zero_counts(P, c, 1, D.nx * D.ny * D.nz, nullptr, 0, st));
            HIPCHK(launch_build(P, c, 1, D.natoms, D.nx * D.ny * D.nz, D.slotCap, st));
            HIPCHK(launch_pairlist(P, c, 1, D.groupCap, 4, st));
            if (uses_atom_lists(C, flags)) HIPCHK(launch_alist(P, c, 1, D.groupCap, st));
            HIPCHK(launch_pairs(P, c, 1, D.groupCap, C.tuning.pairWaves, true, (flags & ATM_EVAL_FP32_PAIRS) == 0,
                                C.tuning.pairKernel, C.tuning.pairSplit, st));
            if (D.excCount > 0) HIPCHK(launch_exceptions(P, c, 1, true, st));
            HIPCHK(launch_set_energy(P, c, energy_stride(C, flags), true, st));
            HIPCHK(hipMemcpyAsync(&U[stt], P.setEnergy + c, sizeof(double), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
        }
        AtmParamsDev p;
        std::memcpy(&p, &C.params[s], sizeof(p));
        SysResult r;
        atm_combine_host(p, U[0], U[1], r);
        atm_result& o = C.results[s];
        o.energy = r.W; o.perturbation_energy = r.usc; o.state1_energy = r.U1; o.state2_energy = r.U2;
        o.raw_perturbation = r.u; o.softcore_derivative = r.fp; o.softplus_derivative = r.bfp; o.hybrid_weight = r.sp;
        if (forces) HIPCHK(launch_blend(P, s, 1, D0.natoms, 1, r.sp, p.direction, false, st));
    }
}

// State 1 only (a plain cutoff LJ + reaction-field evaluation at x): per
// system, the state-1 set's launch sequence, its energy read back to the host,
// F += F1.  Serves NonbondedForce evaluations outside the ATM force.
void eval_state1(Context& C, int flags) {
    flush_params(C);
    const hipStream_t st = C.stream;
    const DevPtrs& P = C.P;
    for (int s = 0; s < C.nsys; ++s) {
        const SetDesc& D = C.sets[2 * s];
        const int c = 2 * s;
        if (flags & ATM_EVAL_CLEAR_FORCES) HIPCHK(launch_fill_zero(P.outF + 3ll * D.atomOff, 3ll * D.natoms, st));
        HIPCHK(launch_zero_counts(P, c, 1, D.nx * D.ny * D.nz, nullptr, 0, st));
        HIPCHK(launch_build(P, c, 1, D.natoms, D.nx * D.ny * D.nz, D.slotCap, st));
        HIPCHK(launch_pairlist(P, c, 1, D.groupCap, 4, st));
        if (uses_atom_lists(C, flags)) HIPCHK(launch_alist(P, c, 1, D.groupCap, st));
        HIPCHK(launch_pairs(P, c, 1, D.groupCap, C.tuning.pairWaves, (flags & ATM_EVAL_FORCES) != 0,
                            (flags & ATM_EVAL_FP32_PAIRS) == 0, C.tuning.pairKernel, C.tuning.pairSplit, st));
        if (D.excCount > 0) HIPCHK(launch_exceptions(P, c, 1, (flags & ATM_EVAL_FORCES) != 0, st));
        HIPCHK(launch_set_energy(P, c, energy_stride(C, flags), false, st));
        double U = 0.0;
        HIPCHK(hipMemcpyAsync(&U, P.setEnergy + c, sizeof(double), hipMemcpyDeviceToHost, st));
        if (flags & ATM_EVAL_FORCES) HIPCHK(launch_blend(P, s, 1, D.natoms, 2, 0.0, 1.0, false, st));
        HIPCHK(hipStreamSynchronize(st));
        atm_result& o = C.results[s];
        o = atm_result{};
        o.energy = U;
        o.state1_energy = U;
    }
}

// Waits for the context stream.  The first call after an evaluation also checks
// the list-overflow flags and copies the results; later calls use the copies
// (every evaluation clears resultsValid).
void fetch_results(Context& C) {
    if (!C.evaluated) fail(ATM_ERR_STATE, "no evaluation has been run");
    HIPCHK(hipStreamSynchronize(C.stream));
    if (C.resultsValid) return;
    std::vector<SetRuntime> rt(2 * C.nsys);
    to_host(C, rt.data(), C.dRt.p, sizeof(SetRuntime) * rt.size());
    bool overflow = false;
    for (int c = 0; c < 2 * C.nsys; ++c) {
        // the flags are sticky over the evaluations since they were last cleared (a chunk
        // of MD steps); the new capacities cover the longest lists of the last evaluation
        if (rt[c].overflow & 1) {
            overflow = true;
            const long long need = (long long)rt[c].maxEntries + rt[c].maxEntries / 8;
            C.entryCapOverride[c] = (int)std::min<long long>(std::max<long long>(2ll * C.sets[c].entryCap, (need + 31) / 32 * 32),
                                                             (long long)C.sets[c].clusterCap + 32);
        }
        if (rt[c].overflow & 2) {
            // an atom list holds at most the 8 atoms of each of the set's clusters
            overflow = true;
            const long long need = (long long)rt[c].maxAtoms + rt[c].maxAtoms / 8;
            C.alistCapOverride[c] = (int)std::min<long long>(std::max<long long>(2ll * C.sets[c].alistCap, (need + 7) / 8 * 8),
                                                             8ll * (C.sets[c].clusterCap + 1));
        }
    }
    if (overflow) {
        // the relayout reallocates and zeroes the runtime records, clearing the flags
        layout(C);
        alloc_layout(C);
        C.evaluated = false;
        C.resultsValid = false;
        fail(ATM_ERR_CAPACITY, "neighbour list capacity exceeded; capacity doubled, re-run the evaluation");
    }
    if (!C.resultsValid) {
        if (!C.resultsOnHost) {
            std::vector<SysResult> r(C.nsys);
            to_host(C, r.data(), C.dRes.p, sizeof(SysResult) * C.nsys);
            for (int s = 0; s < C.nsys; ++s) {
                atm_result& o = C.results[s];
                o.energy = r[s].W; o.perturbation_energy = r[s].usc; o.state1_energy = r[s].U1; o.state2_energy = r[s].U2;
                o.raw_perturbation = r[s].u; o.softcore_derivative = r[s].fp; o.softplus_derivative = r[s].bfp;
                o.hybrid_weight = r[s].sp;
            }
        }
        C.resultsValid = true;
    }
}

void upload_system(Context& C, int s) {
    const size_t off = C.atomOff[s], n = C.sys[s].num_atoms;
    HIPCHK(hipMemcpyAsync(C.dPosq.as<float4>() + off, C.hPosq.data() + off, n * sizeof(float4), hipMemcpyHostToDevice, C.stream));
}


// ------------------------------------------------------------------ restraints

// A restraint set: the atoms it touches are gathered into a compact array
// (positions in, fixed-point forces out), so an evaluation moves only those.
struct RestraintSet {
    int device = 0;
    hipStream_t stream = nullptr;
    int numAtoms = 0;
    std::vector<int> used;              // compact index -> caller's atom index
    std::vector<atm_centroid_restraint> cent;
    std::vector<atm_position_restraint> posr;
    DevBuf dPos, dF, dMembers, dWeights, dCent, dPosr, dE;
    std::vector<double> hPos, hE;
    std::vector<long long> hF;
    ~RestraintSet() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

RestraintSet& restraints_of(void* r) {
    if (!r) fail(ATM_ERR_INVALID_ARGUMENT, "null restraint set");
    return *static_cast<RestraintSet*>(r);
}

// The FP32 positions the pair kernels read, rewritten from the FP64 MD positions.
hipError_t launch_md_init_posq(Context& C);

void rs_to_device(RestraintSet& R, void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, R.stream));
    HIPCHK(hipStreamSynchronize(R.stream));
}

}  // namespace

extern "C" {

const char* atm_version(void) { return "atm-mi355x 0.1.0 (ATMMetaForce 0.3.1 semantics)"; }
int32_t atm_abi_version(void) { return ATM_MI355X_ABI_VERSION; }
const char* atm_last_error(void) { return atm::last_error().c_str(); }

int atm_device_count(int32_t* count) {
    return guarded([&] {
        if (!count) fail(ATM_ERR_INVALID_ARGUMENT, "null count");
        int n = 0;
        hipError_t e = hipGetDeviceCount(&n);
        *count = e == hipSuccess ? n : 0;
    });
}

int atm_create(const atm_system_desc* systems, int32_t num_systems, int32_t device, void** out_context) {
    return guarded([&] {
        if (!out_context) fail(ATM_ERR_INVALID_ARGUMENT, "null output pointer");
        *out_context = nullptr;
        if (!systems || num_systems <= 0) fail(ATM_ERR_INVALID_ARGUMENT, "need at least one system");
        for (int s = 0; s < num_systems; ++s) validate_desc(systems[s], s);
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) fail(ATM_ERR_INVALID_ARGUMENT, "no GPU with index " + std::to_string(device));
        HIPCHK(hipSetDevice(device));
        std::unique_ptr<Context> C(new Context());
        C->device = device;
        HIPCHK(hipStreamCreateWithFlags(&C->own, hipStreamNonBlocking));
        C->stream = C->own;
        C->nsys = num_systems;
        C->sys.assign(systems, systems + num_systems);
        C->tilt.assign(num_systems, std::array<double, 3>{0.0, 0.0, 0.0});
        C->atomOff.resize(num_systems);
        long long off = 0;
        for (int s = 0; s < num_systems; ++s) {
            C->atomOff[s] = (int)off;
            off += systems[s].num_atoms;
            C->maxAtoms = std::max(C->maxAtoms, systems[s].num_atoms);
        }
        if (off >= (1ll << 31) / 3) fail(ATM_ERR_UNSUPPORTED, "batch too large");
        C->totalAtoms = off;
        C->hPosq.assign(off, make_float4(0.f, 0.f, 0.f, 0.f));
        C->hDispl.assign(off, make_float4(0.f, 0.f, 0.f, 0.f));
        C->hLjp.assign(off, make_float2(0.f, 0.f));
        C->hPrm64.assign(off, make_double4(0.0, 0.0, 0.0, 0.0));
        C->params.assign(num_systems, atm_params{0.5, 0.5, 0.0, 0.0, 0.0, 200.0, 100.0, 0.0625, 1.0});
        C->haveParticles.assign(num_systems, 0);
        C->results.assign(num_systems, atm_result{});
        C->dPosq.alloc(sizeof(float4) * off);
        C->dDispl.alloc(sizeof(float4) * off);
        C->dLjp.alloc(sizeof(float2) * off);
        C->dPrm64.alloc(sizeof(double4) * off);
        C->dParams.alloc(sizeof(AtmParamsDev) * num_systems);
        C->dRes.alloc(sizeof(SysResult) * num_systems);
        C->dOutF.alloc(sizeof(double) * 3 * off);
        zero_on_stream(*C, C->dPosq.p, sizeof(float4) * off);
        zero_on_stream(*C, C->dDispl.p, sizeof(float4) * off);
        zero_on_stream(*C, C->dLjp.p, sizeof(float2) * off);
        zero_on_stream(*C, C->dPrm64.p, sizeof(double4) * off);
        zero_on_stream(*C, C->dOutF.p, sizeof(double) * 3 * off);
        to_device(*C, C->dParams.p, C->params.data(), sizeof(AtmParamsDev) * num_systems);
        C->eoff.assign(2ll * num_systems, 0.0);
        C->dEoff.alloc(sizeof(double) * 2 * num_systems);
        to_device(*C, C->dEoff.p, C->eoff.data(), sizeof(double) * 2 * num_systems);
        layout(*C);
        alloc_layout(*C);
        *out_context = C.release();
    });
}

int atm_destroy(void* context) {
    return guarded([&] {
        if (!context) return;
        Context* C = static_cast<Context*>(context);
        (void)hipSetDevice(C->device);
        if (C->stream) (void)hipStreamSynchronize(C->stream);
        delete C;
    });
}

int atm_set_particles(void* context, int32_t sys, const double* charges, const double* sigmas, const double* epsilons,
                      const double* displacements) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (!charges || !sigmas || !epsilons) fail(ATM_ERR_INVALID_ARGUMENT, "null particle array");
        HIPCHK(hipSetDevice(C.device));
        const int n = C.sys[sys].num_atoms, off = C.atomOff[sys];
        const double sqk = std::sqrt(ATM_ONE_4PI_EPS0);
        for (int i = 0; i < n; ++i) {
            if (!(sigmas[i] >= 0.0) || !(epsilons[i] >= 0.0))
                fail(ATM_ERR_INVALID_ARGUMENT, "sigma and epsilon must be non-negative (atom " + std::to_string(i) + ")");
            C.hPosq[off + i].w = (float)(charges[i] * sqk);
            C.hLjp[off + i] = make_float2((float)(0.5 * sigmas[i]), (float)(2.0 * std::sqrt(epsilons[i])));
            C.hPrm64[off + i] = make_double4(0.5 * sigmas[i], 2.0 * std::sqrt(epsilons[i]), charges[i] * sqk,
                                             C.ligMark.empty() ? 0.0 : (double)C.ligMark[off + i]);
            if (displacements)
                C.hDispl[off + i] = make_float4((float)displacements[3 * i], (float)displacements[3 * i + 1],
                                                (float)displacements[3 * i + 2], 0.f);
        }
        upload_system(C, sys);
        // during dynamics the device positions are the integrator's: rewrite them over the
        // host copy just uploaded (which holds only the charges current)
        if (C.mdReady) HIPCHK(launch_md_init_posq(C));
        HIPCHK(hipMemcpyAsync(C.dLjp.as<float2>() + off, C.hLjp.data() + off, n * sizeof(float2), hipMemcpyHostToDevice, C.stream));
        HIPCHK(hipMemcpyAsync(C.dPrm64.as<double4>() + off, C.hPrm64.data() + off, n * sizeof(double4), hipMemcpyHostToDevice,
                              C.stream));
        HIPCHK(hipMemcpyAsync(C.dDispl.as<float4>() + off, C.hDispl.data() + off, n * sizeof(float4), hipMemcpyHostToDevice, C.stream));
        HIPCHK(hipStreamSynchronize(C.stream));
        if (displacements) update_displaced(C);
        C.haveParticles[sys] = 1;
    });
}

int atm_set_energy_offsets(void* context, int32_t sys, double state1, double state2) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (!std::isfinite(state1) || !std::isfinite(state2)) fail(ATM_ERR_INVALID_ARGUMENT, "energy offsets must be finite");
        C.eoff[2 * sys] = state1;
        C.eoff[2 * sys + 1] = state2;
        C.eoffDirty = true;
    });
}

int atm_set_exceptions(void* context, int32_t sys, int32_t count, const int32_t* pairs, const double* charge_prod,
                       const double* sigma, const double* epsilon) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (count < 0) fail(ATM_ERR_INVALID_ARGUMENT, "negative exception count");
        if (count > 0 && (!pairs || !charge_prod || !sigma || !epsilon)) fail(ATM_ERR_INVALID_ARGUMENT, "null exception array");
        HIPCHK(hipSetDevice(C.device));
        const int n = C.sys[sys].num_atoms;
        Context::SysExceptions X;
        if (count > 0) {
            X.mask.assign(n, 0ull);
            X.far.assign(n, std::vector<int>());
            std::vector<int> per(n, 0);
            std::vector<unsigned long long> keys((size_t)count);
            for (int k = 0; k < count; ++k) {
                const int i = pairs[2 * k], j = pairs[2 * k + 1];
                const std::string pre = "exception " + std::to_string(k) + ": ";
                if (i < 0 || j < 0 || i >= n || j >= n || i == j) fail(ATM_ERR_INVALID_ARGUMENT, pre + "invalid atom pair");
                if (!std::isfinite(charge_prod[k]) || !(sigma[k] >= 0.0) || !(epsilon[k] >= 0.0) ||
                    !std::isfinite(sigma[k]) || !std::isfinite(epsilon[k]))
                    fail(ATM_ERR_INVALID_ARGUMENT, pre + "charge product must be finite, sigma and epsilon finite and non-negative");
                keys[k] = (unsigned long long)std::min(i, j) << 32 | (unsigned)std::max(i, j);
                const int d = j - i;
                if (d >= -31 && d <= 31) {
                    X.mask[i] |= 1ull << (32 + d);
                    X.mask[j] |= 1ull << (32 - d);
                } else {
                    X.mask[i] |= 1ull << 32;
                    X.mask[j] |= 1ull << 32;
                    X.far[i].push_back(j);
                    X.far[j].push_back(i);
                }
                if (charge_prod[k] != 0.0 || epsilon[k] != 0.0) {
                    X.pairs.push_back(make_int2(i, j));
                    X.prm.push_back(make_double4(ATM_ONE_4PI_EPS0 * charge_prod[k], sigma[k], 4.0 * epsilon[k], 0.0));
                    ++per[i];
                    ++per[j];
                }
            }
            std::sort(keys.begin(), keys.end());
            if (std::adjacent_find(keys.begin(), keys.end()) != keys.end())
                fail(ATM_ERR_INVALID_ARGUMENT, "an atom pair is listed twice among the exceptions");
            X.maxPerAtom = *std::max_element(per.begin(), per.end());
        }
        C.exc.resize(C.nsys);
        C.exc[sys] = std::move(X);
        layout(C);
        alloc_layout(C);
        C.evaluated = false;
        C.resultsValid = false;
    });
}

int atm_set_displacements(void* context, int32_t sys, const double* d) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (!d) fail(ATM_ERR_INVALID_ARGUMENT, "null displacement array");
        HIPCHK(hipSetDevice(C.device));
        const int n = C.sys[sys].num_atoms, off = C.atomOff[sys];
        for (int i = 0; i < n; ++i)
            C.hDispl[off + i] = make_float4((float)d[3 * i], (float)d[3 * i + 1], (float)d[3 * i + 2], 0.f);
        HIPCHK(hipMemcpyAsync(C.dDispl.as<float4>() + off, C.hDispl.data() + off, n * sizeof(float4), hipMemcpyHostToDevice, C.stream));
        HIPCHK(hipStreamSynchronize(C.stream));
        update_displaced(C);
    });
}

int atm_set_positions(void* context, int32_t sys, const double* x) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (!x) fail(ATM_ERR_INVALID_ARGUMENT, "null position array");
        HIPCHK(hipSetDevice(C.device));
        const int n = C.sys[sys].num_atoms, off = C.atomOff[sys];
        for (int i = 0; i < n; ++i) {
            if (!std::isfinite(x[3 * i]) || !std::isfinite(x[3 * i + 1]) || !std::isfinite(x[3 * i + 2]))
                fail(ATM_ERR_INVALID_ARGUMENT, "non-finite position (atom " + std::to_string(i) + ")");
            float4& p = C.hPosq[off + i];
            p.x = (float)x[3 * i]; p.y = (float)x[3 * i + 1]; p.z = (float)x[3 * i + 2];
        }
        if (C.chargesOnDevice) {
            // the device holds charges the host mirror does not: replace x, y, z only
            std::vector<float> xyz(3 * (size_t)n);
            for (int i = 0; i < n; ++i) {
                xyz[3 * i] = C.hPosq[off + i].x; xyz[3 * i + 1] = C.hPosq[off + i].y; xyz[3 * i + 2] = C.hPosq[off + i].z;
            }
            if (!C.dPosStage.p) C.dPosStage.alloc(sizeof(float) * 3 * C.totalAtoms);
            HIPCHK(hipMemcpyAsync(C.dPosStage.as<float>() + 3ll * off, xyz.data(), sizeof(float) * 3 * n,
                                  hipMemcpyHostToDevice, C.stream));
            HIPCHK(launch_set_positions(C.P, C.dPosq.as<float4>() + off, C.dPosStage.as<float>() + 3ll * off, n, C.stream));
        } else {
            upload_system(C, sys);
        }
        HIPCHK(hipStreamSynchronize(C.stream));
    });
}

int atm_set_positions_device(void* context, const float* positions_dev) {
    return guarded([&] {
        Context& C = ctx_of(context);
        if (!positions_dev) fail(ATM_ERR_INVALID_ARGUMENT, "null device pointer");
        HIPCHK(hipSetDevice(C.device));
        HIPCHK(launch_set_positions(C.P, C.dPosq.as<float4>(), positions_dev, C.totalAtoms, C.stream));
    });
}

int atm_set_particles_device(void* context, const double* charges, const double* sigmas, const double* epsilons) {
    return guarded([&] {
        Context& C = ctx_of(context);
        if (!charges || !sigmas || !epsilons) fail(ATM_ERR_INVALID_ARGUMENT, "null device array");
        HIPCHK(hipSetDevice(C.device));
        if (!C.dFlag.p) C.dFlag.alloc(sizeof(int));
        zero_on_stream(C, C.dFlag.p, sizeof(int));
        HIPCHK(launch_check_particles(charges, sigmas, epsilons, C.totalAtoms, C.dFlag.as<int>(), C.stream));
        int bad = 0;
        to_host(C, &bad, C.dFlag.p, sizeof(int));
        if (bad) fail(ATM_ERR_INVALID_ARGUMENT, std::to_string(bad) + " atoms with a non-finite parameter or a negative "
                                                "sigma or epsilon (nothing was changed)");
        HIPCHK(launch_set_particles(C.P, charges, sigmas, epsilons, C.totalAtoms, std::sqrt(ATM_ONE_4PI_EPS0), C.stream));
        HIPCHK(hipStreamSynchronize(C.stream));
        C.chargesOnDevice = true;
        std::fill(C.haveParticles.begin(), C.haveParticles.end(), 1);
    });
}

int atm_set_box(void* context, int32_t sys, const double* box3) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (!box3) fail(ATM_ERR_INVALID_ARGUMENT, "null box");
        atm_system_desc d = C.sys[sys];
        for (int k = 0; k < 3; ++k) d.box[k] = box3[k];
        validate_desc(d, sys);
        HIPCHK(hipSetDevice(C.device));
        HIPCHK(hipStreamSynchronize(C.stream));
        C.sys[sys] = d;
        C.tilt[sys] = {0.0, 0.0, 0.0};
        layout(C);
        alloc_layout(C);
        C.evaluated = false;
        C.resultsValid = false;
    });
}

int atm_set_box_vectors(void* context, int32_t sys, const double* v) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (!v) fail(ATM_ERR_INVALID_ARGUMENT, "null box vectors");
        for (int k = 0; k < 9; ++k)
            if (!std::isfinite(v[k])) fail(ATM_ERR_INVALID_ARGUMENT, "box vectors must be finite");
        // OpenMM's reduced form: a along x, b in the x-y plane, each vector's tilt at
        // most half the edge of the ones before it
        if (v[1] != 0.0 || v[2] != 0.0 || v[5] != 0.0)
            fail(ATM_ERR_INVALID_ARGUMENT, "box vectors must be in reduced form (a = (ax, 0, 0), b = (bx, by, 0))");
        if (!(v[0] > 0.0 && v[4] > 0.0 && v[8] > 0.0))
            fail(ATM_ERR_INVALID_ARGUMENT, "box vectors must have positive diagonal components");
        if (std::fabs(v[3]) > 0.5 * v[0] || std::fabs(v[6]) > 0.5 * v[0] || std::fabs(v[7]) > 0.5 * v[4])
            fail(ATM_ERR_INVALID_ARGUMENT,
                 "box vectors must be in reduced form (ax >= 2 |bx|, ax >= 2 |cx|, by >= 2 |cy|)");
        atm_system_desc d = C.sys[sys];
        d.box[0] = v[0]; d.box[1] = v[4]; d.box[2] = v[8];
        validate_desc(d, sys);
        HIPCHK(hipSetDevice(C.device));
        HIPCHK(hipStreamSynchronize(C.stream));
        C.sys[sys] = d;
        C.tilt[sys] = {v[3], v[6], v[7]};
        layout(C);
        alloc_layout(C);
        C.evaluated = false;
        C.resultsValid = false;
    });
}

int atm_set_params(void* context, int32_t sys, const atm_params* p) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (!p) fail(ATM_ERR_INVALID_ARGUMENT, "null parameters");
        HIPCHK(hipSetDevice(C.device));
        C.params[sys] = *p;
        C.paramsDirty = true;  // uploaded (all systems at once) before the next evaluation
    });
}

int atm_eval(void* context, int32_t flags) {
    return guarded([&] {
        Context& C = ctx_of(context);
        require_ready(C);
        if ((flags & ATM_EVAL_PARAM_GRADS) &&
            (flags & (ATM_EVAL_FORCES | ATM_EVAL_FP32_PAIRS | ATM_EVAL_PLAIN | ATM_EVAL_STATE1_ONLY)))
            fail(ATM_ERR_INVALID_ARGUMENT,
                 "ATM_EVAL_PARAM_GRADS is an energy-only evaluation of the batched FP64 path");
        HIPCHK(hipSetDevice(C.device));
        C.resultsValid = false;
        C.gradsReady = false;
        C.lastDelta = false;
        if (flags & ATM_EVAL_STATE1_ONLY) {
            eval_state1(C, flags);
            C.resultsOnHost = true;
        } else if (flags & ATM_EVAL_PLAIN) {
            eval_plain(C, flags);
            C.resultsOnHost = true;
        } else {
            eval_batched(C, flags);
            C.resultsOnHost = false;
            C.gradsReady = (flags & ATM_EVAL_PARAM_GRADS) != 0;
        }
        C.evaluated = true;
    });
}

int atm_synchronize(void* context) {
    return guarded([&] {
        Context& C = ctx_of(context);
        HIPCHK(hipSetDevice(C.device));
        HIPCHK(hipStreamSynchronize(C.stream));
    });
}

int atm_get_results(void* context, int32_t sys, atm_result* out) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (!out) fail(ATM_ERR_INVALID_ARGUMENT, "null output");
        HIPCHK(hipSetDevice(C.device));
        fetch_results(C);
        *out = C.results[sys];
    });
}

int atm_get_forces_fixed(void* context, int32_t sys, int64_t* f) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (!f) fail(ATM_ERR_INVALID_ARGUMENT, "null output");
        HIPCHK(hipSetDevice(C.device));
        fetch_results(C);
        const int n = C.sys[sys].num_atoms;
        std::vector<double> soa(3ll * n);
        to_host(C, soa.data(), C.dOutF.as<double>() + 3ll * C.atomOff[sys], sizeof(double) * 3 * n);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) {
                const double v = std::fmin(std::fmax(soa[(size_t)k * n + i] * ATM_FORCE_SCALE, -9.2e18), 9.2e18);
                f[3 * i + k] = (int64_t)std::llrint(v);
            }
    });
}

int atm_get_forces(void* context, int32_t sys, double* f) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (!f) fail(ATM_ERR_INVALID_ARGUMENT, "null output");
        HIPCHK(hipSetDevice(C.device));
        fetch_results(C);
        const int n = C.sys[sys].num_atoms;
        std::vector<double> soa(3ll * n);
        to_host(C, soa.data(), C.dOutF.as<double>() + 3ll * C.atomOff[sys], sizeof(double) * 3 * n);
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) f[3 * i + k] = soa[(size_t)k * n + i];
    });
}

int atm_clear_forces(void* context) {
    return guarded([&] {
        Context& C = ctx_of(context);
        HIPCHK(hipSetDevice(C.device));
        HIPCHK(launch_fill_zero(C.P.outF, 3 * C.totalAtoms, C.stream));
        HIPCHK(hipStreamSynchronize(C.stream));
    });
}

int atm_get_pairs(void* context, int32_t sys, int32_t state, int32_t* pairs, int64_t capacity, int64_t* count) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (state != 0 && state != 1) fail(ATM_ERR_INVALID_ARGUMENT, "state must be 0 or 1");
        if (!count) fail(ATM_ERR_INVALID_ARGUMENT, "null count");
        if (capacity > 0 && !pairs) fail(ATM_ERR_INVALID_ARGUMENT, "null pair buffer");
        HIPCHK(hipSetDevice(C.device));
        fetch_results(C);
        const long long cap = std::max<long long>(capacity, 0);
        auto less = [](const int2& a, const int2& b) { return a.x != b.x ? a.x < b.x : a.y < b.y; };
        // every pair a set's lists hold (the traversal of the tile kernel), or the pairs the
        // displaced-atom kernel evaluates for `st`, sorted
        auto dump = [&](bool displaced, int st) {
            DevBuf ctr;
            ctr.alloc(sizeof(unsigned long long) * 4);
            unsigned long long h[4];
            long long n = 0;
            std::vector<int2> v;
            for (int pass = 0; pass < 2; ++pass) {  // count, then fetch
                DevBuf out;
                out.alloc(sizeof(int2) * std::max<long long>(n, 1));
                zero_on_stream(C, ctr.p, sizeof(unsigned long long) * 4);
                if (displaced)
                    HIPCHK(launch_displaced_dump(C.P, sys, st, C.sets[2 * sys].ligCount, out.as<int2>(), n,
                                                 ctr.as<unsigned long long>(), C.stream));
                else
                    HIPCHK(launch_pairs_dump(C.P, 2 * sys + st, C.sets[2 * sys + st].groupCap, out.as<int2>(), n,
                                             ctr.as<unsigned long long>(), C.stream));
                to_host(C, h, ctr.p, sizeof(h));
                if (pass == 0) { n = (long long)h[0]; continue; }
                v.resize(std::min<long long>((long long)h[0], n));
                if (!v.empty()) to_host(C, v.data(), out.p, sizeof(int2) * v.size());
            }
            std::sort(v.begin(), v.end(), less);
            return v;
        };
        std::vector<int2> v;
        if (state == 1 && C.lastDelta) {
            // x + d was evaluated as the difference from x: the state-0 pairs without those of
            // the displaced atoms at x, with those at x + d
            const std::vector<int2> all = dump(false, 0), d0 = dump(true, 0), d1 = dump(true, 1);
            std::vector<int2> keep;
            std::set_difference(all.begin(), all.end(), d0.begin(), d0.end(), std::back_inserter(keep), less);
            std::merge(keep.begin(), keep.end(), d1.begin(), d1.end(), std::back_inserter(v), less);
        } else {
            v = dump(false, state);
        }
        *count = (int64_t)v.size();
        const long long nout = std::min<long long>((long long)v.size(), cap);
        for (long long k = 0; k < nout; ++k) { pairs[2 * k] = v[k].x; pairs[2 * k + 1] = v[k].y; }
    });
}

int atm_get_stats(void* context, int64_t* pairs_in_cutoff, int64_t* pairs_evaluated, int64_t* list_entries) {
    return guarded([&] {
        Context& C = ctx_of(context);
        HIPCHK(hipSetDevice(C.device));
        fetch_results(C);
        std::vector<SetRuntime> rt(2 * C.nsys);
        to_host(C, rt.data(), C.dRt.p, sizeof(SetRuntime) * rt.size());
        long long nin = 0, nev = 0, nent = 0;
        DevBuf ctr;
        if (pairs_evaluated) ctr.alloc(sizeof(unsigned long long) * 4);
        for (int c = 0; c < 2 * C.nsys; ++c) {
            nin += (long long)rt[c].npairs;  // counted by the pair kernel of the evaluation
            if (C.lastDelta && (c & 1)) continue;  // x + d was not binned (state difference)
            if (pairs_evaluated) {
                zero_on_stream(C, ctr.p, sizeof(unsigned long long) * 4);
                HIPCHK(launch_pairs_dump(C.P, c, C.sets[c].groupCap, nullptr, 0, ctr.as<unsigned long long>(), C.stream));
                unsigned long long h[4];
                to_host(C, h, ctr.p, sizeof(h));
                nev += (long long)h[1];
            }
            if (list_entries) {
                std::vector<int> gc(C.sets[c].groupCap);
                to_host(C, gc.data(), C.dGcount.as<int>() + C.sets[c].groupOff, sizeof(int) * gc.size());
                for (int g = 0; g < rt[c].ngroups; ++g) nent += gc[g];
            }
        }
        if (pairs_in_cutoff) *pairs_in_cutoff = nin;
        if (pairs_evaluated) *pairs_evaluated = nev;
        if (list_entries) *list_entries = nent;
    });
}

int atm_get_list_stats(void* context, int64_t* pairs, int32_t* max_entries, int32_t* capacity) {
    return guarded([&] {
        Context& C = ctx_of(context);
        HIPCHK(hipSetDevice(C.device));
        fetch_results(C);
        std::vector<SetRuntime> rt(2 * C.nsys);
        to_host(C, rt.data(), C.dRt.p, sizeof(SetRuntime) * rt.size());
        for (int c = 0; c < 2 * C.nsys; ++c) {
            if (pairs) pairs[c] = (int64_t)rt[c].npairs;
            if (max_entries) max_entries[c] = rt[c].maxEntries;
            if (capacity) capacity[c] = C.sets[c].entryCap;
        }
    });
}

int atm_set_stream(void* context, void* hip_stream) {
    return guarded([&] {
        Context& C = ctx_of(context);
        HIPCHK(hipSetDevice(C.device));
        HIPCHK(hipStreamSynchronize(C.stream));
        C.stream = hip_stream ? static_cast<hipStream_t>(hip_stream) : C.own;
        C.dropGraphs();
    });
}

int atm_get_tuning(void* context, atm_tuning* out) {
    return guarded([&] {
        Context& C = ctx_of(context);
        if (!out) fail(ATM_ERR_INVALID_ARGUMENT, "null output");
        out->column_scale = C.tuning.colScale;
        out->pair_waves_per_block = C.tuning.pairWaves;
        out->zbins_per_column = C.tuning.zbinsPerColumn;
        out->best_ms = C.tunedMs;
        out->pair_kernel = C.tuning.pairKernel;
        out->list_capacity = C.tuning.listCapacity;
        out->pair_split = C.tuning.pairSplit;
        out->full_state2 = C.tuning.fullState2;
        out->pair_order = C.tuning.pairOrder;
        out->reserved_ = 0;
    });
}

int atm_set_tuning(void* context, const atm_tuning* t) {
    return guarded([&] {
        Context& C = ctx_of(context);
        if (!t) fail(ATM_ERR_INVALID_ARGUMENT, "null tuning");
        if (!(t->column_scale >= 0.5 && t->column_scale <= 2.0)) fail(ATM_ERR_INVALID_ARGUMENT, "column_scale must be in [0.5, 2]");
        if (t->pair_waves_per_block != 1 && t->pair_waves_per_block != 2 && t->pair_waves_per_block != 4 &&
            t->pair_waves_per_block != 8)
            fail(ATM_ERR_INVALID_ARGUMENT, "pair_waves_per_block must be 1, 2, 4 or 8");
        if (t->zbins_per_column < 1 || t->zbins_per_column > 16) fail(ATM_ERR_INVALID_ARGUMENT, "zbins_per_column must be in [1, 16]");
        if (t->pair_kernel < 0 || t->pair_kernel >= kPairKernelCount)
            fail(ATM_ERR_INVALID_ARGUMENT, "pair_kernel must be in [0, " + std::to_string(kPairKernelCount - 1) + "]");
        if (t->pair_kernel != 0 && t->pair_waves_per_block == 8)
            fail(ATM_ERR_INVALID_ARGUMENT, "pair kernels 1-6 take 1, 2 or 4 waves per block");
        if (t->list_capacity < 0) fail(ATM_ERR_INVALID_ARGUMENT, "list_capacity must be >= 0");
        if (t->pair_split != 1 && t->pair_split != 2 && t->pair_split != 4 && t->pair_split != 8)
            fail(ATM_ERR_INVALID_ARGUMENT, "pair_split must be 1, 2, 4 or 8");
        if (t->full_state2 != 0 && t->full_state2 != 1) fail(ATM_ERR_INVALID_ARGUMENT, "full_state2 must be 0 or 1");
        if (t->pair_order < 0 || t->pair_order > 3) fail(ATM_ERR_INVALID_ARGUMENT, "pair_order must be in [0, 3]");
        HIPCHK(hipSetDevice(C.device));
        HIPCHK(hipStreamSynchronize(C.stream));
        const bool relayout = t->column_scale != C.tuning.colScale || t->zbins_per_column != C.tuning.zbinsPerColumn ||
                              (t->pair_kernel == kPairKernelAtomLists) != (C.tuning.pairKernel == kPairKernelAtomLists) ||
                              t->list_capacity != C.tuning.listCapacity;
        C.tuning.colScale = t->column_scale;
        C.tuning.zbinsPerColumn = t->zbins_per_column;
        C.tuning.pairWaves = t->pair_waves_per_block;
        C.tuning.pairKernel = t->pair_kernel;
        C.tuning.pairSplit = t->pair_split;
        C.tuning.fullState2 = t->full_state2;
        C.tuning.pairOrder = t->pair_order;
        C.P.lptMode = t->pair_order;
        if (t->list_capacity != C.tuning.listCapacity) {
            C.entryCapOverride.assign(2 * C.nsys, 0);
            C.alistCapOverride.assign(2 * C.nsys, 0);
        }
        C.tuning.listCapacity = t->list_capacity;
        if (relayout) {
            layout(C);
            alloc_layout(C);
        } else {
            C.dropGraphs();
        }
        C.evaluated = false;
        C.resultsValid = false;
    });
}

int atm_autotune(void* context, int32_t repeats, atm_tuning* out) {
    return guarded([&] {
        Context& C = ctx_of(context);
        require_ready(C);
        HIPCHK(hipSetDevice(C.device));
        const int reps = std::max(1, (int)repeats);
        const double scales[] = {0.8, 0.9, 1.0, 1.15, 1.3};
        const int waves[] = {1, 2, 4};
        // (the atom-list kernel, 5, is left out: it measured slower than the tile kernel on every
        // column scale tried, and its lists need their own buffers)
        // (kernel, group split): the tile kernel with its lists split 1, 2 or 4 ways,
        // the compacted kernel 1 and the filtered kernel
        const int kernels[][2] = {{0, 1}, {0, 2}, {0, 4}, {1, 1}, {kPairKernelFiltered, 1}};
        const int nk = (int)(sizeof(kernels) / sizeof(kernels[0]));
        const int zpcs[] = {2, 4};
        const Tuning orig = C.tuning;
        hipEvent_t e0, e1;
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        const int flags = ATM_EVAL_FORCES | ATM_EVAL_CLEAR_FORCES;
        // Time one configuration (layout already set for its column scale and
        // z-bins): a first evaluation that grows list capacity if needed (and
        // captures the graph), then `n` timed ones; +inf if capacity never fits.
        auto time_config = [&](int n) {
            C.dropGraphs();
            bool ok = true;
            for (int attempt = 0; attempt < 3; ++attempt) {
                eval_batched(C, flags);
                C.evaluated = true;
                C.gradsReady = false;
                C.resultsOnHost = false;
                C.resultsValid = false;
                try {
                    fetch_results(C);
                    ok = true;
                    break;
                } catch (const AtmError& e) {
                    if (e.code != ATM_ERR_CAPACITY) throw;
                    ok = false;
                }
            }
            if (!ok) return 1e300;
            HIPCHK(hipEventRecord(e0, C.stream));
            for (int r = 0; r < n; ++r) eval_batched(C, flags);
            C.resultsValid = false;
            HIPCHK(hipEventRecord(e1, C.stream));
            HIPCHK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e0, e1));
            return (double)ms / n;
        };
        auto set_layout = [&](double sc, int zp) {
            C.tuning.colScale = sc;
            C.tuning.zbinsPerColumn = zp;
            C.entryCapOverride.assign(2 * C.nsys, 0);
            C.alistCapOverride.assign(2 * C.nsys, 0);
            layout(C);
            alloc_layout(C);
        };
        // sweep (at the current group order), then re-time the four fastest with 4x the
        // repeats (a single short timing of near-equal configurations picks by noise)
        std::vector<std::pair<double, Tuning>> timed;
        for (double sc : scales) {
            for (int zp : zpcs) {
                set_layout(sc, zp);
                for (int kv = 0; kv < nk * 3; ++kv) {
                    C.tuning.pairKernel = kernels[kv / 3][0];
                    C.tuning.pairSplit = kernels[kv / 3][1];
                    C.tuning.pairWaves = waves[kv % 3];
                    const double per = time_config(reps);
                    if (per < 1e300) timed.emplace_back(per, C.tuning);
                }
            }
        }
        std::stable_sort(timed.begin(), timed.end(),
                         [](const std::pair<double, Tuning>& a, const std::pair<double, Tuning>& b) { return a.first < b.first; });
        Tuning best = orig;  // kept if no configuration fits the list capacity
        double bestMs = 1e300;
        // (the tile kernel's candidates with each of its four group orders)
        for (size_t i = 0; i < timed.size() && i < 4; ++i) {
            const Tuning t = timed[i].second;
            set_layout(t.colScale, t.zbinsPerColumn);
            for (int order = 0; order <= 3; ++order) {
                if (t.pairKernel != 0 && order != t.pairOrder) continue;
                C.tuning = t;
                C.tuning.pairOrder = t.pairKernel == 0 ? order : t.pairOrder;
                const double per = time_config(4 * reps);
                if (per < bestMs) {
                    bestMs = per;
                    best = C.tuning;
                }
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        C.tuning = best;
        C.tunedMs = bestMs;
        C.entryCapOverride.assign(2 * C.nsys, 0);
        C.alistCapOverride.assign(2 * C.nsys, 0);
        layout(C);
        alloc_layout(C);
        C.evaluated = false;
        C.resultsValid = false;
        if (out) {
            out->column_scale = best.colScale;
            out->pair_waves_per_block = best.pairWaves;
            out->zbins_per_column = best.zbinsPerColumn;
            out->best_ms = bestMs;
            out->pair_kernel = best.pairKernel;
            out->list_capacity = best.listCapacity;
            out->pair_split = best.pairSplit;
            out->full_state2 = best.fullState2;
            out->pair_order = best.pairOrder;
            out->reserved_ = 0;
        }
    });
}

int atm_profile(void* context, int32_t flags, int32_t reps, atm_stage_times* out) {
    return guarded([&] {
        Context& C = ctx_of(context);
        if (!out) fail(ATM_ERR_INVALID_ARGUMENT, "null output");
        if (reps < 1) fail(ATM_ERR_INVALID_ARGUMENT, "reps must be positive");
        require_ready(C);
        HIPCHK(hipSetDevice(C.device));
        flush_params(C);
        hipEvent_t ev[7];
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
        double acc[6] = {0, 0, 0, 0, 0, 0};
        const int key = flags & (ATM_EVAL_FORCES | ATM_EVAL_CLEAR_FORCES | ATM_EVAL_FP32_PAIRS);
        for (int r = 0; r < reps; ++r) {
            enqueue_batched(C, key, ev);
            HIPCHK(hipEventSynchronize(ev[6]));
            for (int k = 0; k < 6; ++k) {
                float ms = 0.f;
                HIPCHK(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
                acc[k] += ms;
            }
        }
        for (auto& e : ev) (void)hipEventDestroy(e);
        C.evaluated = true;
        C.gradsReady = false;
        C.resultsOnHost = false;
        C.resultsValid = false;
        out->build_ms = acc[0] / reps;
        out->pairlist_ms = acc[1] / reps;
        out->pairs_ms = acc[2] / reps;
        out->displaced_ms = acc[3] / reps;
        out->energy_ms = acc[4] / reps;
        out->blend_ms = acc[5] / reps;
        out->total_ms = (acc[0] + acc[1] + acc[2] + acc[3] + acc[4] + acc[5]) / reps;
        out->reps = reps;
    });
}

namespace {

// atm_backward / atm_backward_device: per-atom gradients into host arrays or
// (deviceOut) straight into the caller's device arrays.
void backward_impl(Context& C, const double* grad_energy, double* d_charges, double* d_sigmas, double* d_epsilons,
                   double* d_params, bool deviceOut) {
    if (!grad_energy) fail(ATM_ERR_INVALID_ARGUMENT, "null grad_energy");
    HIPCHK(hipSetDevice(C.device));
    fetch_results(C);
    const int S = C.nsys;
    // per-set weights: dW/dU_state = gW * ((1-sp) for e0's state, sp for the other)
    std::vector<double> wset(2 * S);
    for (int s = 0; s < S; ++s) {
        const atm_result& r = C.results[s];
        const bool fwd = C.params[s].direction > 0.0;
        const double w_e0 = 1.0 - r.hybrid_weight, w_o = r.hybrid_weight;
        wset[2 * s] = grad_energy[s] * (fwd ? w_e0 : w_o);
        wset[2 * s + 1] = grad_energy[s] * (fwd ? w_o : w_e0);
    }
    if (d_charges || d_sigmas || d_epsilons) {
        // after an evaluation with ATM_EVAL_PARAM_GRADS the accumulators hold dU/dp of
        // each set: only the weighted gather remains; otherwise one weighted pass
        bool fused = C.gradsReady;
        if (!fused && C.lastDelta) {
            // after a state-difference evaluation x + d has no lists of its own: evaluate
            // with parameter gradients (both states in full, unit weights), then gather
            // with the weights computed above
            eval_batched(C, ATM_EVAL_PARAM_GRADS);
            C.evaluated = true;
            C.resultsValid = false;
            C.resultsOnHost = false;
            fetch_results(C);  // list overflow: relays out and reports ATM_ERR_CAPACITY
            C.gradsReady = fused = true;
        }
        ensure_grad_buffers(C);
        if (C.dWset.bytes < sizeof(double) * 2 * S) C.dWset.alloc(sizeof(double) * 2 * S);
        HIPCHK(hipMemcpyAsync(C.dWset.p, wset.data(), sizeof(double) * 2 * S, hipMemcpyHostToDevice, C.stream));
        if (!fused)
            HIPCHK(launch_backward(C.P, C.totalSlots, 0, 2 * S, C.maxGroups, C.dWset.as<double>(), C.tuning.pairSplit,
                                   C.stream));
        const size_t nb = sizeof(double) * C.totalAtoms;
        double *oq = d_charges, *os = d_sigmas, *oe = d_epsilons;
        if (!deviceOut) {
            // host outputs: through context buffers (allocated once)
            if (C.dDq.bytes < nb) {
                C.dDq.alloc(nb);
                C.dDsig.alloc(nb);
                C.dDeps.alloc(nb);
            }
            oq = C.dDq.as<double>(); os = C.dDsig.as<double>(); oe = C.dDeps.as<double>();
        }
        HIPCHK(launch_backward_gather(C.P, 0, S, C.maxAtoms, oq, os, oe, fused ? C.dWset.as<double>() : nullptr,
                                      C.stream, fused && C.lastDelta));
        HIPCHK(hipStreamSynchronize(C.stream));
        if (!deviceOut) {
            if (d_charges) to_host(C, d_charges, C.dDq.p, nb);
            if (d_sigmas) to_host(C, d_sigmas, C.dDsig.p, nb);
            if (d_epsilons) to_host(C, d_epsilons, C.dDeps.p, nb);
        }
    }
    if (d_params) {
        for (int s = 0; s < S; ++s) {
            const atm_params& p = C.params[s];
            const atm_result& r = C.results[s];
            const double g = grad_energy[s];
            const double usc = r.perturbation_energy;
            const double x = -p.alpha * (usc - p.u0);
            const double lnee = x > 0.0 ? x + std::log1p(std::exp(-x)) : std::log1p(std::exp(x));
            const double inv_ee = x > 0.0 ? std::exp(-x) / (1.0 + std::exp(-x)) : 1.0 / (1.0 + std::exp(x));
            double dl1 = 0.0, dl2 = usc, da = 0.0, du0 = 0.0;
            if (p.alpha > 0.0) {
                dl1 = -lnee / p.alpha;
                dl2 += lnee / p.alpha;
                du0 = (p.lambda2 - p.lambda1) * (1.0 - inv_ee);
                da = (p.lambda2 - p.lambda1) * (-lnee / (p.alpha * p.alpha) - (usc - p.u0) * (1.0 - inv_ee) / p.alpha);
            }
            d_params[5 * s + 0] = g * dl1;
            d_params[5 * s + 1] = g * dl2;
            d_params[5 * s + 2] = g * da;
            d_params[5 * s + 3] = g * du0;
            d_params[5 * s + 4] = g * 1.0;
        }
    }
}

}  // namespace

int atm_backward(void* context, const double* grad_energy, double* d_charges, double* d_sigmas, double* d_epsilons,
                 double* d_params) {
    return guarded([&] { backward_impl(ctx_of(context), grad_energy, d_charges, d_sigmas, d_epsilons, d_params, false); });
}

int atm_backward_device(void* context, const double* grad_energy, double* d_charges, double* d_sigmas,
                        double* d_epsilons, double* d_params) {
    return guarded([&] { backward_impl(ctx_of(context), grad_energy, d_charges, d_sigmas, d_epsilons, d_params, true); });
}

// ------------------------------------------------------------------ restraints
// Flat-bottom restraints beside the ATM force: python/ATMMetaForceUtils.py:79-138
// (centroid), :707-753 (position).

int atm_restraints_create(int32_t device, void** out) {
    return guarded([&] {
        if (!out) fail(ATM_ERR_INVALID_ARGUMENT, "null output pointer");
        *out = nullptr;
        int ndev = 0;
        HIPCHK(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) fail(ATM_ERR_INVALID_ARGUMENT, "no GPU with index " + std::to_string(device));
        HIPCHK(hipSetDevice(device));
        std::unique_ptr<RestraintSet> R(new RestraintSet());
        R->device = device;
        HIPCHK(hipStreamCreateWithFlags(&R->stream, hipStreamNonBlocking));
        *out = R.release();
    });
}

int atm_restraints_destroy(void* r) {
    return guarded([&] {
        if (!r) return;
        RestraintSet* R = static_cast<RestraintSet*>(r);
        (void)hipSetDevice(R->device);
        (void)hipStreamSynchronize(R->stream);
        delete R;
    });
}

int atm_restraints_set(void* r, int32_t num_atoms, int32_t num_members, const int32_t* member_atoms,
                       const double* member_weights, int32_t num_centroid, const atm_centroid_restraint* centroid,
                       int32_t num_position, const atm_position_restraint* position) {
    return guarded([&] {
        RestraintSet& R = restraints_of(r);
        if (num_atoms < 0 || num_members < 0 || num_centroid < 0 || num_position < 0)
            fail(ATM_ERR_INVALID_ARGUMENT, "negative count");
        if ((num_members && (!member_atoms || !member_weights)) || (num_centroid && !centroid) ||
            (num_position && !position))
            fail(ATM_ERR_INVALID_ARGUMENT, "null array");
        std::vector<int> localOf(num_atoms, -1), used;
        auto local = [&](int a) {
            if (a < 0 || a >= num_atoms) fail(ATM_ERR_INVALID_ARGUMENT, "restraint atom index " + std::to_string(a) + " out of range");
            if (localOf[a] < 0) {
                localOf[a] = (int)used.size();
                used.push_back(a);
            }
            return localOf[a];
        };
        std::vector<int> members(num_members);
        std::vector<double> weights(member_weights, member_weights + num_members);
        for (int m = 0; m < num_members; ++m) {
            if (!(weights[m] > 0.0) || !std::isfinite(weights[m])) fail(ATM_ERR_INVALID_ARGUMENT, "member weights must be positive");
            members[m] = local(member_atoms[m]);
        }
        std::vector<atm_centroid_restraint> cent(centroid, centroid + num_centroid);
        for (const atm_centroid_restraint& c : cent) {
            if (c.group1_count <= 0 || c.group2_count <= 0) fail(ATM_ERR_INVALID_ARGUMENT, "centroid groups must not be empty");
            if (c.group1_first < 0 || c.group1_first + c.group1_count > num_members || c.group2_first < 0 ||
                c.group2_first + c.group2_count > num_members)
                fail(ATM_ERR_INVALID_ARGUMENT, "centroid group outside the member arrays");
            if (!std::isfinite(c.kf) || !std::isfinite(c.tol)) fail(ATM_ERR_INVALID_ARGUMENT, "non-finite restraint parameter");
        }
        std::vector<atm_position_restraint> posr(position, position + num_position);
        for (atm_position_restraint& q : posr) {
            if (!std::isfinite(q.kf) || !std::isfinite(q.tol)) fail(ATM_ERR_INVALID_ARGUMENT, "non-finite restraint parameter");
            q.atom = local(q.atom);
        }
        HIPCHK(hipSetDevice(R.device));
        HIPCHK(hipStreamSynchronize(R.stream));
        R.numAtoms = num_atoms;
        R.used = used;
        R.cent = cent;
        R.posr = posr;
        const size_t U = std::max<size_t>(used.size(), 1);
        R.dPos.alloc(sizeof(double) * 3 * U);
        R.dF.alloc(sizeof(long long) * 3 * U);
        R.dMembers.alloc(sizeof(int) * std::max(num_members, 1));
        R.dWeights.alloc(sizeof(double) * std::max(num_members, 1));
        R.dCent.alloc(sizeof(atm_centroid_restraint) * std::max(num_centroid, 1));
        R.dPosr.alloc(sizeof(atm_position_restraint) * std::max(num_position, 1));
        R.dE.alloc(sizeof(double) * std::max(num_centroid + num_position, 1));
        rs_to_device(R, R.dMembers.p, members.data(), sizeof(int) * num_members);
        rs_to_device(R, R.dWeights.p, weights.data(), sizeof(double) * num_members);
        rs_to_device(R, R.dCent.p, cent.data(), sizeof(atm_centroid_restraint) * num_centroid);
        rs_to_device(R, R.dPosr.p, posr.data(), sizeof(atm_position_restraint) * num_position);
        R.hPos.assign(3 * U, 0.0);
        R.hF.assign(3 * U, 0);
        R.hE.assign(std::max(num_centroid + num_position, 1), 0.0);
    });
}

int atm_restraints_eval(void* r, const double* positions, const double box[3], double* forces, double* energy) {
    return guarded([&] {
        RestraintSet& R = restraints_of(r);
        if (!positions || !box || !energy) fail(ATM_ERR_INVALID_ARGUMENT, "null argument");
        const int nc = (int)R.cent.size(), np = (int)R.posr.size();
        *energy = 0.0;
        if (nc + np == 0) return;
        for (int k = 0; k < 3; ++k)
            if (!(box[k] > 0.0)) fail(ATM_ERR_INVALID_ARGUMENT, "box edges must be positive");
        const size_t U = R.used.size();
        for (size_t u = 0; u < U; ++u)
            for (int k = 0; k < 3; ++k) R.hPos[3 * u + k] = positions[3ll * R.used[u] + k];
        HIPCHK(hipSetDevice(R.device));
        HIPCHK(hipMemcpyAsync(R.dPos.p, R.hPos.data(), sizeof(double) * 3 * U, hipMemcpyHostToDevice, R.stream));
        HIPCHK(hipMemsetAsync(R.dF.p, 0, sizeof(long long) * 3 * U, R.stream));
        RestraintPtrs P;
        P.pos = R.dPos.as<double>();
        P.fsum = R.dF.as<unsigned long long>();
        P.members = R.dMembers.as<int>();
        P.weights = R.dWeights.as<double>();
        P.cent = R.dCent.as<atm_centroid_restraint>();
        P.posr = R.dPosr.as<atm_position_restraint>();
        P.energy = R.dE.as<double>();
        P.ncent = nc;
        P.npos = np;
        for (int k = 0; k < 3; ++k) P.box[k] = box[k];
        HIPCHK(launch_restraints(P, R.stream));
        HIPCHK(hipMemcpyAsync(R.hE.data(), R.dE.p, sizeof(double) * (nc + np), hipMemcpyDeviceToHost, R.stream));
        HIPCHK(hipMemcpyAsync(R.hF.data(), R.dF.p, sizeof(long long) * 3 * U, hipMemcpyDeviceToHost, R.stream));
        HIPCHK(hipStreamSynchronize(R.stream));
        double e = 0.0;
        for (int k = 0; k < nc + np; ++k) e += R.hE[k];  // fixed order
        *energy = e;
        if (forces)
            for (size_t u = 0; u < U; ++u)
                for (int k = 0; k < 3; ++k) forces[3ll * R.used[u] + k] += (double)R.hF[3 * u + k] / ATM_FORCE_SCALE;
    });
}

// ------------------------------------------------------------------ batched Langevin dynamics
// The MD loop of an ATM run (example/abfe/abfe.py:101-102, 149-160) for every
// system of the context at once.

namespace {
const double kBoltzmann = 0.00831446261815324;  // kJ/mol/K

MDPtrs md_ptrs(Context& C) {
    MDPtrs M;
    M.x = C.dMx.as<double>();
    M.v = C.dMv.as<double>();
    M.invm = C.dMinvm.as<double>();
    M.posq = C.dPosq.as<float4>();
    return M;
}

hipError_t launch_md_init_posq(Context& C) { return launch_md_posq(C.P, md_ptrs(C), C.totalAtoms, C.stream); }

// One batched evaluation with forces, re-run with doubled list capacity until
// no list overflows (positions unchanged).
void md_evaluate(Context& C) {
    for (int attempt = 0;; ++attempt) {
        eval_batched(C, ATM_EVAL_FORCES | ATM_EVAL_CLEAR_FORCES);
        C.evaluated = true;
        C.gradsReady = false;
        C.resultsOnHost = false;
        C.resultsValid = false;
        try {
            fetch_results(C);
            return;
        } catch (const AtmError& e) {
            if (e.code != ATM_ERR_CAPACITY || attempt >= 4) throw;
        }
    }
}
}  // namespace

int atm_md_init(void* context, const double* masses, const double* positions, double temperature, uint64_t seed) {
    return guarded([&] {
        Context& C = ctx_of(context);
        require_ready(C);
        if (!masses || !positions) fail(ATM_ERR_INVALID_ARGUMENT, "null array");
        if (!(temperature >= 0.0)) fail(ATM_ERR_INVALID_ARGUMENT, "temperature must be >= 0");
        HIPCHK(hipSetDevice(C.device));
        const long long n = C.totalAtoms;
        std::vector<double> invm(n);
        for (long long i = 0; i < n; ++i) {
            if (!(masses[i] >= 0.0) || !std::isfinite(masses[i])) fail(ATM_ERR_INVALID_ARGUMENT, "masses must be >= 0");
            invm[i] = masses[i] > 0.0 ? 1.0 / masses[i] : 0.0;
        }
        for (long long i = 0; i < 3 * n; ++i)
            if (!std::isfinite(positions[i])) fail(ATM_ERR_INVALID_ARGUMENT, "non-finite position");
        C.masses.assign(masses, masses + n);
        C.dMx.alloc(sizeof(double) * 3 * n);
        C.dMv.alloc(sizeof(double) * 3 * n);
        C.dMxSnap.alloc(sizeof(double) * 3 * n);
        C.dMvSnap.alloc(sizeof(double) * 3 * n);
        C.dMinvm.alloc(sizeof(double) * n);
        to_device(C, C.dMx.p, positions, sizeof(double) * 3 * n);
        to_device(C, C.dMinvm.p, invm.data(), sizeof(double) * n);
        for (long long i = 0; i < n; ++i) {
            C.hPosq[i].x = (float)positions[3 * i];
            C.hPosq[i].y = (float)positions[3 * i + 1];
            C.hPosq[i].z = (float)positions[3 * i + 2];
        }
        C.mdSeed = seed;
        C.mdStep = 0;
        HIPCHK(launch_md_init(C.P, md_ptrs(C), n, kBoltzmann * temperature, seed, C.stream));
        md_evaluate(C);
        C.mdReady = true;
    });
}

int atm_md_step(void* context, int32_t steps, double dt, double friction, double temperature) {
    return guarded([&] {
        Context& C = ctx_of(context);
        if (!C.mdReady) fail(ATM_ERR_STATE, "atm_md_init has not been called");
        if (steps < 0 || !(dt > 0.0) || !(friction >= 0.0) || !(temperature >= 0.0))
            fail(ATM_ERR_INVALID_ARGUMENT, "steps >= 0, dt > 0, friction >= 0 and temperature >= 0 are required");
        HIPCHK(hipSetDevice(C.device));
        const double a = std::exp(-friction * dt), kT = kBoltzmann * temperature;
        const size_t bytes = sizeof(double) * 3 * C.totalAtoms;
        const int chunk = 50;  // steps between overflow checks (and snapshots for a replay)
        for (int done = 0; done < steps;) {
            const int n = std::min(chunk, steps - done);
            HIPCHK(hipMemcpyAsync(C.dMxSnap.p, C.dMx.p, bytes, hipMemcpyDeviceToDevice, C.stream));
            HIPCHK(hipMemcpyAsync(C.dMvSnap.p, C.dMv.p, bytes, hipMemcpyDeviceToDevice, C.stream));
            for (int k = 0; k < n; ++k) {
                HIPCHK(launch_md_step(C.P, md_ptrs(C), C.nsys, C.maxAtoms, dt, a, kT, C.mdSeed, C.mdStep + k, C.stream));
                eval_batched(C, ATM_EVAL_FORCES | ATM_EVAL_CLEAR_FORCES);
            }
            C.evaluated = true;
            C.gradsReady = false;
            C.resultsOnHost = false;
            C.resultsValid = false;
            try {
                fetch_results(C);
            } catch (const AtmError& e) {
                if (e.code != ATM_ERR_CAPACITY) throw;
                // a list overflowed somewhere in these steps: capacity has been doubled;
                // restore the snapshot and replay them (same noise counters, same trajectory)
                HIPCHK(hipMemcpyAsync(C.dMx.p, C.dMxSnap.p, bytes, hipMemcpyDeviceToDevice, C.stream));
                HIPCHK(hipMemcpyAsync(C.dMv.p, C.dMvSnap.p, bytes, hipMemcpyDeviceToDevice, C.stream));
                HIPCHK(launch_md_init_posq(C));
                md_evaluate(C);
                continue;
            }
            C.mdStep += n;
            done += n;
        }
    });
}

int atm_md_get(void* context, int32_t sys, double* positions, double* velocities, double* kinetic_energy) {
    return guarded([&] {
        Context& C = ctx_of(context);
        check_sys(C, sys);
        if (!C.mdReady) fail(ATM_ERR_STATE, "atm_md_init has not been called");
        HIPCHK(hipSetDevice(C.device));
        const long long off = C.atomOff[sys], n = C.sys[sys].num_atoms;
        std::vector<double> v(3 * n);
        if (positions) to_host(C, positions, C.dMx.as<double>() + 3 * off, sizeof(double) * 3 * n);
        if (velocities || kinetic_energy) to_host(C, v.data(), C.dMv.as<double>() + 3 * off, sizeof(double) * 3 * n);
        if (velocities) std::copy(v.begin(), v.end(), velocities);
        if (kinetic_energy) {
            double ke = 0.0;
            for (long long i = 0; i < n; ++i)
                ke += 0.5 * C.masses[off + i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
            *kinetic_energy = ke;
        }
    });
}

}  // extern "C"
