// atm_eval_noalloc: one system's ATM evaluation with no device allocation.
//
// The caller passes its inputs (positions, particle parameters, displacements)
// and outputs (forces, result record, status word) as device arrays.  Every
// other buffer the evaluation's kernels use -- the binned and sorted atoms,
// cell and column tables, cluster boxes, group lists, 64-bit fixed-point force
// sums, FP64 side sums, energies and the set descriptors -- is carved out of
// `scratch`, a file-scope device array of kNoallocScratchBytes that each GPU
// holds until the process ends.  The kernels are those of atm_eval (one
// system, both coordinate sets, default launch configuration), so the results
// equal atm_eval's on the same inputs.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstddef>
#include <mutex>
#include <string>
#include "atm_internal.h"
#include "atm_kernels.h"
#include "atm_capi_util.h"
#include "atm_layout.h"
#include "../../include/atm_mi355x.h"

namespace atm {
constexpr size_t kNoallocScratchBytes = (size_t)ATM_NOALLOC_SCRATCH_MIB << 20;
}  // namespace atm

// Scratch of atm_eval_noalloc.  In host code this name is not a device
// address: kernels get the address hipGetSymbolAddress returns for the current
// device (noalloc_scratch below), never the symbol itself.
static __device__ __align__(256) unsigned char scratch[atm::kNoallocScratchBytes];

using namespace atm;

namespace {

static_assert(sizeof(SysResult) == sizeof(atm_result), "SysResult mirrors atm_result");
static_assert(offsetof(atm_result, energy) == offsetof(SysResult, W) &&
                  offsetof(atm_result, perturbation_energy) == offsetof(SysResult, usc) &&
                  offsetof(atm_result, state1_energy) == offsetof(SysResult, U1) &&
                  offsetof(atm_result, state2_energy) == offsetof(SysResult, U2) &&
                  offsetof(atm_result, raw_perturbation) == offsetof(SysResult, u) &&
                  offsetof(atm_result, softcore_derivative) == offsetof(SysResult, fp) &&
                  offsetof(atm_result, softplus_derivative) == offsetof(SysResult, bfp) &&
                  offsetof(atm_result, hybrid_weight) == offsetof(SysResult, sp),
              "SysResult mirrors atm_result");

constexpr int kMaxDevices = 64;
std::mutex gScratchMu;
void* gScratch[kMaxDevices] = {};

// The device address of `scratch` on the current device `dev`, looked up once
// per device (each GPU holds its own copy of the array).
void* noalloc_scratch(int dev) {
    if (dev < 0 || dev >= kMaxDevices) fail(ATM_ERR_INVALID_ARGUMENT, "device index out of range");
    std::lock_guard<std::mutex> lock(gScratchMu);
    if (!gScratch[dev]) {
        void* p = nullptr;
        HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) fail(ATM_ERR_DEVICE, "no device address for the no-allocation scratch");
        gScratch[dev] = p;
    }
    return gScratch[dev];
}

// Sequential 256-byte-aligned carving of the scratch; `need` counts every byte.
struct Carve {
    char* base;
    size_t need = 0;
    template <typename T>
    T* take(long long n) {
        const size_t off = (need + 255) & ~(size_t)255;
        need = off + sizeof(T) * (size_t)std::max<long long>(n, 1);
        return reinterpret_cast<T*>(base + off);
    }
};

constexpr int kStatusBadInput = 4;  // rt[0].overflow bit: an input failed validation

// Per-call set-up, one block: descriptors, parameters and cleared runtime
// records of the two coordinate sets.
__global__ void k_noalloc_init(DevPtrs P, SetDesc d0, SetDesc d1, AtmParamsDev prm) {
    const int i = threadIdx.x;
    if (i < 2) {
        const_cast<SetDesc*>(P.desc)[i] = i == 0 ? d0 : d1;
        SetRuntime r{};
        r.big = 1;  // the FP64 side sums hold another call's values: have k_gather_bbox clear them
        P.rt[i] = r;
    }
    if (i == 0) *const_cast<AtmParamsDev*>(P.params) = prm;
}

// The per-atom inputs in the evaluation's formats (as atm_set_particles and
// atm_set_positions convert them).
// An invalid input sets a flag that becomes the status.
__global__ void k_noalloc_atoms(DevPtrs P, const double* x, const double* q, const double* sig, const double* eps,
                                const double* disp, int n, double sqk) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
    const double qi = q[i], si = sig[i], ei = eps[i];
    const double dx = disp[3 * i], dy = disp[3 * i + 1], dz = disp[3 * i + 2];
    const bool ok = isfinite(xi) && isfinite(yi) && isfinite(zi) && isfinite(qi) && si >= 0.0 && ei >= 0.0 &&
                    isfinite(si) && isfinite(ei) && isfinite(dx) && isfinite(dy) && isfinite(dz);
    if (!ok) atomicOr(&P.rt[0].overflow, kStatusBadInput);
    const double se = ok ? ei : 0.0;
    const_cast<float4*>(P.posq)[i] = make_float4(ok ? (float)xi : 0.f, ok ? (float)yi : 0.f, ok ? (float)zi : 0.f,
                                                 ok ? (float)(qi * sqk) : 0.f);
    const_cast<float2*>(P.ljp)[i] = make_float2(ok ? (float)(0.5 * si) : 0.f, (float)(2.0 * sqrt(se)));
    const_cast<double4*>(P.prm64)[i] = make_double4(ok ? 0.5 * si : 0.0, 2.0 * sqrt(se), ok ? qi * sqk : 0.0, 0.0);
    const_cast<float4*>(P.displ)[i] =
        ok ? make_float4((float)dx, (float)dy, (float)dz, 0.f) : make_float4(0.f, 0.f, 0.f, 0.f);
}

// Status word, and the system's force added to (or written into) the caller's
// atom-major array.  A failed evaluation leaves the forces untouched and sets
// every result field to NaN.
__global__ void k_noalloc_finish(DevPtrs P, int n, int forces, int clear, double* f, int32_t* status) {
    const int ov = P.rt[0].overflow | P.rt[1].overflow;
    const bool bad = (ov & kStatusBadInput) != 0;
    const bool full = (ov & 3) != 0;
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        *status = bad ? ATM_ERR_INVALID_ARGUMENT : (full ? ATM_ERR_CAPACITY : ATM_OK);
        if (bad || full) {
            const double nan = __builtin_nan("");
            *P.res = SysResult{nan, nan, nan, nan, nan, nan, nan, nan};
        }
    }
    if (!forces || bad || full || i >= n) return;
    for (int k = 0; k < 3; ++k) {
        const double v = P.outF[(long long)k * n + i];
        f[3 * i + k] = clear ? v : f[3 * i + k] + v;
    }
}

}  // namespace

extern "C" int atm_eval_noalloc(const atm_system_desc* system, const atm_params* params, const double* positions,
                                const double* charges, const double* sigmas, const double* epsilons,
                                const double* displacements, int32_t flags, double* forces, atm_result* result,
                                int32_t* status, void* hip_stream) {
    return guarded([&] {
        if (!system || !params) fail(ATM_ERR_INVALID_ARGUMENT, "null system description or parameters");
        if (!positions || !charges || !sigmas || !epsilons || !displacements || !result || !status)
            fail(ATM_ERR_INVALID_ARGUMENT, "null device array");
        const bool withForces = (flags & ATM_EVAL_FORCES) != 0;
        if (withForces && !forces) fail(ATM_ERR_INVALID_ARGUMENT, "null force array with ATM_EVAL_FORCES");
        if (flags & ~(ATM_EVAL_FORCES | ATM_EVAL_CLEAR_FORCES | ATM_EVAL_FP32_PAIRS))
            fail(ATM_ERR_INVALID_ARGUMENT, "atm_eval_noalloc takes ATM_EVAL_FORCES, ATM_EVAL_CLEAR_FORCES and "
                                           "ATM_EVAL_FP32_PAIRS only");
        validate_desc(*system, 0);
        const int n = system->num_atoms;
        if (n > ATM_NOALLOC_MAX_ATOMS)
            fail(ATM_ERR_UNSUPPORTED, "atm_eval_noalloc takes at most " + std::to_string(ATM_NOALLOC_MAX_ATOMS) +
                                          " atoms (" + std::to_string(n) + " given)");
        int dev = 0;
        HIPCHK(hipGetDevice(&dev));
        void* base = noalloc_scratch(dev);
        const hipStream_t st = static_cast<hipStream_t>(hip_stream);

        // the two coordinate sets, laid out as atm_create lays out a one-system batch
        const Tuning tu;
        SetDesc D[2];
        long long cellOff = 0, colOff = 0, slotOff = 0, clOff = 0, gOff = 0, eOff = 0;
        for (int stt = 0; stt < 2; ++stt) {
            SetDesc d = make_set(*system, tu, 0, stt);
            d.fixMax = fixed_point_limit(d);
            d.atomOff = 0;
            d.setAtomOff = stt * n;
            d.cellOff = (int)cellOff;
            d.colOff = (int)colOff;
            d.slotOff = (int)slotOff;
            d.clusterOff = (int)clOff;
            d.groupOff = (int)gOff;
            d.entryOff = eOff;
            d.alistCap = 0;
            d.alistOff = 0;
            cellOff += (long long)d.nx * d.ny * d.nz;
            colOff += (long long)d.nx * d.ny;
            slotOff += d.slotCap;
            clOff += d.clusterCap;
            gOff += d.groupCap;
            eOff += (long long)d.groupCap * d.entryCap;
            D[stt] = d;
        }
        Carve cv{static_cast<char*>(base)};
        DevPtrs P{};
        P.desc = cv.take<SetDesc>(2);
        P.rt = cv.take<SetRuntime>(2);
        P.posq = cv.take<float4>(n);
        P.displ = cv.take<float4>(n);
        P.ljp = cv.take<float2>(n);
        P.prm64 = cv.take<double4>(n);
        P.params = cv.take<AtmParamsDev>(1);
        P.res = reinterpret_cast<SysResult*>(result);
        P.outF = cv.take<double>(3ll * n);
        P.bpos = cv.take<float4>(2ll * n);
        P.bcell = cv.take<int>(2ll * n);
        P.brank = cv.take<int>(2ll * n);
        P.slotOf = cv.take<int>(2ll * n);
        P.cellCount = cv.take<int>(cellOff);
        P.cellStart = cv.take<int>(cellOff);
        P.colStart = cv.take<int>(colOff);
        P.colCount = cv.take<int>(colOff);
        P.spos = cv.take<float4>(slotOff);
        P.sprm = cv.take<float2>(slotOff);
        P.sprm64 = cv.take<double4>(slotOff);
        P.sidx = cv.take<int>(slotOff);
        P.sslot = cv.take<int2>(slotOff);
        P.sf = cv.take<long long>(3 * slotOff);
        P.sfn = slotOff;
        P.sd = cv.take<double>(3 * slotOff);
        P.sdMask = cv.take<unsigned long long>(slotOff / 64 + 1);
        P.bbc = cv.take<float4>(clOff);
        P.bbh = cv.take<float4>(clOff);
        P.ginfo = cv.take<int4>(gOff);
        P.gcount = cv.take<int>(gOff);
        P.genergy = cv.take<double>(8 * gOff);
        P.entries = cv.take<uint32_t>(eOff);
        P.setEnergy = cv.take<double>(2);
        P.alist = nullptr;
        P.acount = nullptr;
        P.gq = P.gsig = P.geps = nullptr;
        if (cv.need > kNoallocScratchBytes)
            fail(ATM_ERR_UNSUPPORTED, "system needs " + std::to_string(cv.need >> 20) + " MiB of scratch; the "
                                      "no-allocation entry point has " + std::to_string(ATM_NOALLOC_SCRATCH_MIB) + " MiB");

        AtmParamsDev prm{params->lambda1, params->lambda2, params->alpha, params->u0, params->w0,
                         params->umax, params->ubcore, params->acore, params->direction};
        const int maxCells = std::max(D[0].nx * D[0].ny * D[0].nz, D[1].nx * D[1].ny * D[1].nz);
        const int maxSlots = std::max(D[0].slotCap, D[1].slotCap);
        const int maxGroups = std::max(D[0].groupCap, D[1].groupCap);
        hipLaunchKernelGGL(k_noalloc_init, dim3(1), dim3(64), 0, st, P, D[0], D[1], prm);
        HIPCHK(hipGetLastError());
        hipLaunchKernelGGL(k_noalloc_atoms, dim3((n + 255) / 256), dim3(256), 0, st, P, positions, charges, sigmas,
                           epsilons, displacements, n, std::sqrt(ATM_ONE_4PI_EPS0));
        HIPCHK(hipGetLastError());
        HIPCHK(launch_zero_counts(P, 0, 2, maxCells, nullptr, 0, st));
        HIPCHK(launch_build(P, 0, 2, n, maxCells, maxSlots, st));
        HIPCHK(launch_pairlist(P, 0, 2, maxGroups, 4, st));
        HIPCHK(launch_pairs(P, 0, 2, maxGroups, tu.pairWaves, withForces, (flags & ATM_EVAL_FP32_PAIRS) == 0,
                            tu.pairKernel, tu.pairSplit, st));
        HIPCHK(launch_energy(P, 0, 1, pair_energy_stride(tu.pairKernel, (flags & ATM_EVAL_FP32_PAIRS) == 0, tu.pairSplit),
                             st));
        if (withForces) HIPCHK(launch_blend(P, 0, 1, n, 0, 0.0, 0.0, true, st));
        hipLaunchKernelGGL(k_noalloc_finish, dim3((std::max(n, 1) + 255) / 256), dim3(256), 0, st, P, n,
                           withForces ? 1 : 0, (flags & ATM_EVAL_CLEAR_FORCES) ? 1 : 0, forces, status);
        HIPCHK(hipGetLastError());
    });
}
