// gfx950 kernels of the PME variable group: OpenMM's NonbondedForce with
// particle-mesh Ewald evaluated on x and on x + d, then the ATM combine.
//
// Per coordinate set (state s = 0: x, 1: x + d):
//   k_pme_fill3      clear the cell counts, charge grid and force sums
//   k_pme_bin        (state 1: form x + d) wrap into the box, cell of edge
//                    >= rc, rank in the cell
//   k_pme_scan       cell starts (one block)
//   k_pme_scatter    atom indices into cell order
//   k_pme_cellsort   each cell ordered by atom index (deterministic sums), and
//                    positions / parameters / exclusion keys gathered into
//                    cell order
//   k_pme_direct     erfc(alpha r)/r Coulomb + LJ over minimum-image pairs with
//                    r < rc that are not exceptions; one wave per (64 atoms of
//                    a cell, the cell or one of its 13 forward neighbours, a
//                    quarter of that cell's atoms): each pair once
//   k_pme_spread_tiles  order-5 B-spline charge spreading per cell into an LDS
//                    tile of the grid, 64-bit fixed point (k_pme_spread: the
//                    same straight into the grid, when a tile does not fit)
//   k_pme_dft_lines  one axis of the 3-D complex DFT, lines staged in LDS (run
//                    3x forward, the first reading the fixed-point charge grid,
//                    3x back; k_pme_q2c + k_pme_dft for lines too long for LDS)
//   k_pme_convolve   influence function, reciprocal energy, convolution
//   k_pme_gather     reciprocal forces from the potential grid
//   k_pme_exceptions exception pairs: own Coulomb + LJ, minus erf(alpha r)/r
// and once per evaluation:
//   k_pme_finish     fixed-order sums of every energy slot, U1, U2, combine
//   k_pme_blend      F += (1 - sp) F1 + sp F2 (direction-aware), FP64 out
//
// Everything is FP64: positions, pair function, grid, transforms.  Forces are
// summed in 64-bit fixed point (2^-24 kJ/mol/nm) with integer atomics, energies
// in per-block slots reduced in a fixed order, so results are bit-reproducible.
//
// The model restated is OpenMM's PME NonbondedForce as the reference's
// end-to-end test builds it (python/tests/test_abfe.py:42-43, through
// AmberPrmtopFile.createSystem), cloned into the ATM inner contexts at
// openmmapi/src/ATMMetaForceImpl.cpp:58-63 and combined by
// platforms/reference/src/ReferenceATMMetaForceKernels.cpp:56-110.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "atm_pme.h"
#include "atm_math.h"

namespace atm {

namespace {

constexpr double kFix = 16777216.0;          // 2^24 force units per kJ/mol/nm
constexpr double kGridFix = 68719476736.0;   // 2^36 charge-grid units per e
constexpr int kOrder = 5;

__device__ __forceinline__ void add_fixed(unsigned long long* p, double f) {
    atomicAdd(p, (unsigned long long)__double2ll_rn(f * kFix));
}

// Grid index v in [-K, 2K) wrapped into [0, K) (no integer division).
__device__ __forceinline__ int wrap_index(int v, int K) {
    v = v < 0 ? v + K : v;
    return v >= K ? v - K : v;
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block sum (fixed order) of one value per thread; valid in thread 0.
template <int NT>
__device__ double block_sum(double v, double* red) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < NT / 64; ++k) s += red[k];
    __syncthreads();
    return s;
}

__device__ __forceinline__ int wrap_cell(double x, double L, double invL, int m, double* w) {
    double v = x - L * floor(x * invL);
    if (v < 0.0) v += L;
    if (v >= L) v -= L;
    *w = v;
    int c = (int)(v * invL * m);
    return c < 0 ? 0 : (c >= m ? m - 1 : c);
}

// Order-5 cardinal B-spline weights M5(w + j) and derivatives M5'(w + j),
// j = 0..4, for w in [0, 1) (two-term recursion M_k(t) = t/(k-1) M_{k-1}(t)
// + (k-t)/(k-1) M_{k-1}(t-1), from M_2(w) = w, M_2(w+1) = 1 - w).
__device__ __forceinline__ void bspline5(double w, double th[kOrder], double dth[kOrder]) {
    double a[kOrder] = {w, 1.0 - w, 0.0, 0.0, 0.0};
    for (int k = 3; k <= kOrder; ++k) {
        double nw[kOrder];
        const double inv = 1.0 / (k - 1);
        for (int j = 0; j < kOrder; ++j) {
            const double am1 = j > 0 ? a[j - 1] : 0.0;
            nw[j] = j < k ? ((w + j) * a[j] + (k - w - j) * am1) * inv : 0.0;
        }
        if (k == kOrder)
            for (int j = 0; j < kOrder; ++j) dth[j] = a[j] - (j > 0 ? a[j - 1] : 0.0);
        for (int j = 0; j < kOrder; ++j) a[j] = nw[j];
    }
    for (int j = 0; j < kOrder; ++j) th[j] = a[j];
}

}  // namespace

// Cell of every atom and its rank among the cell's atoms (any order: the cell
// sort fixes it).  Ranks come from a block histogram in LDS and one global
// atomic per (block, occupied cell), so the few hundred atoms of a cell do not
// serialise on one counter; grids with more than kPmeBinLds cells use global
// atomics directly.
// d: when not null, the positions binned are x + d (the displaced state), also
// written to P.x2 for the kernels after this one.
__global__ void __launch_bounds__(256) k_pme_bin(PmeDesc D, PmePtrs P, const double* x, const double* d) {
    __shared__ int hist[kPmeBinLds];
    __shared__ int base[kPmeBinLds];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool lds = D.ncell <= kPmeBinLds;
    if (lds) {
        for (int c = threadIdx.x; c < D.ncell; c += blockDim.x) hist[c] = 0;
        __syncthreads();
    }
    int c = -1, r = 0;
    if (i < D.n) {
        double w[3], xi[3] = {x[3 * i], x[3 * i + 1], x[3 * i + 2]};
        if (d) {
            for (int k = 0; k < 3; ++k) {
                xi[k] += d[3 * i + k];
                P.x2[3 * i + k] = xi[k];
            }
        }
        const int cx = wrap_cell(xi[0], D.L[0], D.invL[0], D.m[0], &w[0]);
        const int cy = wrap_cell(xi[1], D.L[1], D.invL[1], D.m[1], &w[1]);
        const int cz = wrap_cell(xi[2], D.L[2], D.invL[2], D.m[2], &w[2]);
        c = (cx * D.m[1] + cy) * D.m[2] + cz;
        P.cellOf[i] = c;
        P.wpos[i] = make_double4(w[0], w[1], w[2], 0.0);
        r = lds ? atomicAdd(&hist[c], 1) : atomicAdd(&P.cellCount[c], 1);
    }
    if (lds) {
        __syncthreads();
        for (int k = threadIdx.x; k < D.ncell; k += blockDim.x)
            base[k] = hist[k] ? atomicAdd(&P.cellCount[k], hist[k]) : 0;
        __syncthreads();
        if (c >= 0) r += base[c];
    }
    if (i < D.n) P.rankOf[i] = r;
}

__global__ void __launch_bounds__(1024) k_pme_scan(PmeDesc D, PmePtrs P) {
    __shared__ int part[1024];
    const int t = threadIdx.x;
    const int per = (D.ncell + 1023) / 1024;
    const int b0 = t * per, b1 = min(b0 + per, D.ncell);
    int s = 0, mx = 0;
    for (int c = b0; c < b1; ++c) { s += P.cellCount[c]; mx = max(mx, P.cellCount[c]); }
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int c = b0; c < b1; ++c) { P.cellStart[c] = run; run += P.cellCount[c]; }
    if (t == 1023) P.cellStart[D.ncell] = part[1023];
    if (mx > D.cellCap) atomicOr(P.overflow, 1);
}

__global__ void k_pme_scatter(PmeDesc D, PmePtrs P) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.n) return;
    P.cellAtoms[P.cellStart[P.cellOf[i]] + P.rankOf[i]] = i;
}

// One block per cell: order its atoms by index and gather them into cell order.
__global__ void __launch_bounds__(256) k_pme_cellsort(PmeDesc D, PmePtrs P) {
    __shared__ int ids[kPmeCellMax];
    const int c = blockIdx.x;
    const int s0 = P.cellStart[c], cnt = min(P.cellStart[c + 1] - s0, kPmeCellMax);
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) ids[k] = P.cellAtoms[s0 + k];
    __syncthreads();
    for (int k = threadIdx.x; k < cnt; k += blockDim.x) {
        const int a = ids[k];
        int r = 0;
        for (int l = 0; l < cnt; ++l) r += ids[l] < a;
        const int slot = s0 + r;
        const double4 w = P.wpos[a];
        const double2 pr = P.prm[a];
        P.cpos[slot] = make_double4(w.x, w.y, w.z, pr.x);
        P.cprm[slot] = make_double2(P.prm2[a].x, P.prm2[a].y);
        const int4 ex = P.exkey[a];
        P.cex[slot] = make_int4(ex.x, ex.y, ex.z, a);
    }
}

// Half shell: blockIdx.x = 0 is cell c = blockIdx.y itself (pairs with the
// j slot above the i slot), 1..13 its forward neighbour cells (all pairs);
// blockIdx.z = chunk of 64 atoms of c (one per lane).  Every pair is visited
// once: +f on i, -f on j, the whole pair energy.
//  test:  the wave walks the neighbour cell's atoms staged in LDS; each lane
//         tests its atom against the common j (FP64 r^2, the image shift known
//         from the cell offset, exclusion bitmap for atoms of one exclusion
//         group) and the in-range pairs are appended to a ring in LDS
//         (ballot + mbcnt, so the order is fixed);
//  dense: whenever 64 pairs are pending, each lane evaluates one of them
//         (erfc Coulomb + LJ in FP64); forces go to per-atom LDS accumulators
//         of the chunk (i) and of the staged batch (j) in 64-bit fixed point
//         (integer atomics: order-independent), flushed to global memory.
__global__ void __launch_bounds__(64) k_pme_direct(PmeDesc D, PmePtrs P, unsigned long long* F, double* eslot) {
    __shared__ double4 sI[64];
    __shared__ double2 sIp[64];
    __shared__ unsigned long long sF[3][64];
    __shared__ uint32_t sQ[128];
    __shared__ double4 sJ[kPmeJBatch];  // the neighbour cell's atoms of the current batch
    __shared__ double2 sJp[kPmeJBatch];
    __shared__ int4 sJx[kPmeJBatch];    // exclusion group, rank, -, atom
    __shared__ unsigned long long sJF[3][kPmeJBatch];
    const int lane = threadIdx.x;
    // x: neighbour offset (fastest) and part of the neighbour cell's atoms; y: cell; z: 64-atom chunk
    const int nb = blockIdx.x % kPmeHalfShell, part = blockIdx.x / kPmeHalfShell;
    const int c = blockIdx.y, chunk = blockIdx.z;
    const int cz = c % D.m[2], cy = (c / D.m[2]) % D.m[1], cx = c / (D.m[1] * D.m[2]);
    // forward offsets: (a, b, e) lexicographically above (0, 0, 0); nb = 0 is the cell itself
    int ox = 0, oy = 0, oz = 0;
    if (nb > 0) {
        const int f = nb + 13;  // 14..26 in the 3x3x3 numbering (a + 1) 9 + (b + 1) 3 + (e + 1)
        ox = f / 9 - 1; oy = (f / 3) % 3 - 1; oz = f % 3 - 1;
    }
    int nx = cx + ox, ny = cy + oy, nz = cz + oz;
    double sx = 0.0, sy = 0.0, sz = 0.0;  // image of the neighbour cell: x_j + s
    if (nx < 0) { nx += D.m[0]; sx = -D.L[0]; } else if (nx >= D.m[0]) { nx -= D.m[0]; sx = D.L[0]; }
    if (ny < 0) { ny += D.m[1]; sy = -D.L[1]; } else if (ny >= D.m[1]) { ny -= D.m[1]; sy = D.L[1]; }
    if (nz < 0) { nz += D.m[2]; sz = -D.L[2]; } else if (nz >= D.m[2]) { nz -= D.m[2]; sz = D.L[2]; }
    const int cn = (nx * D.m[1] + ny) * D.m[2] + nz;
    const int s0 = P.cellStart[c], s1 = P.cellStart[c + 1];
    const int blockSlot = (chunk * gridDim.y + c) * gridDim.x + blockIdx.x;
    if (s0 + chunk * 64 >= s1) {  // no atom in this chunk: the whole block leaves
        if (lane == 0) eslot[blockSlot] = 0.0;
        return;
    }
    const int si = s0 + chunk * 64 + lane;
    const bool active = si < s1;
    double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
    int4 xi = make_int4(-1, 0, 0, -1);
    uint32_t word0 = 0u;  // first word of this atom's exclusion row (groups of up to 32 atoms)
    if (active) {
        pi = P.cpos[si];
        xi = P.cex[si];
        sIp[lane] = P.cprm[si];
        word0 = P.exbits[xi.z];
    }
    pi.x -= sx; pi.y -= sy; pi.z -= sz;  // (x_i - s) - x_j = x_i - (x_j + s)
    sI[lane] = pi;
    sF[0][lane] = 0ull; sF[1][lane] = 0ull; sF[2][lane] = 0ull;
    const int j0 = P.cellStart[cn], j1 = P.cellStart[cn + 1];
    // in the cell itself only j slots above the chunk's first slot can pair with it;
    // this wave takes part `part` of kPmeJSplit contiguous parts of the j range
    const int jall = nb == 0 ? s0 + chunk * 64 + 1 : j0;
    const int jlen = max(0, j1 - jall);
    const int jstart = jall + (jlen * part) / kPmeJSplit, jend = jall + (jlen * (part + 1)) / kPmeJSplit;
    const double a = D.alpha, a2pi = D.twoAlphaSqrtPi, rc2 = D.rc2;
    double e = 0.0;
    int head = 0, tail = 0;  // ring [head, tail) of pending pairs (wave-uniform)
    const uint64_t below = lane == 0 ? 0ull : (~0ull >> (64 - lane));

    // evaluate `count` queued pairs, one per lane (j data from the staged batch)
    auto dense = [&](int count) {
        if (lane < count) {
            const uint32_t code = sQ[(head + lane) & 127];
            const int il = code >> 16, k = (int)(code & 0xFFFFu);
            const double4 p = sI[il];
            const double2 ip = sIp[il];
            const double4 pj = sJ[k];
            const double2 jp = sJp[k];
            const double dx = p.x - pj.x, dy = p.y - pj.y, dz = p.z - pj.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const double rinv = 1.0 / sqrt(r2);
            const double r = r2 * rinv, r2inv = rinv * rinv;
            const double qq = p.w * pj.w;
            const double ec = qq * erfc(a * r) * rinv;
            const double sg = ip.x + jp.x, e4 = ip.y * jp.y;
            const double s2 = sg * sg * r2inv, s6 = s2 * s2 * s2;
            const double t = e4 * s6;
            const double elj = fma(t, s6, -t);
            const double fs = (ec + qq * a2pi * exp(-a * a * r2) + fma(12.0, elj, 6.0 * t)) * r2inv;
            e += ec + elj;
            const long long fx = __double2ll_rn(fs * dx * kFix), fy = __double2ll_rn(fs * dy * kFix);
            const long long fz = __double2ll_rn(fs * dz * kFix);
            atomicAdd(&sF[0][il], (unsigned long long)fx);
            atomicAdd(&sF[1][il], (unsigned long long)fy);
            atomicAdd(&sF[2][il], (unsigned long long)fz);
            atomicAdd(&sJF[0][k], (unsigned long long)(-fx));
            atomicAdd(&sJF[1][k], (unsigned long long)(-fy));
            atomicAdd(&sJF[2][k], (unsigned long long)(-fz));
        }
        head += count;
    };

    for (int jb = jstart; jb < jend; jb += kPmeJBatch) {
        const int nbat = min(kPmeJBatch, jend - jb);
        __syncthreads();  // the previous batch is fully evaluated and flushed
        for (int k = lane; k < nbat; k += 64) {
            sJ[k] = P.cpos[jb + k];
            sJp[k] = P.cprm[jb + k];
            sJx[k] = P.cex[jb + k];
            sJF[0][k] = 0ull; sJF[1][k] = 0ull; sJF[2][k] = 0ull;
        }
        __syncthreads();
        for (int k = 0; k < nbat; ++k) {
            const int sj = jb + k;
            const double4 pj = sJ[k];
            const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
            const double r2 = dx * dx + dy * dy + dz * dz;
            bool in = active && r2 < rc2 && (nb != 0 || sj > si);
            if (__ballot(in) == 0ull) continue;
            const int4 xj = sJx[k];
            if (in && xj.x == xi.x) {  // same exclusion group: look the pair up
                const uint32_t bits = (xj.y >> 5) == 0 ? word0 : P.exbits[xi.z + (xj.y >> 5)];
                if ((bits >> (xj.y & 31)) & 1u) in = false;
            }
            const uint64_t m = __ballot(in);
            if (in) sQ[(tail + __popcll(m & below)) & 127] = ((uint32_t)lane << 16) | (uint32_t)k;
            tail += __popcll(m);
            if (tail - head >= 64) {
                __syncthreads();
                dense(64);
                __syncthreads();
            }
        }
        if (tail > head) {  // drain: the next batch replaces the staged j data
            __syncthreads();
            dense(tail - head);
        }
        __syncthreads();
        for (int k = lane; k < nbat; k += 64) {  // j forces of the batch
            const int atom = sJx[k].w;
            for (int q = 0; q < 3; ++q) {
                const unsigned long long v = sJF[q][k];
                if (v) atomicAdd(F + (long long)q * D.n + atom, v);
            }
        }
    }
    __syncthreads();
    if (active) {
        const int atom = xi.w;
        for (int k = 0; k < 3; ++k) {
            const unsigned long long v = sF[k][lane];
            if (v) atomicAdd(F + (long long)k * D.n + atom, v);
        }
    }
    e = wave_sum(e);
    if (lane == 0) eslot[blockSlot] = e;
}

__global__ void k_pme_spread(PmeDesc D, PmePtrs P, const double* x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= D.n) return;
    const double q = P.prm[i].y;
    if (q == 0.0) return;
    double th[3][kOrder], dth[kOrder];
    int g0[3];
    for (int d = 0; d < 3; ++d) {
        double f = x[3 * i + d] * D.invL[d];
        f -= floor(f);
        const double u = f * D.K[d];
        int g = (int)floor(u);
        const double w = u - g;
        g0[d] = g;
        bspline5(w, th[d], dth);
    }
    for (int a = 0; a < kOrder; ++a) {
        const int ia = ((g0[0] - a) % D.K[0] + D.K[0]) % D.K[0];
        for (int b = 0; b < kOrder; ++b) {
            const int ib = ((g0[1] - b) % D.K[1] + D.K[1]) % D.K[1];
            const double qab = q * th[0][a] * th[1][b];
            for (int cc = 0; cc < kOrder; ++cc) {
                const int ic = ((g0[2] - cc) % D.K[2] + D.K[2]) % D.K[2];
                atomicAdd(reinterpret_cast<unsigned long long*>(P.qgrid) + ((long long)ia * D.K[1] + ib) * D.K[2] + ic,
                          (unsigned long long)__double2ll_rn(qab * th[2][cc] * kGridFix));
            }
        }
    }
}

__global__ void k_pme_q2c(PmeDesc D, PmePtrs P) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= D.ngrid) return;
    P.g0[k] = make_double2((double)P.qgrid[k] * (1.0 / kGridFix), 0.0);
}

// One axis of an unnormalised complex DFT: out[..k..] = sum_j in[..j..] e^{sign 2 pi i j k / K}.
__global__ void k_pme_dft(PmeDesc D, const double2* in, double2* out, const double2* tw, int axis, double sign) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= D.ngrid) return;
    const int K = D.K[axis];
    const int stride = axis == 2 ? 1 : (axis == 1 ? D.K[2] : D.K[1] * D.K[2]);
    const int k = (idx / stride) % K;
    const int base = idx - k * stride;
    double re = 0.0, im = 0.0;
    int t = 0;
    for (int j = 0; j < K; ++j) {
        const double2 v = in[base + j * stride];
        const double2 w = tw[t];
        const double s = sign * w.y;
        re += v.x * w.x - v.y * s;
        im += v.y * w.x + v.x * s;
        t += k;
        if (t >= K) t -= K;
    }
    out[idx] = make_double2(re, im);
}

// One wave per grid line along `axis` (four lines per block): the line and the
// axis' twiddles are staged in LDS, then lane k (k + 64, ...) forms output k.
// Dynamic LDS: 5 K complex values.
// qfix: when not null, the input is the fixed-point charge grid (real), read in
// place of `in` (the first forward pass needs no separate conversion kernel).
__global__ void __launch_bounds__(256) k_pme_dft_lines(PmeDesc D, const double2* in, double2* out, const double2* tw,
                                                      int axis, double sign, const long long* qfix) {
    extern __shared__ double2 sm[];
    const int K = D.K[axis];
    const int stride = axis == 2 ? 1 : (axis == 1 ? D.K[2] : D.K[1] * D.K[2]);
    const int nlines = D.ngrid / K;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double2* stw = sm;
    double2* line = sm + K + w * K;
    for (int t = threadIdx.x; t < K; t += 256) stw[t] = tw[t];
    const int L = blockIdx.x * 4 + w;
    int base = 0;
    if (L < nlines) {
        base = (L / stride) * K * stride + L % stride;
        if (qfix)
            for (int j = lane; j < K; j += 64) line[j] = make_double2((double)qfix[base + j * stride] * (1.0 / kGridFix), 0.0);
        else
            for (int j = lane; j < K; j += 64) line[j] = in[base + j * stride];
    }
    __syncthreads();
    if (L >= nlines) return;
    for (int k = lane; k < K; k += 64) {
        double re = 0.0, im = 0.0;
        int t = 0;
        for (int j = 0; j < K; ++j) {
            const double2 v = line[j];
            const double2 c = stw[t];
            const double sn = sign * c.y;
            re += v.x * c.x - v.y * sn;
            im += v.y * c.x + v.x * sn;
            t += k;
            if (t >= K) t -= K;
        }
        out[base + k * stride] = make_double2(re, im);
    }
}

// Charge spreading of one cell's atoms into an LDS tile of the grid (64-bit
// fixed point, integer LDS atomics), then one global atomic per nonzero tile
// point.  The tile covers grid indices [o, o + T) per axis with
// o = floor(c K / m) - 5 and T = D.tile[axis]; a point outside it (rounding at
// the cell edge) goes straight to the global grid.
__global__ void __launch_bounds__(256) k_pme_spread_tiles(PmeDesc D, PmePtrs P) {
    extern __shared__ unsigned long long tile[];
    // block (c, part): part `part` of kPmeSpreadSplit contiguous parts of cell c's atoms
    const int c = blockIdx.x / kPmeSpreadSplit, part = blockIdx.x % kPmeSpreadSplit;
    const int cz = c % D.m[2], cy = (c / D.m[2]) % D.m[1], cx = c / (D.m[1] * D.m[2]);
    const int o[3] = {(cx * D.K[0]) / D.m[0] - 5, (cy * D.K[1]) / D.m[1] - 5, (cz * D.K[2]) / D.m[2] - 5};
    const int T0 = D.tile[0], T1 = D.tile[1], T2 = D.tile[2], TT = T0 * T1 * T2;
    for (int k = threadIdx.x; k < TT; k += 256) tile[k] = 0ull;
    __syncthreads();
    const int cs = P.cellStart[c], call = min(P.cellStart[c + 1] - cs, D.cellCap);
    const int s0 = cs + (call * part) / kPmeSpreadSplit, cnt = cs + (call * (part + 1)) / kPmeSpreadSplit - s0;
    // task = (atom, first B-spline index a): 25 points each
    for (int task = threadIdx.x; task < cnt * kOrder; task += 256) {
        const int sl = s0 + task / kOrder, a = task % kOrder;
        const double4 p = P.cpos[sl];
        const double q = P.prm[P.cex[sl].w].y;
        if (q == 0.0) continue;
        double th[3][kOrder], dth[kOrder];
        int g0[3];
        const double pv[3] = {p.x, p.y, p.z};
        for (int d = 0; d < 3; ++d) {
            const double u = pv[d] * D.invL[d] * D.K[d];
            const int g = (int)floor(u);
            g0[d] = g;
            bspline5(u - g, th[d], dth);
        }
        const int la = g0[0] - a - o[0];
        for (int b = 0; b < kOrder; ++b) {
            const int lb = g0[1] - b - o[1];
            const double qab = q * th[0][a] * th[1][b];
            for (int cc = 0; cc < kOrder; ++cc) {
                const int lc = g0[2] - cc - o[2];
                const unsigned long long v = (unsigned long long)__double2ll_rn(qab * th[2][cc] * kGridFix);
                if (la >= 0 && la < T0 && lb >= 0 && lb < T1 && lc >= 0 && lc < T2) {
                    atomicAdd(&tile[(la * T1 + lb) * T2 + lc], v);
                } else {
                    const int ia = wrap_index(g0[0] - a, D.K[0]);
                    const int ib = wrap_index(g0[1] - b, D.K[1]);
                    const int ic = wrap_index(g0[2] - cc, D.K[2]);
                    atomicAdd(reinterpret_cast<unsigned long long*>(P.qgrid) + ((long long)ia * D.K[1] + ib) * D.K[2] + ic, v);
                }
            }
        }
    }
    __syncthreads();
    // tile index o + l lies in [-5, K + 3): one wrap per axis
    for (int k = threadIdx.x; k < TT; k += 256) {
        const unsigned long long v = tile[k];
        if (v == 0ull) continue;
        const int lc = k % T2, lb = (k / T2) % T1, la = k / (T1 * T2);
        const int ia = wrap_index(o[0] + la, D.K[0]);
        const int ib = wrap_index(o[1] + lb, D.K[1]);
        const int ic = wrap_index(o[2] + lc, D.K[2]);
        atomicAdd(reinterpret_cast<unsigned long long*>(P.qgrid) + ((long long)ia * D.K[1] + ib) * D.K[2] + ic, v);
    }
}

__global__ void __launch_bounds__(256) k_pme_convolve(PmeDesc D, PmePtrs P, double2* g, double* eslot) {
    __shared__ double red[4];
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (idx < D.ngrid) {
        const int c = idx % D.K[2], b = (idx / D.K[2]) % D.K[1], a = idx / (D.K[1] * D.K[2]);
        const double mx = (a <= D.K[0] / 2 ? a : a - D.K[0]) * D.invL[0];
        const double my = (b <= D.K[1] / 2 ? b : b - D.K[1]) * D.invL[1];
        const double mz = (c <= D.K[2] / 2 ? c : c - D.K[2]) * D.invL[2];
        const double m2 = mx * mx + my * my + mz * mz;
        double w = 0.0;
        if (idx != 0) {
            const double B = P.bmod[a] * P.bmod[D.K[0] + b] * P.bmod[D.K[0] + D.K[1] + c];
            w = exp(-D.pi2OverAlpha2 * m2) / (m2 * B);
        }
        const double2 f = g[idx];
        e = w * (f.x * f.x + f.y * f.y);
        g[idx] = make_double2(w * f.x, w * f.y);
    }
    const double s = block_sum<256>(e, red);
    if (threadIdx.x == 0) eslot[blockIdx.x] = s * D.recipEnergyScale;
}

// Reciprocal forces, 8 lanes per atom: lane a < 5 of the octet sums the 25 grid
// points of B-spline index a, then the octet adds its partial gradients
// (xor shuffles 1, 2, 4; lanes 5-7 add zeros).  Eight atoms per wave.
__global__ void __launch_bounds__(256) k_pme_gather(PmeDesc D, PmePtrs P, const double* x, const double2* phi,
                                                   unsigned long long* F) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int i = t >> 3, a = t & 7;
    const bool on = i < D.n;
    const double q = on ? P.prm[i].y : 0.0;
    double gx = 0.0, gy = 0.0, gz = 0.0;
    if (on && q != 0.0 && a < kOrder) {
        double th[3][kOrder], dth[3][kOrder];
        int g0[3];
        for (int d = 0; d < 3; ++d) {
            double f = x[3 * i + d] * D.invL[d];
            f -= floor(f);
            const double u = f * D.K[d];
            const int g = (int)floor(u);
            g0[d] = g;
            bspline5(u - g, th[d], dth[d]);
        }
        int ic[kOrder];  // g0 in [0, K]: g0 - k in [-4, K]
#pragma unroll
        for (int cc = 0; cc < kOrder; ++cc) ic[cc] = wrap_index(g0[2] - cc, D.K[2]);
        // th[0][a], dth[0][a] with a run-time a: select from the unrolled values
        double t0 = th[0][0], dt0 = dth[0][0];
#pragma unroll
        for (int k = 1; k < kOrder; ++k)
            if (a == k) { t0 = th[0][k]; dt0 = dth[0][k]; }
        const int ia = wrap_index(g0[0] - a, D.K[0]);
#pragma unroll
        for (int b = 0; b < kOrder; ++b) {
            const int ib = wrap_index(g0[1] - b, D.K[1]);
            const long long row = ((long long)ia * D.K[1] + ib) * D.K[2];
#pragma unroll
            for (int cc = 0; cc < kOrder; ++cc) {
                const double p = phi[row + ic[cc]].x;
                gx += dt0 * th[1][b] * th[2][cc] * p;
                gy += t0 * dth[1][b] * th[2][cc] * p;
                gz += t0 * th[1][b] * dth[2][cc] * p;
            }
        }
    }
#pragma unroll
    for (int o = 1; o < 8; o <<= 1) {
        gx += __shfl_xor(gx, o, 64);
        gy += __shfl_xor(gy, o, 64);
        gz += __shfl_xor(gz, o, 64);
    }
    if (on && q != 0.0 && a < 3) {
        const double s = -q * D.recipPhiScale;
        const double g = a == 0 ? gx * D.K[0] * D.invL[0] : (a == 1 ? gy * D.K[1] * D.invL[1] : gz * D.K[2] * D.invL[2]);
        add_fixed(F + (long long)a * D.n + i, s * g);
    }
}

__global__ void __launch_bounds__(256) k_pme_exceptions(PmeDesc D, PmePtrs P, const double* x, unsigned long long* F,
                                                        double* eslot) {
    __shared__ double red[4];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double e = 0.0;
    if (k < D.nex) {
        const int2 pr = P.expairs[k];
        const double4 ep = P.exprm[k];  // (k_e qq', sigma', 4 eps', k_e q_i q_j)
        const double dx = x[3 * pr.x] - x[3 * pr.y];
        const double dy = x[3 * pr.x + 1] - x[3 * pr.y + 1];
        const double dz = x[3 * pr.x + 2] - x[3 * pr.y + 2];
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double rinv = 1.0 / sqrt(r2), r = r2 * rinv, r2inv = rinv * rinv;
        double fs = 0.0;
        if (ep.x != 0.0) { e += ep.x * rinv; fs += ep.x * rinv * r2inv; }
        if (ep.z != 0.0) {
            const double s2 = ep.y * ep.y * r2inv, s6 = s2 * s2 * s2;
            e += ep.z * (s6 * s6 - s6);
            fs += ep.z * (12.0 * s6 * s6 - 6.0 * s6) * r2inv;
        }
        if (ep.w != 0.0) {
            const double ar = D.alpha * r;
            const double ec = ep.w * erf(ar) * rinv;
            e -= ec;
            fs -= (ec - ep.w * D.twoAlphaSqrtPi * exp(-ar * ar)) * r2inv;
        }
        if (fs != 0.0) {
            add_fixed(F + pr.x, fs * dx); add_fixed(F + D.n + pr.x, fs * dy); add_fixed(F + 2 * D.n + pr.x, fs * dz);
            add_fixed(F + pr.y, -fs * dx); add_fixed(F + D.n + pr.y, -fs * dy); add_fixed(F + 2 * D.n + pr.y, -fs * dz);
        }
    }
    const double s = block_sum<256>(e, red);
    if (threadIdx.x == 0) eslot[blockIdx.x] = s;
}

// Fixed-order sums of the energy slots of both states; U = sum + constants;
// combine.  parts[s][0..4]: direct, exceptions, reciprocal, self, dispersion.
// single: state 0 only (W = U1, hybrid weight 0).
__global__ void __launch_bounds__(1024) k_pme_finish(PmeDesc D, PmePtrs P, int single) {
    __shared__ double red[16];
    double U[2] = {0.0, 0.0};
    for (int s = 0; s < (single ? 1 : 2); ++s) {
        const double* sl = P.eslot + (long long)s * D.nslot;
        const int ranges[4] = {0, D.nslotDirect, D.nslotDirect + D.nslotEx, D.nslot};
        double parts[3];
        for (int r = 0; r < 3; ++r) {
            // four independent partial sums per thread (fixed order): their loads overlap
            double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
            int k = ranges[r] + (int)threadIdx.x;
            for (; k + 3 * 1024 < ranges[r + 1]; k += 4 * 1024) {
                v0 += sl[k];
                v1 += sl[k + 1024];
                v2 += sl[k + 2 * 1024];
                v3 += sl[k + 3 * 1024];
            }
            for (; k < ranges[r + 1]; k += 1024) v0 += sl[k];
            parts[r] = block_sum<1024>((v0 + v1) + (v2 + v3), red);
        }
        if (threadIdx.x == 0) {
            P.parts[5 * s + 0] = parts[0];
            P.parts[5 * s + 1] = parts[1];
            P.parts[5 * s + 2] = parts[2];
            P.parts[5 * s + 3] = D.selfEnergy;
            P.parts[5 * s + 4] = D.dispersionEnergy;
            U[s] = parts[0] + parts[1] + parts[2] + D.selfEnergy + D.dispersionEnergy + (P.eoff && !single ? P.eoff[s] : 0.0);
        }
    }
    if (threadIdx.x == 0) {
        SysResult r;
        if (single) {
            r.W = U[0]; r.usc = 0.0; r.U1 = U[0]; r.U2 = 0.0; r.u = 0.0; r.fp = 1.0; r.bfp = 0.0; r.sp = 0.0;
        } else {
            atm_combine(*P.params, U[0], U[1], r);
        }
        *P.result = r;
    }
}

__global__ void k_pme_blend(PmeDesc D, PmePtrs P, int single, double* out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= 3 * D.n) return;
    const SysResult r = *P.result;
    const double a = single ? 0.0 : (P.params->direction > 0.0 ? r.sp : 1.0 - r.sp);  // weight of state 2 (x + d)
    const double f1 = (double)(long long)P.F[k] * (1.0 / kFix);
    const double f2 = single ? 0.0 : (double)(long long)P.F[3ll * D.n + k] * (1.0 / kFix);
    const int comp = k / D.n, i = k % D.n;
    out[3 * i + comp] = (1.0 - a) * f1 + a * f2;
}

__global__ void k_pme_fill(unsigned long long* p, long long n) {
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) p[k] = 0ull;
}

// Three buffers cleared by one launch (each launch of a graph costs a few us):
// thread k clears word k of the concatenation a | b | c.
__global__ void k_pme_fill3(unsigned long long* a, long long na, unsigned long long* b, long long nb,
                            unsigned long long* c, long long nc) {
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < na) { a[k] = 0ull; return; }
    k -= na;
    if (k < nb) { b[k] = 0ull; return; }
    k -= nb;
    if (k < nc) c[k] = 0ull;
}

static inline unsigned blocks_of(long long n, int t) { return (unsigned)((n + t - 1) / t); }

hipError_t launch_pme_fill(unsigned long long* p, long long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    k_pme_fill<<<blocks_of(n, 256), 256, 0, st>>>(p, n);
    return hipGetLastError();
}

hipError_t launch_pme_state(const PmeDesc& D, const PmePtrs& P, int state, hipStream_t st) {
    const double* x = state == 0 ? P.x : P.x2;
    unsigned long long* F = P.F + 3ll * D.n * state;
    double* es = P.eslot + (long long)D.nslot * state;
    {
        const long long nc = (D.ncell + 1) / 2, ng = D.ngrid, nf = 3ll * D.n;
        k_pme_fill3<<<blocks_of(nc + ng + nf, 256), 256, 0, st>>>(reinterpret_cast<unsigned long long*>(P.cellCount), nc,
                                                                  reinterpret_cast<unsigned long long*>(P.qgrid), ng, F, nf);
    }
    // 256 threads: the histogram loops assume it; state 1 forms x + d here
    k_pme_bin<<<blocks_of(D.n, 256), 256, 0, st>>>(D, P, state == 1 ? P.x : x, state == 1 ? P.displ : nullptr);
    k_pme_scan<<<1, 1024, 0, st>>>(D, P);
    k_pme_scatter<<<blocks_of(D.n, 256), 256, 0, st>>>(D, P);
    k_pme_cellsort<<<D.ncell, 256, 0, st>>>(D, P);
    // neighbour offset fastest, chunk slowest: consecutive workgroups (dealt round-robin to the
    // XCDs) carry comparable work, and the mostly empty late chunks come last
    k_pme_direct<<<dim3(kPmeHalfShell * kPmeJSplit, D.ncell, D.chunks), 64, 0, st>>>(D, P, F, es);
    k_pme_exceptions<<<blocks_of(D.nex, 256) > 0 ? blocks_of(D.nex, 256) : 1, 256, 0, st>>>(D, P, x, F,
                                                                                               es + D.nslotDirect);
    if (D.tileBytes > 0)  // the cell's grid tile fits in LDS
        k_pme_spread_tiles<<<D.ncell * kPmeSpreadSplit, 256, D.tileBytes, st>>>(D, P);
    else
        k_pme_spread<<<blocks_of(D.n, 256), 256, 0, st>>>(D, P, x);
    const unsigned gb = blocks_of(D.ngrid, 256);
    auto dft = [&](const double2* in, double2* out, int axis, double sign, const long long* qfix = nullptr) {
        const int K = D.K[axis];
        if (K <= kPmeDftLdsMax) {
            k_pme_dft_lines<<<blocks_of(D.ngrid / K, 4), 256, 5 * K * sizeof(double2), st>>>(D, in, out, P.tw + D.twOff[axis],
                                                                                          axis, sign, qfix);
        } else {
            if (qfix) k_pme_q2c<<<gb, 256, 0, st>>>(D, P);  // into g0 = in
            k_pme_dft<<<gb, 256, 0, st>>>(D, in, out, P.tw + D.twOff[axis], axis, sign);
        }
    };
    // first pass straight from the fixed-point charge grid
    dft(P.g0, P.g1, 2, -1.0, reinterpret_cast<const long long*>(P.qgrid));
    dft(P.g1, P.g0, 1, -1.0);
    dft(P.g0, P.g1, 0, -1.0);
    k_pme_convolve<<<gb, 256, 0, st>>>(D, P, P.g1, es + D.nslotDirect + D.nslotEx);
    dft(P.g1, P.g0, 0, 1.0);
    dft(P.g0, P.g1, 1, 1.0);
    dft(P.g1, P.g0, 2, 1.0);
    k_pme_gather<<<blocks_of(8ll * D.n, 256), 256, 0, st>>>(D, P, x, P.g0, F);
    return hipGetLastError();
}

hipError_t launch_pme_finish(const PmeDesc& D, const PmePtrs& P, bool single, double* forcesOut, hipStream_t st) {
    k_pme_finish<<<1, 1024, 0, st>>>(D, P, single ? 1 : 0);
    if (forcesOut) k_pme_blend<<<blocks_of(3ll * D.n, 256), 256, 0, st>>>(D, P, single ? 1 : 0, forcesOut);
    return hipGetLastError();
}

// ------------------------------------------------------------------ bonded terms
// HarmonicBondForce, HarmonicAngleForce and PeriodicTorsionForce of the
// system's group 0 (the AMBER terms test_abfe.py keeps beside the ATM force).

__global__ void __launch_bounds__(256) k_bonded(BondedPtrs B, double* eslot) {
    __shared__ double red[4];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double eb = 0.0, ea = 0.0, et = 0.0;
    const double* x = B.x;
    auto diff = [&](int i, int j, double d[3]) {
        for (int c = 0; c < 3; ++c) d[c] = x[3 * i + c] - x[3 * j + c];
    };
    auto addf = [&](int i, double s, const double d[3]) {
        for (int c = 0; c < 3; ++c) add_fixed(B.F + 3ll * i + c, s * d[c]);
    };
    if (k < B.nb) {
        const int2 a = B.bonds[k];
        const double2 p = B.bprm[k];  // r0, k
        double d[3];
        diff(a.x, a.y, d);
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double dr = r - p.x;
        eb = 0.5 * p.y * dr * dr;
        const double s = -p.y * dr / r;
        addf(a.x, s, d);
        addf(a.y, -s, d);
    } else if (k < B.nb + B.na) {
        const int4 a = B.angles[k - B.nb];
        const double2 p = B.aprm[k - B.nb];  // theta0, k
        double v1[3], v2[3];
        diff(a.x, a.y, v1);
        diff(a.z, a.y, v2);
        const double r1 = sqrt(v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2]);
        const double r2 = sqrt(v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2]);
        double c = (v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2]) / (r1 * r2);
        c = fmin(1.0, fmax(-1.0, c));
        const double th = acos(c);
        const double dth = th - p.x;
        ea = 0.5 * p.y * dth * dth;
        const double sn = fmax(sqrt(1.0 - c * c), 1e-300);
        const double g = p.y * dth / sn;  // -dE/dtheta * (-1/sin)
        double f1[3], f3[3];
        for (int q = 0; q < 3; ++q) {
            f1[q] = g * (v2[q] / (r1 * r2) - c * v1[q] / (r1 * r1));
            f3[q] = g * (v1[q] / (r1 * r2) - c * v2[q] / (r2 * r2));
        }
        const double one[3] = {1.0, 1.0, 1.0};
        for (int q = 0; q < 3; ++q) {
            add_fixed(B.F + 3ll * a.x + q, f1[q]);
            add_fixed(B.F + 3ll * a.z + q, f3[q]);
            add_fixed(B.F + 3ll * a.y + q, -(f1[q] + f3[q]) * one[q]);
        }
    } else if (k < B.nb + B.na + B.nt) {
        const int t = k - B.nb - B.na;
        const int4 a = B.torsions[t];
        const double4 p = B.tprm[t];  // periodicity, phase, k
        double rij[3], rkj[3], rkl[3];
        diff(a.x, a.y, rij);
        diff(a.z, a.y, rkj);
        diff(a.z, a.w, rkl);
        const double m[3] = {rij[1] * rkj[2] - rij[2] * rkj[1], rij[2] * rkj[0] - rij[0] * rkj[2],
                             rij[0] * rkj[1] - rij[1] * rkj[0]};
        const double nv[3] = {rkj[1] * rkl[2] - rkj[2] * rkl[1], rkj[2] * rkl[0] - rkj[0] * rkl[2],
                              rkj[0] * rkl[1] - rkj[1] * rkl[0]};
        const double m2 = m[0] * m[0] + m[1] * m[1] + m[2] * m[2];
        const double n2 = nv[0] * nv[0] + nv[1] * nv[1] + nv[2] * nv[2];
        const double rkj2 = rkj[0] * rkj[0] + rkj[1] * rkj[1] + rkj[2] * rkj[2];
        const double nrkj = sqrt(rkj2);
        double cp = (m[0] * nv[0] + m[1] * nv[1] + m[2] * nv[2]) / sqrt(m2 * n2);
        cp = fmin(1.0, fmax(-1.0, cp));
        const double sgn = rij[0] * nv[0] + rij[1] * nv[1] + rij[2] * nv[2];
        const double phi = sgn < 0.0 ? -acos(cp) : acos(cp);
        et = p.z * (1.0 + cos(p.x * phi - p.y));
        const double dV = -p.z * p.x * sin(p.x * phi - p.y);
        double fi[3], fl[3], fj[3], fk[3];
        const double pp = (rij[0] * rkj[0] + rij[1] * rkj[1] + rij[2] * rkj[2]) / rkj2;
        const double qq = (rkl[0] * rkj[0] + rkl[1] * rkj[1] + rkl[2] * rkj[2]) / rkj2;
        for (int q = 0; q < 3; ++q) {
            fi[q] = -dV * nrkj / m2 * m[q];
            fl[q] = dV * nrkj / n2 * nv[q];
        }
        for (int q = 0; q < 3; ++q) {
            const double sv = pp * fi[q] - qq * fl[q];
            fj[q] = fi[q] - sv;
            fk[q] = fl[q] + sv;
        }
        for (int q = 0; q < 3; ++q) {
            add_fixed(B.F + 3ll * a.x + q, fi[q]);
            add_fixed(B.F + 3ll * a.y + q, -fj[q]);
            add_fixed(B.F + 3ll * a.z + q, -fk[q]);
            add_fixed(B.F + 3ll * a.w + q, fl[q]);
        }
    }
    const double sb = block_sum<256>(eb, red);
    const double sa = block_sum<256>(ea, red);
    const double stt = block_sum<256>(et, red);
    if (threadIdx.x == 0) {
        eslot[3 * blockIdx.x] = sb;
        eslot[3 * blockIdx.x + 1] = sa;
        eslot[3 * blockIdx.x + 2] = stt;
    }
}

hipError_t launch_bonded(const BondedPtrs& B, double* eslot, int nblocks, hipStream_t st) {
    if (nblocks <= 0) return hipSuccess;
    k_bonded<<<nblocks, 256, 0, st>>>(B, eslot);
    return hipGetLastError();
}

}  // namespace atm
