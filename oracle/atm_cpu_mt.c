/*
 * atm_cpu_mt.c -- multi-threaded CPU evaluation of the ATM energy and forces:
 * the "CPU-equivalent" baseline timed beside the MI355X library.
 *
 * TEST INFRASTRUCTURE ONLY.  Loaded by bench.py's cpu_baseline leg and by
 * tests/; no product code path calls it.
 *
 * The algorithm is atm_oracle.c's (copyState's x + d, then the LJ +
 * reaction-field cutoff sums on x and x + d, the soft-core/softplus combine of
 * platforms/reference/src/ReferenceATMMetaForceKernels.cpp:56-98 and the blend
 * of :100-110), organised the way a multi-core CPU MD code runs it:
 *   - per evaluation and coordinate set, a cell list with cells of edge >= rc
 *     (at least 3 per dimension), built by a counting sort into cell order;
 *     the sorted copy holds FP32 wrapped coordinates and per-atom parameters
 *     (q sqrt(k_e), sigma/2, 2 sqrt(eps)) in SoA arrays;
 *   - OpenMP over i-cells (dynamic schedule); each i-cell visits itself
 *     (j > i) and its 13 forward neighbour cells, the image shift of a
 *     neighbour known from the cell offset, so the inner j loop is a plain
 *     contiguous loop the compiler vectorises;
 *   - FP32 pair math, FP64 energy sums per thread, forces summed in per-thread
 *     FP64 arrays and reduced over threads at the end.
 * Results agree with the FP64 oracle to FP32 rounding (pairs within an ulp of
 * rc may be decided differently); it is a speed baseline, not a checker.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define MT_KE 138.935456

typedef struct oracle_params {
    double lambda1, lambda2, alpha, u0, w0, umax, ubcore, acore, direction;
} oracle_params;

typedef struct oracle_result {
    double W, usc, U1, U2, u, fp, bfp, sp;
    long long npairs1, npairs2;
    double abs1, abs2;
} oracle_result;

/* from atm_oracle.c (libatm_oracle.so) */
void oracle_combine(const oracle_params* p, double U1, double U2, oracle_result* r);

typedef struct {
    int n, nthreads;
    int m[3], ncell;
    int *cellStart, *cellCount, *order, *cellOf;
    float *xs, *ys, *zs, *qs, *hs, *es;
    double* fbuf; /* nthreads x 3 x n */
} mt_work;

static int work_init(mt_work* w, int n, int nthreads, const double* box, double rc) {
    memset(w, 0, sizeof(*w));
    w->n = n;
    w->nthreads = nthreads;
    for (int k = 0; k < 3; ++k) {
        w->m[k] = (int)floor(box[k] / rc);
        if (w->m[k] < 3) return -1;
    }
    w->ncell = w->m[0] * w->m[1] * w->m[2];
    w->cellStart = (int*)malloc(sizeof(int) * (w->ncell + 1));
    w->cellCount = (int*)malloc(sizeof(int) * w->ncell);
    w->order = (int*)malloc(sizeof(int) * n);
    w->cellOf = (int*)malloc(sizeof(int) * n);
    w->xs = (float*)malloc(sizeof(float) * n);
    w->ys = (float*)malloc(sizeof(float) * n);
    w->zs = (float*)malloc(sizeof(float) * n);
    w->qs = (float*)malloc(sizeof(float) * n);
    w->hs = (float*)malloc(sizeof(float) * n);
    w->es = (float*)malloc(sizeof(float) * n);
    w->fbuf = (double*)malloc(sizeof(double) * 3 * (size_t)n * nthreads);
    if (!w->cellStart || !w->cellCount || !w->order || !w->cellOf || !w->xs || !w->ys || !w->zs || !w->qs ||
        !w->hs || !w->es || !w->fbuf)
        return -2;
    return 0;
}

static void work_free(mt_work* w) {
    free(w->cellStart); free(w->cellCount); free(w->order); free(w->cellOf);
    free(w->xs); free(w->ys); free(w->zs); free(w->qs); free(w->hs); free(w->es);
    free(w->fbuf);
}

/* Energy of one coordinate set; F (3n, AoS, overwritten) if non-NULL. */
static double nonbonded_mt(mt_work* w, const double* x, const double* q, const double* sig, const double* eps,
                           const double* box, double rc, double eps_rf, double* F, long long* npairs) {
    const int n = w->n, nt = w->nthreads, *m = w->m;
    const float Lf[3] = {(float)box[0], (float)box[1], (float)box[2]};
    /* cell of every atom (wrapped), counting sort into cell order */
    memset(w->cellCount, 0, sizeof(int) * w->ncell);
    for (int i = 0; i < n; ++i) {
        int ci[3];
        for (int k = 0; k < 3; ++k) {
            double v = x[3 * i + k] - box[k] * floor(x[3 * i + k] / box[k]);
            int c = (int)(v / box[k] * m[k]);
            ci[k] = c < 0 ? 0 : (c >= m[k] ? m[k] - 1 : c);
        }
        const int c = (ci[0] * m[1] + ci[1]) * m[2] + ci[2];
        w->cellOf[i] = c;
        w->cellCount[c]++;
    }
    w->cellStart[0] = 0;
    for (int c = 0; c < w->ncell; ++c) w->cellStart[c + 1] = w->cellStart[c] + w->cellCount[c];
    memset(w->cellCount, 0, sizeof(int) * w->ncell);
    for (int i = 0; i < n; ++i) {
        const int c = w->cellOf[i];
        const int s = w->cellStart[c] + w->cellCount[c]++;
        w->order[s] = i;
        w->xs[s] = (float)(x[3 * i] - box[0] * floor(x[3 * i] / box[0]));
        w->ys[s] = (float)(x[3 * i + 1] - box[1] * floor(x[3 * i + 1] / box[1]));
        w->zs[s] = (float)(x[3 * i + 2] - box[2] * floor(x[3 * i + 2] / box[2]));
        w->qs[s] = (float)(q[i] * sqrt(MT_KE));
        w->hs[s] = (float)(0.5 * sig[i]);
        w->es[s] = (float)(2.0 * sqrt(eps[i]));
    }
    const float rc2 = (float)(rc * rc);
    const double krf64 = (eps_rf - 1.0) / ((2.0 * eps_rf + 1.0) * rc * rc * rc);
    const double crf64 = 3.0 * eps_rf / ((2.0 * eps_rf + 1.0) * rc);
    const float krf = (float)krf64;
    double Etot = 0.0;
    long long np = 0;
    /* the 13 forward neighbour offsets (lexicographically positive) */
    int off[13][3], no = 0;
    for (int a = -1; a <= 1; ++a)
        for (int b = -1; b <= 1; ++b)
            for (int e = -1; e <= 1; ++e)
                if (a > 0 || (a == 0 && (b > 0 || (b == 0 && e > 0)))) {
                    off[no][0] = a; off[no][1] = b; off[no][2] = e; ++no;
                }
#pragma omp parallel num_threads(nt) reduction(+ : Etot, np)
    {
        const int t = omp_get_thread_num();
        double* fx = w->fbuf + (size_t)t * 3 * n;
        double* fy = fx + n;
        double* fz = fy + n;
        memset(fx, 0, sizeof(double) * 3 * n);
        double eE = 0.0, eQQ = 0.0, eQR2 = 0.0;
#pragma omp for schedule(dynamic, 4)
        for (int c = 0; c < w->ncell; ++c) {
            const int cz = c % m[2], cy = (c / m[2]) % m[1], cx = c / (m[1] * m[2]);
            const int i0 = w->cellStart[c], i1 = w->cellStart[c + 1];
            for (int nb = -1; nb < 13; ++nb) {
                int j0, j1;
                float sx = 0.f, sy = 0.f, sz = 0.f;
                if (nb < 0) {
                    j0 = i0; j1 = i1;
                } else {
                    int nx = cx + off[nb][0], ny = cy + off[nb][1], nz = cz + off[nb][2];
                    if (nx >= m[0]) { nx -= m[0]; sx = Lf[0]; } else if (nx < 0) { nx += m[0]; sx = -Lf[0]; }
                    if (ny >= m[1]) { ny -= m[1]; sy = Lf[1]; } else if (ny < 0) { ny += m[1]; sy = -Lf[1]; }
                    if (nz >= m[2]) { nz -= m[2]; sz = Lf[2]; } else if (nz < 0) { nz += m[2]; sz = -Lf[2]; }
                    const int cj = (nx * m[1] + ny) * m[2] + nz;
                    j0 = w->cellStart[cj]; j1 = w->cellStart[cj + 1];
                }
                for (int i = i0; i < i1; ++i) {
                    const float xi = w->xs[i] - sx, yi = w->ys[i] - sy, zi = w->zs[i] - sz;
                    const float qi = w->qs[i], hi = w->hs[i], ei = w->es[i];
                    float fxi = 0.f, fyi = 0.f, fzi = 0.f;
                    double e1 = 0.0, e2 = 0.0, e3 = 0.0;
                    long long cnt = 0;
                    const int js = nb < 0 ? i + 1 : j0;
#pragma omp simd reduction(+ : fxi, fyi, fzi, e1, e2, e3, cnt)
                    for (int j = js; j < j1; ++j) {
                        const float dx = xi - w->xs[j], dy = yi - w->ys[j], dz = zi - w->zs[j];
                        const float r2 = dx * dx + dy * dy + dz * dz;
                        if (r2 < rc2) {
                            const float rinv = 1.0f / sqrtf(r2);
                            const float r2inv = rinv * rinv;
                            const float s = hi + w->hs[j];
                            const float s2 = s * s * r2inv;
                            const float s6 = s2 * s2 * s2;
                            const float e4 = ei * w->es[j];
                            const float qq = qi * w->qs[j];
                            const float elj = e4 * (s6 * s6 - s6);
                            const float fs = (e4 * s6 * (12.0f * s6 - 6.0f) + qq * (rinv - 2.0f * krf * r2)) * r2inv;
                            e1 += (double)(elj + qq * rinv);
                            e2 += (double)qq;
                            e3 += (double)(qq * r2);
                            fxi += fs * dx; fyi += fs * dy; fzi += fs * dz;
                            fx[j] -= (double)(fs * dx);
                            fy[j] -= (double)(fs * dy);
                            fz[j] -= (double)(fs * dz);
                            cnt += 1;
                        }
                    }
                    fx[i] += fxi; fy[i] += fyi; fz[i] += fzi;
                    eE += e1; eQQ += e2; eQR2 += e3;
                    np += cnt;
                }
            }
        }
        Etot += eE + krf64 * eQR2 - crf64 * eQQ;
    }
    if (F) {
#pragma omp parallel for num_threads(nt) schedule(static)
        for (int s = 0; s < n; ++s) {
            double a = 0.0, b = 0.0, d = 0.0;
            for (int t = 0; t < nt; ++t) {
                const double* f = w->fbuf + (size_t)t * 3 * n;
                a += f[s]; b += f[n + s]; d += f[2 * n + s];
            }
            const int i = w->order[s];
            F[3 * i] = a; F[3 * i + 1] = b; F[3 * i + 2] = d;
        }
    }
    if (npairs) *npairs = np;
    return Etot;
}

/* One ATM evaluation on `nthreads` threads (<= 0: OpenMP's default): U1, F1
 * at x; U2, F2 at x + d; combine; F (3n, overwritten with the ATM force) if
 * non-NULL.  Returns 0, or < 0 if a box edge holds fewer than 3 cells of rc. */
int cpu_mt_atm_eval(int n, const double* x, const double* d, const double* q, const double* sig, const double* eps,
                    const double* box, double rc, double eps_rf, const oracle_params* p, int nthreads, double* F,
                    oracle_result* r) {
    if (nthreads <= 0) nthreads = omp_get_max_threads();
    mt_work w;
    int rc0 = work_init(&w, n, nthreads, box, rc);
    double* x2 = (double*)malloc(sizeof(double) * 3 * n);
    double* F1 = F ? (double*)malloc(sizeof(double) * 3 * n) : NULL;
    double* F2 = F ? (double*)malloc(sizeof(double) * 3 * n) : NULL;
    if (rc0 != 0 || !x2 || (F && (!F1 || !F2))) {
        work_free(&w); free(x2); free(F1); free(F2);
        return rc0 != 0 ? rc0 : -2;
    }
    for (int i = 0; i < 3 * n; ++i) x2[i] = x[i] + d[i];
    long long n1 = 0, n2 = 0;
    const double U1 = nonbonded_mt(&w, x, q, sig, eps, box, rc, eps_rf, F1, &n1);
    const double U2 = nonbonded_mt(&w, x2, q, sig, eps, box, rc, eps_rf, F2, &n2);
    oracle_combine(p, U1, U2, r);
    r->npairs1 = n1;
    r->npairs2 = n2;
    r->abs1 = r->abs2 = 0.0;
    if (F) {
        const double sp = r->sp, a = p->direction > 0.0 ? sp : 1.0 - sp;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int k = 0; k < 3 * n; ++k) F[k] = a * F2[k] + (1.0 - a) * F1[k];
    }
    work_free(&w);
    free(x2); free(F1); free(F2);
    return 0;
}

int cpu_mt_max_threads(void) { return omp_get_max_threads(); }
