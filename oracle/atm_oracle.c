/*
 * atm_oracle.c -- FP64 CPU restatement of the ATMMetaForce evaluation.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker of the MI355X library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 * No product code path calls it.
 *
 * What it restates, and from where (paths relative to the reference tree):
 *   oracle_softcore      SoftCoreF, platforms/reference/src/ReferenceATMMetaForceKernels.cpp:26-37
 *   oracle_combine       execute, ReferenceATMMetaForceKernels.cpp:56-98 (u, e0, soft-core,
 *                        softplus, bfp, W), written exactly as the reference evaluates it
 *                        (ee = 1 + exp(-alpha (u_sc - u0)), log(ee))
 *   oracle_blend         the hybrid force, ReferenceATMMetaForceKernels.cpp:100-110
 *   oracle_atm_eval      ATMMetaForceImpl::calcForcesAndEnergy, openmmapi/src/ATMMetaForceImpl.cpp:90-128,
 *                        with copyState's pos2 = pos + displ[i] (ReferenceATMMetaForceKernels.cpp:116-124)
 *   oracle_nonbonded     the inner contexts' variable-group energy: OpenMM NonbondedForce,
 *                        CutoffPeriodic with reaction field (external to the reference; reached at
 *                        ATMMetaForceImpl.cpp:113,116).  OpenMM is not vendored; the formulas are
 *                        its documented ones:
 *                          E_LJ = 4 eps_ij [(s_ij/r)^12 - (s_ij/r)^6], s_ij = (s_i+s_j)/2, eps_ij = sqrt(eps_i eps_j)
 *                          E_C  = k_e q_i q_j (1/r + k_rf r^2 - c_rf),
 *                          k_rf = (e_rf-1)/((2 e_rf+1) rc^3), c_rf = 3 e_rf/((2 e_rf+1) rc)
 *                        for pairs i<j with minimum-image r^2 < rc^2 (strict), k_e = 138.935456,
 *                        that are not NonbondedForce exceptions; no dispersion correction.
 *                        Each exception (i, j, q', s', e') adds k_e q'/r + 4 e' [(s'/r)^12 - (s'/r)^6]
 *                        at the plain (non-periodic) distance, with no cutoff (OpenMM's documented
 *                        exception terms; q' = e' = 0 is an exclusion).  The build's convention for
 *                        the reaction-field method, the same as its PME path's.
 *   triclinic boxes      oracle_pairs_tri, oracle_atm_eval_tri: pairs take OpenMM's triclinic image
 *                        rule (tri_image; box vectors copied by copyState,
 *                        ReferenceATMMetaForceKernels.cpp:126-129)
 *   oracle_atm_grad      analytic dW/d(q, sigma, epsilon) and dW/d(lambda1, lambda2, alpha, u0, w0)
 *                        (no reference counterpart: the reference has no gradients).
 *
 * How each part is pinned (tests/test_oracle.py): closed-form two-atom values (E_LJ(sigma) = 0,
 * F_LJ(2^(1/6) sigma) = 0, E_LJ = -eps there, E_C(rc) = 0, F_C(rc) = k_e q q / rc^2 * 3/(2 e_rf + 1),
 * e_rf = 1 gives the shifted Coulomb), F = -grad U by central differences, sum F = 0, translation
 * invariance, cell list == brute force, soft-core invariants (f(ub) = ub, f'(ub) = 1, f -> umax,
 * f' vs central differences, monotone), softplus special cases (lambda1 = lambda2, direction
 * symmetry, dW/du_sc = bfp), blend end points, gradients vs central differences, and the
 * reference's serialization values.  Parity of the pair term against an installed OpenMM is
 * unpinned (OpenMM is absent).
 *
 * Summation: energies use Neumaier-compensated FP64 sums; forces plain FP64.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_KE 138.935456

typedef struct oracle_params {
    double lambda1, lambda2, alpha, u0, w0, umax, ubcore, acore, direction;
} oracle_params;

typedef struct oracle_result {
    double W, usc, U1, U2, u, fp, bfp, sp;
    long long npairs1, npairs2;
    double abs1, abs2; /* sum over pairs of |e_ij| in each state: the energies' condition scale */
} oracle_result;

/* ------------------------------------------------------------------ ATM scalars */

double oracle_softcore(double u, double umax, double acore, double ubcore, double* fp) {
    if (u <= ubcore) {
        *fp = 1.0;
        return u;
    }
    double gu = (u - ubcore) / (acore * (umax - ubcore));
    double zeta = 1.0 + 2.0 * gu * (gu + 1.0);
    double zetap = pow(zeta, acore);
    double s = 4.0 * (2.0 * gu + 1.0) / zeta;
    *fp = s * zetap / pow(1.0 + zetap, 2);
    return (umax - ubcore) * (zetap - 1.0) / (zetap + 1.0) + ubcore;
}

void oracle_combine(const oracle_params* p, double U1, double U2, oracle_result* r) {
    double fp;
    double u = p->direction > 0 ? U2 - U1 : U1 - U2;
    double e0 = p->direction > 0 ? U1 : U2;
    double usc = oracle_softcore(u, p->umax, p->acore, p->ubcore, &fp);
    double ebias = 0.0;
    double ee = 1.0 + exp(-p->alpha * (usc - p->u0));
    if (p->alpha > 0) ebias = ((p->lambda2 - p->lambda1) / p->alpha) * log(ee);
    ebias += p->lambda2 * usc + p->w0;
    double bfp = (p->lambda2 - p->lambda1) / ee + p->lambda1;
    r->W = e0 + ebias;
    r->usc = usc;
    r->U1 = U1;
    r->U2 = U2;
    r->u = u;
    r->fp = fp;
    r->bfp = bfp;
    r->sp = bfp * fp;
}

/* F += sp F2 + (1-sp) F1 (direction > 0) or sp F1 + (1-sp) F2 */
void oracle_blend(int n, double sp, double direction, const double* F1, const double* F2, double* F) {
    for (int i = 0; i < 3 * n; ++i) {
        if (direction > 0) F[i] += sp * F2[i] + (1.0 - sp) * F1[i];
        else F[i] += sp * F1[i] + (1.0 - sp) * F2[i];
    }
}

/* ------------------------------------------------------------------ pairs */

static inline double mimg(double d, double L) { return d - L * floor(d / L + 0.5); }

/* Image of a difference vector in a triclinic box in OpenMM's reduced form,
 * a = (box[0], 0, 0), b = (tilt[0], box[1], 0), c = (tilt[1], tilt[2], box[2]):
 * OpenMM's rule for triclinic boxes (ReferenceForce::getDeltaRPeriodic in OpenMM, external
 * to the reference tree, reached through the inner contexts' NonbondedForce at
 * ATMMetaForceImpl.cpp:113,116 with the box vectors copyState copies,
 * ReferenceATMMetaForceKernels.cpp:126-129): subtract c floor(dz/cz + 1/2), then
 * b floor(dy/by + 1/2), then a floor(dx/ax + 1/2). */
static inline void tri_image(const double* box, const double* tilt, double* dx, double* dy, double* dz) {
    double s = floor(*dz / box[2] + 0.5);
    *dx -= tilt[1] * s;
    *dy -= tilt[2] * s;
    *dz -= box[2] * s;
    s = floor(*dy / box[1] + 0.5);
    *dx -= tilt[0] * s;
    *dy -= box[1] * s;
    *dx -= box[0] * floor(*dx / box[0] + 0.5);
}

typedef struct {
    double s, c;
} nsum; /* Neumaier sum */

static inline void nsum_add(nsum* a, double v) {
    double t = a->s + v;
    if (fabs(a->s) >= fabs(v)) a->c += (a->s - t) + v;
    else a->c += (v - t) + a->s;
    a->s = t;
}

/* Cell grid: m[k] = floor(L_k / rc) cells per edge; usable when every m[k] >= 3. */
typedef struct {
    int m[3];
    int* head;
    int* next;
} grid;

static int grid_build(grid* g, int n, const double* x, const double* box, double rc) {
    for (int k = 0; k < 3; ++k) {
        g->m[k] = (int)floor(box[k] / rc);
        if (g->m[k] < 3) return 0;
    }
    int nc = g->m[0] * g->m[1] * g->m[2];
    g->head = (int*)malloc(sizeof(int) * nc);
    g->next = (int*)malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int c = 0; c < nc; ++c) g->head[c] = -1;
    for (int i = n - 1; i >= 0; --i) {
        int ci[3];
        for (int k = 0; k < 3; ++k) {
            double w = x[3 * i + k] - box[k] * floor(x[3 * i + k] / box[k]);
            int c = (int)(w / box[k] * g->m[k]);
            if (c >= g->m[k]) c = g->m[k] - 1;
            if (c < 0) c = 0;
            ci[k] = c;
        }
        int c = (ci[0] * g->m[1] + ci[1]) * g->m[2] + ci[2];
        g->next[i] = g->head[c];
        g->head[c] = i;
    }
    return 1;
}

static void grid_free(grid* g) {
    free(g->head);
    free(g->next);
}

typedef void (*pair_fn)(void* ctx, int i, int j, double dx, double dy, double dz, double r2);

/* Exception pairs as sorted keys i << 32 | j (i < j). */
typedef struct {
    unsigned long long* key;
    long long n;
} excl_set;

static int cmp_u64(const void* a, const void* b) {
    unsigned long long x = *(const unsigned long long*)a, y = *(const unsigned long long*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

static void excl_build(excl_set* e, int nexc, const int* pairs) {
    e->n = nexc;
    e->key = nexc > 0 ? (unsigned long long*)malloc(sizeof(unsigned long long) * nexc) : NULL;
    for (int k = 0; k < nexc; ++k) {
        unsigned long long a = (unsigned)pairs[2 * k], b = (unsigned)pairs[2 * k + 1];
        e->key[k] = a < b ? (a << 32 | b) : (b << 32 | a);
    }
    if (nexc > 1) qsort(e->key, nexc, sizeof(unsigned long long), cmp_u64);
}

static int excl_has(const excl_set* e, int i, int j) {
    if (!e || e->n == 0) return 0;
    unsigned long long k = (unsigned long long)(unsigned)i << 32 | (unsigned)j;
    return bsearch(&k, e->key, e->n, sizeof(unsigned long long), cmp_u64) != NULL;
}

/* Visit every pair i<j with minimum-image r^2 < rc^2 that is not in `ex` exactly once, in a
 * fixed order.  brute != 0 forces the O(N^2) loop.  tilt (b_x, c_x, c_y; NULL: orthorhombic)
 * makes the box triclinic: pairs take tri_image's image, in the O(N^2) loop. */
static long long visit_pairs_tilt(int n, const double* x, const double* box, const double* tilt, double rc,
                                  int brute, const excl_set* ex, pair_fn fn, void* ctx) {
    const double rc2 = rc * rc;
    long long cnt = 0;
    if (tilt) {
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) {
                double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
                tri_image(box, tilt, &dx, &dy, &dz);
                double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 < rc2 && !excl_has(ex, i, j)) {
                    ++cnt;
                    if (fn) fn(ctx, i, j, dx, dy, dz, r2);
                }
            }
        return cnt;
    }
    grid g;
    if (!brute && grid_build(&g, n, x, box, rc)) {
        /* for each atom i, scan the 27 neighbouring cells for j > i */
        for (int i = 0; i < n; ++i) {
            int ci[3];
            for (int k = 0; k < 3; ++k) {
                double w = x[3 * i + k] - box[k] * floor(x[3 * i + k] / box[k]);
                int c = (int)(w / box[k] * g.m[k]);
                if (c >= g.m[k]) c = g.m[k] - 1;
                if (c < 0) c = 0;
                ci[k] = c;
            }
            for (int a = -1; a <= 1; ++a)
                for (int b = -1; b <= 1; ++b)
                    for (int e = -1; e <= 1; ++e) {
                        int cx = (ci[0] + a + g.m[0]) % g.m[0];
                        int cy = (ci[1] + b + g.m[1]) % g.m[1];
                        int cz = (ci[2] + e + g.m[2]) % g.m[2];
                        for (int j = g.head[(cx * g.m[1] + cy) * g.m[2] + cz]; j >= 0; j = g.next[j]) {
                            if (j <= i) continue;
                            double dx = mimg(x[3 * i] - x[3 * j], box[0]);
                            double dy = mimg(x[3 * i + 1] - x[3 * j + 1], box[1]);
                            double dz = mimg(x[3 * i + 2] - x[3 * j + 2], box[2]);
                            double r2 = dx * dx + dy * dy + dz * dz;
                            if (r2 < rc2 && !excl_has(ex, i, j)) {
                                ++cnt;
                                if (fn) fn(ctx, i, j, dx, dy, dz, r2);
                            }
                        }
                    }
        }
        grid_free(&g);
        return cnt;
    }
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) {
            double dx = mimg(x[3 * i] - x[3 * j], box[0]);
            double dy = mimg(x[3 * i + 1] - x[3 * j + 1], box[1]);
            double dz = mimg(x[3 * i + 2] - x[3 * j + 2], box[2]);
            double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 < rc2 && !excl_has(ex, i, j)) {
                ++cnt;
                if (fn) fn(ctx, i, j, dx, dy, dz, r2);
            }
        }
    return cnt;
}

static long long visit_pairs(int n, const double* x, const double* box, double rc, int brute, const excl_set* ex,
                             pair_fn fn, void* ctx) {
    return visit_pairs_tilt(n, x, box, NULL, rc, brute, ex, fn, ctx);
}

typedef struct {
    int* out;
    long long cap, k;
} pairs_ctx;

static void pairs_cb(void* vc, int i, int j, double dx, double dy, double dz, double r2) {
    (void)dx; (void)dy; (void)dz; (void)r2;
    pairs_ctx* c = (pairs_ctx*)vc;
    if (c->k < c->cap) {
        c->out[2 * c->k] = i;
        c->out[2 * c->k + 1] = j;
    }
    c->k++;
}

/* In-cutoff pairs (i<j) that are not exceptions (nexc pairs, 2 indices each), in visiting
 * order; returns the count. */
long long oracle_pairs_ex(int n, const double* x, const double* box, double rc, int brute, int nexc,
                          const int* exc, int* pairs, long long cap) {
    pairs_ctx c = {pairs, pairs ? cap : 0, 0};
    excl_set e;
    excl_build(&e, nexc, exc);
    long long r = visit_pairs(n, x, box, rc, brute, &e, pairs_cb, &c);
    free(e.key);
    return r;
}

long long oracle_pairs(int n, const double* x, const double* box, double rc, int brute, int* pairs, long long cap) {
    return oracle_pairs_ex(n, x, box, rc, brute, 0, NULL, pairs, cap);
}

/* oracle_pairs_ex in a triclinic box: edges box[3] = (a_x, b_y, c_z), tilt[3] = (b_x, c_x, c_y). */
long long oracle_pairs_tri(int n, const double* x, const double* box, const double* tilt, double rc, int nexc,
                           const int* exc, int* pairs, long long cap) {
    pairs_ctx c = {pairs, pairs ? cap : 0, 0};
    excl_set e;
    excl_build(&e, nexc, exc);
    long long r = visit_pairs_tilt(n, x, box, tilt, rc, 1, &e, pairs_cb, &c);
    free(e.key);
    return r;
}

typedef struct {
    const double *q, *sig, *eps;
    double krf, crf;
    double* F;
    nsum E, Elj, Ec, Eabs;
} nb_ctx;

static void nb_cb(void* vc, int i, int j, double dx, double dy, double dz, double r2) {
    nb_ctx* c = (nb_ctx*)vc;
    double r = sqrt(r2);
    double inv_r = 1.0 / r;
    double sig = 0.5 * (c->sig[i] + c->sig[j]);
    double eps = sqrt(c->eps[i] * c->eps[j]);
    double sr2 = sig * sig / r2;
    double sr6 = sr2 * sr2 * sr2;
    double elj = 4.0 * eps * (sr6 * sr6 - sr6);
    double qq = ORACLE_KE * c->q[i] * c->q[j];
    double ec = qq * (inv_r + c->krf * r2 - c->crf);
    /* -dE/dr / r */
    double f_over_r = 4.0 * eps * (12.0 * sr6 * sr6 - 6.0 * sr6) / r2 + qq * (inv_r / r2 - 2.0 * c->krf);
    if (c->F) {
        c->F[3 * i] += f_over_r * dx;
        c->F[3 * i + 1] += f_over_r * dy;
        c->F[3 * i + 2] += f_over_r * dz;
        c->F[3 * j] -= f_over_r * dx;
        c->F[3 * j + 1] -= f_over_r * dy;
        c->F[3 * j + 2] -= f_over_r * dz;
    }
    nsum_add(&c->E, elj + ec);
    nsum_add(&c->Elj, elj);
    nsum_add(&c->Ec, ec);
    nsum_add(&c->Eabs, fabs(elj) + fabs(ec));
}

/* The exception terms: k_e q'/r + 4 e' [(s'/r)^12 - (s'/r)^6] per listed pair at the plain
 * distance (no cutoff, no periodic image); forces added to F (may be NULL); energy added to
 * E, sum of |terms| to Eabs. */
static void exception_terms(const double* x, int nexc, const int* exc, const double* qq, const double* es,
                            const double* ee, double* F, nsum* E, nsum* Eabs) {
    for (int k = 0; k < nexc; ++k) {
        const int i = exc[2 * k], j = exc[2 * k + 1];
        if (qq[k] == 0.0 && ee[k] == 0.0) continue;  /* an exclusion */
        double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
        double r2 = dx * dx + dy * dy + dz * dz;
        double inv_r = 1.0 / sqrt(r2);
        double sr2 = es[k] * es[k] / r2;
        double sr6 = sr2 * sr2 * sr2;
        double elj = 4.0 * ee[k] * (sr6 * sr6 - sr6);
        double ec = ORACLE_KE * qq[k] * inv_r;
        double f_over_r = 4.0 * ee[k] * (12.0 * sr6 * sr6 - 6.0 * sr6) / r2 + ec / r2;
        if (F) {
            F[3 * i] += f_over_r * dx; F[3 * i + 1] += f_over_r * dy; F[3 * i + 2] += f_over_r * dz;
            F[3 * j] -= f_over_r * dx; F[3 * j + 1] -= f_over_r * dy; F[3 * j + 2] -= f_over_r * dz;
        }
        nsum_add(E, elj + ec);
        nsum_add(Eabs, fabs(elj) + fabs(ec));
    }
}

/* Energy (returned) and forces (F, n*3, overwritten; may be NULL) of the LJ +
 * reaction-field cutoff interaction with nexc exceptions (exc: 2 indices each; qq: k_e-free
 * charge products, es: sigma, ee: epsilon).  parts (may be NULL, 4 doubles) receives
 * E_LJ and E_C of the cutoff pairs, the sum of |terms| over the cutoff pairs and the
 * exceptions, and the exception energy. */
static double nonbonded_tilt(int n, const double* x, const double* q, const double* sig, const double* eps,
                             const double* box, const double* tilt, double rc, double eps_rf, int brute, int nexc,
                             const int* exc, const double* qq, const double* es, const double* ee, double* F,
                             long long* npairs, double* parts) {
    nb_ctx c;
    memset(&c, 0, sizeof(c));
    c.q = q;
    c.sig = sig;
    c.eps = eps;
    c.krf = (eps_rf - 1.0) / ((2.0 * eps_rf + 1.0) * rc * rc * rc);
    c.crf = 3.0 * eps_rf / ((2.0 * eps_rf + 1.0) * rc);
    c.F = F;
    if (F) memset(F, 0, sizeof(double) * 3 * n);
    excl_set e;
    excl_build(&e, nexc, exc);
    long long np = visit_pairs_tilt(n, x, box, tilt, rc, brute, &e, nb_cb, &c);
    free(e.key);
    nsum Ex = {0.0, 0.0};
    exception_terms(x, nexc, exc, qq, es, ee, F, &Ex, &c.Eabs);
    if (npairs) *npairs = np;
    if (parts) {
        parts[0] = c.Elj.s + c.Elj.c;
        parts[1] = c.Ec.s + c.Ec.c;
        parts[2] = c.Eabs.s + c.Eabs.c;
        parts[3] = Ex.s + Ex.c;
    }
    nsum_add(&c.E, Ex.s);
    nsum_add(&c.E, Ex.c);
    return c.E.s + c.E.c;
}

double oracle_nonbonded_ex(int n, const double* x, const double* q, const double* sig, const double* eps,
                           const double* box, double rc, double eps_rf, int brute, int nexc, const int* exc,
                           const double* qq, const double* es, const double* ee, double* F, long long* npairs,
                           double* parts) {
    return nonbonded_tilt(n, x, q, sig, eps, box, NULL, rc, eps_rf, brute, nexc, exc, qq, es, ee, F, npairs, parts);
}

double oracle_nonbonded(int n, const double* x, const double* q, const double* sig, const double* eps,
                        const double* box, double rc, double eps_rf, int brute, double* F, long long* npairs,
                        double* parts) {
    double p4[4];
    double E = oracle_nonbonded_ex(n, x, q, sig, eps, box, rc, eps_rf, brute, 0, NULL, NULL, NULL, NULL, F, npairs,
                                   p4);
    if (parts) { parts[0] = p4[0]; parts[1] = p4[1]; parts[2] = p4[2]; }
    return E;
}

/* One ATM evaluation: U1,F1 at x; U2,F2 at x + d (entry i displaces atom i);
 * combine; F (n*3, overwritten with the ATM force; may be NULL). */
static int atm_eval_tilt(int n, const double* x, const double* d, const double* q, const double* sig,
                         const double* eps, const double* box, const double* tilt, double rc, double eps_rf,
                         int nexc, const int* exc, const double* qq, const double* es, const double* ee,
                         const oracle_params* p, int brute, double* F, oracle_result* r) {
    double* x2 = (double*)malloc(sizeof(double) * 3 * n);
    double* F1 = F ? (double*)malloc(sizeof(double) * 3 * n) : NULL;
    double* F2 = F ? (double*)malloc(sizeof(double) * 3 * n) : NULL;
    if (!x2 || (F && (!F1 || !F2))) return -1;
    for (int i = 0; i < 3 * n; ++i) x2[i] = x[i] + d[i];
    long long n1 = 0, n2 = 0;
    double parts1[4], parts2[4];
    double U1 = nonbonded_tilt(n, x, q, sig, eps, box, tilt, rc, eps_rf, brute, nexc, exc, qq, es, ee, F1, &n1, parts1);
    double U2 = nonbonded_tilt(n, x2, q, sig, eps, box, tilt, rc, eps_rf, brute, nexc, exc, qq, es, ee, F2, &n2, parts2);
    oracle_combine(p, U1, U2, r);
    r->npairs1 = n1;
    r->npairs2 = n2;
    r->abs1 = parts1[2];
    r->abs2 = parts2[2];
    if (F) {
        memset(F, 0, sizeof(double) * 3 * n);
        oracle_blend(n, r->sp, p->direction, F1, F2, F);
    }
    free(x2);
    free(F1);
    free(F2);
    return 0;
}

int oracle_atm_eval_ex(int n, const double* x, const double* d, const double* q, const double* sig,
                       const double* eps, const double* box, double rc, double eps_rf, int nexc, const int* exc,
                       const double* qq, const double* es, const double* ee, const oracle_params* p, int brute,
                       double* F, oracle_result* r) {
    return atm_eval_tilt(n, x, d, q, sig, eps, box, NULL, rc, eps_rf, nexc, exc, qq, es, ee, p, brute, F, r);
}

/* oracle_atm_eval_ex in a triclinic box (tilt as for oracle_pairs_tri; O(N^2) pair loop). */
int oracle_atm_eval_tri(int n, const double* x, const double* d, const double* q, const double* sig,
                        const double* eps, const double* box, const double* tilt, double rc, double eps_rf, int nexc,
                        const int* exc, const double* qq, const double* es, const double* ee, const oracle_params* p,
                        double* F, oracle_result* r) {
    return atm_eval_tilt(n, x, d, q, sig, eps, box, tilt, rc, eps_rf, nexc, exc, qq, es, ee, p, 1, F, r);
}

int oracle_atm_eval(int n, const double* x, const double* d, const double* q, const double* sig, const double* eps,
                    const double* box, double rc, double eps_rf, const oracle_params* p, int brute, double* F,
                    oracle_result* r) {
    return oracle_atm_eval_ex(n, x, d, q, sig, eps, box, rc, eps_rf, 0, NULL, NULL, NULL, NULL, p, brute, F, r);
}

/* ------------------------------------------------------------------ gradients */

typedef struct {
    const double *q, *sig, *eps;
    double krf, crf, w;
    double *dq, *dsig, *deps;
} gr_ctx;

static void gr_cb(void* vc, int i, int j, double dx, double dy, double dz, double r2) {
    (void)dx; (void)dy; (void)dz;
    gr_ctx* c = (gr_ctx*)vc;
    double r = sqrt(r2);
    double sig = 0.5 * (c->sig[i] + c->sig[j]);
    double eij = sqrt(c->eps[i] * c->eps[j]);
    double sr2 = sig * sig / r2;
    double sr6 = sr2 * sr2 * sr2;
    double phi = ORACLE_KE * (1.0 / r + c->krf * r2 - c->crf);
    c->dq[i] += c->w * phi * c->q[j];
    c->dq[j] += c->w * phi * c->q[i];
    /* dE/dsig_ij = 4 eps (12 sr12 - 6 sr6)/sig = 4 eps (12 sig^11/r^12 - 6 sig^5/r^6), which
     * is 0 at sig_ij = 0 (the quotient form is 0/0 there); dsig_ij/dsig_i = 1/2 */
    double dsg = sig > 0.0 ? 0.5 * 4.0 * eij * (12.0 * sr6 * sr6 - 6.0 * sr6) / sig : 0.0;
    c->dsig[i] += c->w * dsg;
    c->dsig[j] += c->w * dsg;
    /* dE/deps_i = E_LJ / (2 eps_i) = 2 sqrt(eps_j/eps_i) (sr12 - sr6).  At eps_i = 0 the
     * derivative does not exist (sqrt(eps_i eps_j) has an infinite slope there); this
     * restatement returns 0 for such an atom, a convention the device path follows. */
    double lj = sr6 * sr6 - sr6;
    if (c->eps[i] > 0) c->deps[i] += c->w * 2.0 * sqrt(c->eps[j] / c->eps[i]) * lj;
    if (c->eps[j] > 0) c->deps[j] += c->w * 2.0 * sqrt(c->eps[i] / c->eps[j]) * lj;
}

/* dW/d(q, sigma, eps) (n each, overwritten) and dW/d(lambda1, lambda2, alpha, u0, w0)
 * (dparams, 5) at the evaluation described by (x, d, ..., p), using the result r of
 * oracle_atm_eval.  dW/dU1 and dW/dU2 use sp, the weight the reference's force uses. */
void oracle_atm_grad_ex(int n, const double* x, const double* d, const double* q, const double* sig,
                        const double* eps, const double* box, double rc, double eps_rf, int nexc, const int* exc,
                        const oracle_params* p, const oracle_result* r, double* dq, double* dsig, double* deps,
                        double* dparams) {
    double* x2 = (double*)malloc(sizeof(double) * 3 * n);
    for (int i = 0; i < 3 * n; ++i) x2[i] = x[i] + d[i];
    memset(dq, 0, sizeof(double) * n);
    memset(dsig, 0, sizeof(double) * n);
    memset(deps, 0, sizeof(double) * n);
    gr_ctx c;
    c.q = q; c.sig = sig; c.eps = eps;
    c.krf = (eps_rf - 1.0) / ((2.0 * eps_rf + 1.0) * rc * rc * rc);
    c.crf = 3.0 * eps_rf / ((2.0 * eps_rf + 1.0) * rc);
    c.dq = dq; c.dsig = dsig; c.deps = deps;
    int fwd = p->direction > 0;
    /* exception terms carry their own parameters: no per-atom derivative */
    excl_set e;
    excl_build(&e, nexc, exc);
    c.w = fwd ? 1.0 - r->sp : r->sp; /* dW/dU1 */
    visit_pairs(n, x, box, rc, 0, &e, gr_cb, &c);
    c.w = fwd ? r->sp : 1.0 - r->sp; /* dW/dU2 */
    visit_pairs(n, x2, box, rc, 0, &e, gr_cb, &c);
    free(e.key);
    free(x2);
    if (dparams) {
        double usc = r->usc;
        double ee = 1.0 + exp(-p->alpha * (usc - p->u0));
        double L = log(ee);
        double dl1 = 0, dl2 = usc, da = 0, du0 = 0;
        if (p->alpha > 0) {
            dl1 = -L / p->alpha;
            dl2 += L / p->alpha;
            du0 = (p->lambda2 - p->lambda1) * (1.0 - 1.0 / ee);
            da = (p->lambda2 - p->lambda1) * (-L / (p->alpha * p->alpha) - (usc - p->u0) * (1.0 - 1.0 / ee) / p->alpha);
        }
        dparams[0] = dl1;
        dparams[1] = dl2;
        dparams[2] = da;
        dparams[3] = du0;
        dparams[4] = 1.0;
    }
}

void oracle_atm_grad(int n, const double* x, const double* d, const double* q, const double* sig, const double* eps,
                     const double* box, double rc, double eps_rf, const oracle_params* p, const oracle_result* r,
                     double* dq, double* dsig, double* deps, double* dparams) {
    oracle_atm_grad_ex(n, x, d, q, sig, eps, box, rc, eps_rf, 0, NULL, p, r, dq, dsig, deps, dparams);
}
