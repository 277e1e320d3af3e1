/*
 * sph_oracle.c -- FP64 CPU restatement of sigmaSPH's particle time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the parity oracle for the HIP path and the timed
 * CPU baseline of bench.py; nothing in the product (sigmasph_amd/) loads it.  It follows
 * the reference's phase structure literally -- one loop per reference kernel, in the
 * reference's order -- so that it is an independent restatement of the algorithm and
 * not a copy of the fused GPU schedule.  Every function names the reference lines it
 * follows (paths relative to the reference root).
 *
 * Deliberate, documented choices where the reference leaves something open:
 *   - FP64 throughout (the reference is float32: run_simulation.py:20).
 *   - Canonical neighbour predicate r2 < h*h with r2 = dx*dx + dy*dy + dz*dz evaluated
 *     without contraction (the reference tests norm(r) < h in f32,
 *     particle_system.py:1406).
 *   - Within a cell, particles are ordered by ascending stable id (the reference's
 *     counting sort leaves that order to the scheduler, particle_system.py:1328).
 *   - Cells are visited z-outer, y, x-inner; then the y, z and corner image sweeps.
 *   - The 3x3 symmetric eigenproblem (ti.sym_eig, third-party Taichi, absent here) is
 *     solved by cyclic Jacobi rotations.
 *
 * Build: make -C oracle  (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "../include/sigmasph.h"

/* ------------------------------------------------------------------------------------ */
/* state                                                                                */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    int64_t n;
    int32_t *pid, *obj;
    double *x, *x0, *u, *v, *shift;      /* [n][3] */
    double *V, *m, *p;                   /* [n]    */
    double *acc, *lam_grad;              /* [n][3] */
    double *lam, *div, *de_visc;         /* [n]    */
    double *L6;                          /* [n][6] (00,11,22,01,02,12) */
    double *scal;                        /* [SPH_NSCAL] */
} or_state;

typedef struct {
    int64_t n;
    int32_t *key;     /* bucket of each particle (sorted order)            */
    int32_t *cell;    /* [n][3] binned cell of each particle               */
    int64_t nbuck;    /* n_box + 2^hash_bits + 1 (last: off-grid)          */
    int64_t *start;   /* [nbuck + 1]                                        */
} or_grid;

/* ------------------------------------------------------------------------------------ */
/* cells and keys: particle_system.py:1220-1234 (pos_to_index, on_grid), 1283-1293      */
/* ------------------------------------------------------------------------------------ */
static int or_isfinite3(const double* x) { return isfinite(x[0]) && isfinite(x[1]) && isfinite(x[2]); }

static int32_t or_cell1(const sph_dims* d, int k, double x) {
    double c = floor((x - d->origin[k]) / d->cell[k]);
    if (c > 1073741823.0) c = 1073741823.0;
    if (c < -1073741823.0) c = -1073741823.0;
    int32_t ci = (int32_t)c;
    if (d->flags & (SPH_PERIODIC_X << k)) {
        /* positions are wrapped into [origin, origin + period); guard the rounding */
        if (ci < 0) ci = 0;
        if (ci >= d->box_n[k]) ci = d->box_n[k] - 1;
    }
    return ci;
}

static uint32_t or_hash(const int32_t* c, int32_t bits) {
    /* the XOR hash of particle_system.py:1283-1293 */
    uint32_t h = ((uint32_t)c[0] * 73856093u) ^ ((uint32_t)c[1] * 19349663u)
               ^ ((uint32_t)c[2] * 83492791u);
    return h & ((1u << bits) - 1u);
}

static int64_t or_nbox(const sph_dims* d) {
    return (int64_t)d->box_n[0] * d->box_n[1] * d->box_n[2];
}

/* bucket of a cell: dense index inside the box, else the hashed overflow table */
static int64_t or_bucket(const sph_dims* d, const int32_t* c) {
    int in = 1;
    for (int k = 0; k < 3; ++k)
        if (c[k] < d->box_lo[k] || c[k] >= d->box_lo[k] + d->box_n[k]) in = 0;
    if (in)
        return ((int64_t)(c[2] - d->box_lo[2]) * d->box_n[1] + (c[1] - d->box_lo[1])) * d->box_n[0]
               + (c[0] - d->box_lo[0]);
    return or_nbox(d) + or_hash(c, d->hash_bits);
}

/* ------------------------------------------------------------------------------------ */
/* the sort: update_grid_id + prefix sum + counting_sort (particle_system.py:1300-1345)  */
/* ------------------------------------------------------------------------------------ */
typedef struct { int64_t key; int32_t pid; int64_t idx; } or_sortrec;

static int or_cmp(const void* a, const void* b) {
    const or_sortrec* p = (const or_sortrec*)a;
    const or_sortrec* q = (const or_sortrec*)b;
    if (p->key != q->key) return p->key < q->key ? -1 : 1;
    if (p->pid != q->pid) return p->pid < q->pid ? -1 : 1;
    return 0;
}

static void or_permute_d(double* a, const int64_t* perm, int64_t n, int w) {
    double* t = (double*)malloc(sizeof(double) * n * w);
    for (int64_t k = 0; k < n; ++k)
        for (int c = 0; c < w; ++c) t[k * w + c] = a[perm[k] * w + c];
    memcpy(a, t, sizeof(double) * n * w);
    free(t);
}
static void or_permute_i(int32_t* a, const int64_t* perm, int64_t n) {
    int32_t* t = (int32_t*)malloc(sizeof(int32_t) * n);
    for (int64_t k = 0; k < n; ++k) t[k] = a[perm[k]];
    memcpy(a, t, sizeof(int32_t) * n);
    free(t);
}

/* Sorts the state in place by (bucket, pid) and fills the grid.  Returns the number of
 * particles quarantined off-grid. */
int64_t oracle_sort(const sph_dims* d, or_state* s, or_grid* g) {
    int64_t n = s->n;
    int64_t nbox = or_nbox(d);
    int64_t T = (int64_t)1 << d->hash_bits;
    int64_t offgrid = nbox + T;
    or_sortrec* r = (or_sortrec*)malloc(sizeof(or_sortrec) * n);
    int32_t* cell = (int32_t*)malloc(sizeof(int32_t) * 3 * n);
    int64_t noff = 0, nout = 0;
    double moff = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double* xi = s->x + 3 * i;
        int32_t c[3];
        for (int k = 0; k < 3; ++k) c[k] = isfinite(xi[k]) ? or_cell1(d, k, xi[k]) : 0;
        memcpy(cell + 3 * i, c, sizeof c);
        int64_t key;
        if (!or_isfinite3(xi)) { key = offgrid; ++noff; moff += s->m[i]; }
        else { key = or_bucket(d, c); if (key >= nbox) ++nout; }
        r[i].key = key; r[i].pid = s->pid[i]; r[i].idx = i;
    }
    qsort(r, n, sizeof(or_sortrec), or_cmp);
    int64_t* perm = (int64_t*)malloc(sizeof(int64_t) * n);
    for (int64_t k = 0; k < n; ++k) perm[k] = r[k].idx;
    or_permute_i(s->pid, perm, n);
    or_permute_i(s->obj, perm, n);
    or_permute_d(s->x, perm, n, 3);
    or_permute_d(s->x0, perm, n, 3);
    or_permute_d(s->u, perm, n, 3);
    or_permute_d(s->v, perm, n, 3);
    or_permute_d(s->shift, perm, n, 3);
    or_permute_d(s->V, perm, n, 1);
    or_permute_d(s->m, perm, n, 1);
    or_permute_d(s->p, perm, n, 1);
    free(g->key); free(g->cell); free(g->start);
    g->n = n;
    g->nbuck = offgrid + 1;
    g->key = (int32_t*)malloc(sizeof(int32_t) * n);
    g->cell = (int32_t*)malloc(sizeof(int32_t) * 3 * n);
    g->start = (int64_t*)calloc(g->nbuck + 1, sizeof(int64_t));
    for (int64_t k = 0; k < n; ++k) {
        g->key[k] = (int32_t)r[k].key;
        memcpy(g->cell + 3 * k, cell + 3 * r[k].idx, 3 * sizeof(int32_t));
        g->start[r[k].key + 1]++;
    }
    for (int64_t b = 0; b < g->nbuck; ++b) g->start[b + 1] += g->start[b];
    s->scal[SPH_S_OFF_GRID_NUM] = (double)noff;
    s->scal[SPH_S_OFF_GRID_MASS] = moff;
    s->scal[SPH_S_OUT_OF_BOX] = (double)nout;
    free(r); free(cell); free(perm);
    return noff;
}

/* ------------------------------------------------------------------------------------ */
/* neighbour traversal: for_all_neighbors (particle_system.py:1348-1490)                 */
/* ------------------------------------------------------------------------------------ */
typedef void (*or_task)(void* ctx, int64_t i, int64_t j, int mir, const double* r);

/* minimum image on periodic axes (a build-side extension: the reference has none) */
static void or_min_image(const sph_dims* d, double* r) {
    for (int k = 0; k < 3; ++k) {
        if (!(d->flags & (SPH_PERIODIC_X << k))) continue;
        double L = d->period[k];
        if (r[k] > 0.5 * L) r[k] -= L;
        else if (r[k] < -0.5 * L) r[k] += L;
    }
}

static double or_dist2(const double* r) {
    /* no contraction: built with -ffp-contract=off */
    double a = r[0] * r[0];
    double b = r[1] * r[1];
    double c = r[2] * r[2];
    return (a + b) + c;
}

static int32_t or_wrap_cell(const sph_dims* d, int k, int32_t c) {
    if (d->flags & (SPH_PERIODIC_X << k)) {
        int32_t nk = d->box_n[k];
        c %= nk; if (c < 0) c += nk;
    }
    return c;
}

/* visit every particle binned in cell c; `xt` is the point tested against (x_i, or its
 * mirror image), r = x_i - M x_j is what the task receives */
static void or_visit_cell(const sph_dims* d, const or_state* s, const or_grid* g,
                          int64_t i, const int32_t* c, int mir, const double* xt, double h2,
                          or_task task, void* ctx) {
    int64_t b = or_bucket(d, c);
    int checked = b >= or_nbox(d);
    for (int64_t j = g->start[b]; j < g->start[b + 1]; ++j) {
        if (checked) {
            const int32_t* cj = g->cell + 3 * j;
            if (cj[0] != c[0] || cj[1] != c[1] || cj[2] != c[2]) continue;
        }
        if (mir == 0 && j == i) continue;
        const double* xj = s->x + 3 * j;
        double t[3] = { xt[0] - xj[0], xt[1] - xj[1], xt[2] - xj[2] };
        or_min_image(d, t);
        if (!(or_dist2(t) < h2)) continue;
        /* r = x_i - M x_j */
        double r[3];
        const double* xi = s->x + 3 * i;
        double sy = (mir & 1) ? -1.0 : 1.0, sz = (mir & 2) ? -1.0 : 1.0;
        r[0] = xi[0] - xj[0];
        r[1] = xi[1] - sy * xj[1];
        r[2] = xi[2] - sz * xj[2];
        or_min_image(d, r);
        task(ctx, i, j, mir, r);
    }
}

void or_for_all_neighbors(const sph_dims* d, const or_state* s, const or_grid* g, int64_t i,
                          or_task task, void* ctx) {
    const double* xi = s->x + 3 * i;
    if (!or_isfinite3(xi)) return;                    /* alive gate, 1381 */
    double h2 = d->h * d->h;
    int32_t cc[3];
    for (int k = 0; k < 3; ++k) cc[k] = or_cell1(d, k, xi[k]);
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int32_t c[3] = { or_wrap_cell(d, 0, cc[0] + dx), or_wrap_cell(d, 1, cc[1] + dy),
                                 or_wrap_cell(d, 2, cc[2] + dz) };
                or_visit_cell(d, s, g, i, c, 0, xi, h2, task, ctx);
            }
    /* images across y = 0 (1419-1440): only the row of cells at y-index 0 */
    if ((d->flags & SPH_MIRROR_Y) && xi[1] < d->h) {
        double xm[3] = { xi[0], -xi[1], xi[2] };
        for (int dz = -1; dz <= 1; ++dz)
            for (int dx = -1; dx <= 1; ++dx) {
                int32_t c[3] = { or_wrap_cell(d, 0, cc[0] + dx), 0, or_wrap_cell(d, 2, cc[2] + dz) };
                or_visit_cell(d, s, g, i, c, 1, xm, h2, task, ctx);
            }
    }
    if ((d->flags & SPH_MIRROR_Z) && xi[2] < d->h) {             /* 1442-1463 */
        double xm[3] = { xi[0], xi[1], -xi[2] };
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                int32_t c[3] = { or_wrap_cell(d, 0, cc[0] + dx), or_wrap_cell(d, 1, cc[1] + dy), 0 };
                or_visit_cell(d, s, g, i, c, 2, xm, h2, task, ctx);
            }
    }
    if ((d->flags & SPH_MIRROR_Y) && (d->flags & SPH_MIRROR_Z) &&  /* 1469-1490 */
        xi[1] < d->h && xi[2] < d->h) {
        double xm[3] = { xi[0], -xi[1], -xi[2] };
        for (int dx = -1; dx <= 1; ++dx) {
            int32_t c[3] = { or_wrap_cell(d, 0, cc[0] + dx), 0, 0 };
            or_visit_cell(d, s, g, i, c, 3, xm, h2, task, ctx);
        }
    }
}

/* neighbour pairs as (pid_i, pid_j, mir); returns the count (writes at most cap) */
typedef struct { const or_state* s; int32_t* out; int64_t cap; int64_t cnt; } or_pairctx;
static void or_pair_task(void* c, int64_t i, int64_t j, int mir, const double* r) {
    (void)r;
    or_pairctx* p = (or_pairctx*)c;
    if (p->cnt < p->cap) {
        p->out[3 * p->cnt + 0] = p->s->pid[i];
        p->out[3 * p->cnt + 1] = p->s->pid[j];
        p->out[3 * p->cnt + 2] = mir;
    }
    p->cnt++;
}

int64_t oracle_neighbor_pairs(const sph_dims* d, const or_state* s, const or_grid* g,
                              int32_t* out, int64_t cap) {
    or_pairctx c = { s, out, cap, 0 };
    for (int64_t i = 0; i < s->n; ++i) or_for_all_neighbors(d, s, g, i, or_pair_task, &c);
    return c.cnt;
}

/* ------------------------------------------------------------------------------------ */
/* Wendland C2 kernel and gradient (sph_base.py:42-97)                                   */
/* ------------------------------------------------------------------------------------ */
static double or_alpha(const sph_dims* d) {
    double h = d->h;
    return (d->flags & SPH_TWO_D) ? 7.0 / (M_PI * h * h) : 21.0 / (2.0 * M_PI * h * h * h);
}

double oracle_wendland(const sph_dims* d, double r) {
    double h = d->h, q = r / h;
    if (q < 1.0) {
        double a = 1.0 - q;
        return or_alpha(d) * a * a * a * a * (4.0 * q + 1.0);
    }
    return 0.0;
}

void oracle_wendland_grad(const sph_dims* d, const double* r, double* g) {
    double h = d->h;
    double rn = sqrt(or_dist2(r));
    double q = rn / h;
    g[0] = g[1] = g[2] = 0.0;
    if (q < 1.0 && rn > 1.0e-5) {
        double a = 1.0 - q;
        double f = -(20.0 * or_alpha(d) / h * q * a * a * a);
        for (int k = 0; k < 3; ++k) g[k] = f * r[k] / rn;
    }
}

/* ------------------------------------------------------------------------------------ */
/* symmetric eigenproblem and the regularised inverse (spd_inverse.py)                   */
/* ------------------------------------------------------------------------------------ */
/* cyclic Jacobi on a symmetric 3x3: A is overwritten with its eigenvalues on the diagonal,
 * columns of Vm are the eigenvectors */
void oracle_jacobi3(double A[3][3], double Vm[3][3]) {
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) Vm[a][b] = (a == b) ? 1.0 : 0.0;
    static const int P[3] = { 0, 0, 1 }, Q[3] = { 1, 2, 2 };
    for (int sweep = 0; sweep < 32; ++sweep) {
        double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        double dia = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off <= 1.0e-32 * dia || off == 0.0) break;
        for (int r = 0; r < 3; ++r) {
            int p = P[r], q = Q[r];
            double apq = A[p][q];
            if (apq == 0.0) continue;
            double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
            double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
            for (int k = 0; k < 3; ++k) {               /* A <- A J */
                double akp = A[k][p], akq = A[k][q];
                A[k][p] = c * akp - sn * akq;
                A[k][q] = sn * akp + c * akq;
            }
            for (int k = 0; k < 3; ++k) {               /* A <- J^T A */
                double apk = A[p][k], aqk = A[q][k];
                A[p][k] = c * apk - sn * aqk;
                A[q][k] = sn * apk + c * aqk;
            }
            A[p][q] = A[q][p] = 0.0;
            for (int k = 0; k < 3; ++k) {               /* V <- V J */
                double vkp = Vm[k][p], vkq = Vm[k][q];
                Vm[k][p] = c * vkp - sn * vkq;
                Vm[k][q] = sn * vkp + c * vkq;
            }
        }
    }
}

/* _weight (spd_inverse.py:45-50) */
double oracle_spd_weight(double e, double tol) {
    double u = (e - tol) / tol;
    if (u < 0.0) u = 0.0;
    if (u > 1.0) u = 1.0;
    double beta = u * u * (3.0 - 2.0 * u);
    return beta / (e > tol ? e : tol) + (1.0 - beta);
}

/* spd_inverse_3x3 (spd_inverse.py:87-101) and min eigenvalue of the symmetrised E
 * (DSPH.py:680-686).  E is row-major 3x3; L out as (00,11,22,01,02,12). */
void oracle_spd_inverse3(const double* E, double tol, double* L6, double* min_eig) {
    double A[3][3], Vm[3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) A[a][b] = 0.5 * (E[3 * a + b] + E[3 * b + a]);
    oracle_jacobi3(A, Vm);
    double L[3][3] = { { 0 } };
    double me = A[0][0];
    for (int k = 0; k < 3; ++k) {
        if (A[k][k] < me) me = A[k][k];
        double w = oracle_spd_weight(A[k][k], tol);
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) L[a][b] += w * Vm[a][k] * Vm[b][k];
    }
    L6[0] = L[0][0]; L6[1] = L[1][1]; L6[2] = L[2][2];
    L6[3] = L[0][1]; L6[4] = L[0][2]; L6[5] = L[1][2];
    *min_eig = me;
}

/* eig_sym_2x2 (DSPH.py:35-47): smaller eigenvalue of [[a, b], [b, d]] */
static double or_eig2_min(double a, double b, double d) {
    double det = sqrt((a - d) * (a - d) + 4.0 * b * b);
    return 0.5 * (a + d - det);
}

/* spd_inverse_2x2 (spd_inverse.py:53-84), embedded in 3x3 with L22 = 1 (DSPH.py:663-668) */
void oracle_spd_inverse2(const double* E2, double tol, double* L6) {
    double a = E2[0], b = 0.5 * (E2[1] + E2[2]), d = E2[3];
    double disc = sqrt(fmax((a - d) * (a - d) + 4.0 * b * b, 0.0));
    double e1 = 0.5 * (a + d + disc), e2 = 0.5 * (a + d - disc);
    double n1[2] = { 1.0, 0.0 };
    double scale = fabs(a) + fabs(d) + 1e-30;
    if (disc > 1e-9 * scale) {
        double v1[2] = { b, e1 - a }, v2[2] = { e1 - d, b };
        double l1 = sqrt(v1[0] * v1[0] + v1[1] * v1[1]);
        double l2 = sqrt(v2[0] * v2[0] + v2[1] * v2[1]);
        if (l1 >= l2) { n1[0] = v1[0] / l1; n1[1] = v1[1] / l1; }
        else          { n1[0] = v2[0] / l2; n1[1] = v2[1] / l2; }
    }
    double n2[2] = { -n1[1], n1[0] };
    double w1 = oracle_spd_weight(e1, tol), w2 = oracle_spd_weight(e2, tol);
    L6[0] = w1 * n1[0] * n1[0] + w2 * n2[0] * n2[0];
    L6[1] = w1 * n1[1] * n1[1] + w2 * n2[1] * n2[1];
    L6[2] = 1.0;
    L6[3] = w1 * n1[0] * n1[1] + w2 * n2[0] * n2[1];
    L6[4] = 0.0;
    L6[5] = 0.0;
}

/* ------------------------------------------------------------------------------------ */
/* EOS: tait_pressure (EOS001_tait.py:49-55) with the clamp of DSPH.py:456-468            */
/* ------------------------------------------------------------------------------------ */
static double or_pressure(const sph_fluid_params* P, double rho) {
    double p = P->stiffness * (pow(rho / P->rho0, P->gamma) - 1.0);
    if (!P->allow_negative_pressure && p < 0.0) p = 0.0;
    return p;
}

/* tait_specific_energy (EOS001_tait.py:102-121) */
double oracle_tait_energy(double rho, double rho0, double S, double gamma) {
    double x = rho / rho0;
    if (fabs(gamma - 1.0) < 1.0e-6) return (S / rho0) * (log(x) + 1.0 / x - 1.0);
    return (S / rho0) * ((pow(x, gamma - 1.0) - 1.0) / (gamma - 1.0) + 1.0 / x - 1.0);
}

/* quadrature weight and its mass (particle_system.py:990-1004) */
static double or_w(const sph_dims* d, const sph_fluid_params* P, const or_state* s, int64_t p) {
    if (d->flags & SPH_AXISYMMETRIC) return s->V[p] / fmax(s->x[3 * p + 1], P->axi_r_min);
    return s->V[p];
}
static double or_mw(const sph_dims* d, const sph_fluid_params* P, const or_state* s, int64_t p) {
    if (d->flags & SPH_AXISYMMETRIC) return s->m[p] / fmax(s->x[3 * p + 1], P->axi_r_min);
    return s->m[p];
}
static void or_mirror(int mir, double* a) {
    if (mir & 1) a[1] = -a[1];
    if (mir & 2) a[2] = -a[2];
}

/* ------------------------------------------------------------------------------------ */
/* the fluid phases                                                                      */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    const sph_dims* d; const sph_fluid_params* P; or_state* s;
    double acc[9];     /* generic accumulator */
    double dt;
} or_tctx;

/* compute_L_task (DSPH.py:563-589) */
static void or_L_task(void* c, int64_t i, int64_t j, int mir, const double* r) {
    (void)mir;
    or_tctx* t = (or_tctx*)c;
    double gW[3];
    oracle_wendland_grad(t->d, r, gW);
    double Vj = or_w(t->d, t->P, t->s, j);
    double dx[3] = { -r[0], -r[1], -r[2] };            /* x_j - x_i */
    if (t->d->flags & SPH_TWO_D) {
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) t->acc[2 * a + b] += Vj * (dx[a] * gW[b]);
    } else {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) t->acc[3 * a + b] += Vj * (dx[a] * gW[b]);
    }
}

/* compute_lam_grad_task (DSPH.py:591-609) */
static void or_lamgrad_task(void* c, int64_t i, int64_t j, int mir, const double* r) {
    (void)mir;
    or_tctx* t = (or_tctx*)c;
    double gW[3];
    oracle_wendland_grad(t->d, r, gW);
    double Vj = or_w(t->d, t->P, t->s, j);
    double ld = t->s->lam[j] - t->s->lam[i];
    const double* L = t->s->L6 + 6 * i;
    double Lg[3] = { L[0] * gW[0] + L[3] * gW[1] + L[4] * gW[2],
                     L[3] * gW[0] + L[1] * gW[1] + L[5] * gW[2],
                     L[4] * gW[0] + L[5] * gW[1] + L[2] * gW[2] };
    for (int k = 0; k < 3; ++k) t->acc[k] += Vj * ld * Lg[k];
}

/* compute_lam (DSPH.py:611-702) */
static void or_compute_lam(const sph_dims* d, const sph_fluid_params* P, or_state* s,
                           const or_grid* g) {
    int64_t n = s->n;
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        or_tctx t = { d, P, s, { 0 }, 0.0 };
        or_for_all_neighbors(d, s, g, i, or_L_task, &t);
        double lam;
        if (d->flags & SPH_TWO_D) {
            double e2 = or_eig2_min(t.acc[0], t.acc[1], t.acc[3]);
            lam = fmin(1.0, fmax(0.0, e2));
            oracle_spd_inverse2(t.acc, P->l_eig_tol_lam, s->L6 + 6 * i);
        } else {
            double me;
            oracle_spd_inverse3(t.acc, P->l_eig_tol_lam, s->L6 + 6 * i, &me);
            lam = fmin(1.0, fmax(0.0, me));
        }
        s->lam[i] = lam;
    }
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        or_tctx t = { d, P, s, { 0 }, 0.0 };
        or_for_all_neighbors(d, s, g, i, or_lamgrad_task, &t);
        if (d->flags & SPH_TWO_D) t.acc[2] = 0.0;
        for (int k = 0; k < 3; ++k) s->lam_grad[3 * i + k] = t.acc[k];
    }
}

/* compute_depsvol_dt_task (DSPH.py:1208-1338); gamma-SPH is not part of this path */
static void or_depsvol_task(void* c, int64_t i, int64_t j, int mir, const double* r) {
    or_tctx* t = (or_tctx*)c;
    const or_state* s = t->s;
    const sph_fluid_params* P = t->P;
    double gW[3];
    oracle_wendland_grad(t->d, r, gW);
    double vj[3] = { s->v[3 * j], s->v[3 * j + 1], s->v[3 * j + 2] };
    or_mirror(mir, vj);
    double rho_i = s->m[i] / s->V[i], rho_j = s->m[j] / s->V[j];
    double vi[3] = { s->v[3 * i], s->v[3 * i + 1], s->v[3 * i + 2] };
    double wj = or_w(t->d, P, s, j);
    if (P->pst_enabled) {
        double dui[3], duj[3];
        for (int k = 0; k < 3; ++k) { dui[k] = s->shift[3 * i + k] / t->dt; duj[k] = s->shift[3 * j + k] / t->dt; }
        or_mirror(mir, duj);
        for (int k = 0; k < 3; ++k) { vi[k] += dui[k]; vj[k] += duj[k]; }
        double fl = 0.0;
        for (int k = 0; k < 3; ++k) fl += (rho_j * duj[k] - rho_i * dui[k]) * gW[k];
        t->acc[0] += wj * fl / rho_i;
    }
    double cont = 0.0;
    for (int k = 0; k < 3; ++k) cont += (vi[k] - vj[k]) * gW[k];
    t->acc[0] += wj * cont;
    if (P->delta_sph > 0.0) {
        double h = t->d->h, eta2 = 0.01 * h * h;
        double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        double pg = 0.0;
        for (int k = 0; k < 3; ++k) pg += (2.0 * (rho_j - rho_i) * (-r[k]) / (r2 + eta2)) * gW[k];
        double diffusion = P->delta_sph * h * P->c0 * wj * pg;
        double diff_rate = diffusion / rho_i;
        t->acc[0] += diff_rate;
        t->acc[1] += diff_rate;
    }
}

/* compute_volume_evolution (DSPH.py:1432-1495) */
static void or_volume(const sph_dims* d, const sph_fluid_params* P, or_state* s, const or_grid* g) {
    int64_t n = s->n;
    double dt = s->scal[SPH_S_DT];
    double* wd = (double*)malloc(sizeof(double) * n);
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        or_tctx t = { d, P, s, { 0 }, dt };
        or_for_all_neighbors(d, s, g, i, or_depsvol_task, &t);
        s->div[i] = t.acc[0];
        if (d->flags & SPH_AXISYMMETRIC) s->div[i] -= s->v[3 * i + 1] / fmax(s->x[3 * i + 1], P->axi_r_min);
        wd[i] = -s->V[i] * s->p[i] * t.acc[1];
    }
    double w = 0.0;
    for (int64_t i = 0; i < n; ++i) w += wd[i];
    for (int64_t i = 0; i < n; ++i) s->V[i] += dt * (-s->V[i] * s->div[i]);
    s->scal[SPH_S_DIFF_WORK] += dt * w;
    free(wd);
}

/* compute_pressure_accel_task (DSPH.py:1537-1640) */
static void or_press_task(void* c, int64_t i, int64_t j, int mir, const double* r) {
    or_tctx* t = (or_tctx*)c;
    const or_state* s = t->s;
    const sph_fluid_params* P = t->P;
    const sph_dims* d = t->d;
    double gW[3];
    oracle_wendland_grad(d, r, gW);
    double vj[3] = { s->v[3 * j], s->v[3 * j + 1], s->v[3 * j + 2] };
    or_mirror(mir, vj);
    const double* vi = s->v + 3 * i;
    double wi = or_w(d, P, s, i), wj = or_w(d, P, s, j), mwi = or_mw(d, P, s, i);
    double fp = -wi * wj * (s->p[i] + s->p[j]);
    for (int k = 0; k < 3; ++k) t->acc[k] += fp * gW[k] / mwi;
    double dd = 2.0 * (3 + 2);                       /* ps.dim is 3 in every mode */
    double vxy = 0.0;
    for (int k = 0; k < 3; ++k) vxy += (vi[k] - vj[k]) * r[k];
    double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    double eta2 = 0.01 * d->h * d->h;
    double fv[3];
    for (int k = 0; k < 3; ++k) fv[k] = dd * P->viscosity * wj * vxy / (r2 + eta2) * gW[k];
    for (int k = 0; k < 3; ++k) t->acc[k] += fv[k];
    if (P->track_visc_work) t->acc[3] += vi[0] * fv[0] + vi[1] * fv[1] + vi[2] * fv[2];
    if (P->pst_enabled && P->pst_ale_momentum) {
        double dui[3], duj[3];
        for (int k = 0; k < 3; ++k) { dui[k] = s->shift[3 * i + k] / t->dt; duj[k] = s->shift[3 * j + k] / t->dt; }
        or_mirror(mir, duj);
        double rho_i = s->m[i] / s->V[i], rho_j = s->m[j] / s->V[j];
        double djg = duj[0] * gW[0] + duj[1] * gW[1] + duj[2] * gW[2];
        double dig = dui[0] * gW[0] + dui[1] * gW[1] + dui[2] * gW[2];
        double drd = 0.0;
        for (int k = 0; k < 3; ++k) drd += (rho_j * duj[k] - rho_i * dui[k]) * gW[k];
        for (int k = 0; k < 3; ++k) {
            double dv = (rho_j * djg) * vj[k] - (rho_i * dig) * vi[k];
            t->acc[k] += wj * (dv - vi[k] * drd) / rho_i;
        }
    }
}

/* compute_dt_task (DSPH.py:1340-1363) */
static void or_dt_task(void* c, int64_t i, int64_t j, int mir, const double* r) {
    (void)mir; (void)r;
    or_tctx* t = (or_tctx*)c;
    const or_state* s = t->s;
    double dp = fabs(s->p[i] - s->p[j]);
    double drho = fabs(s->m[i] / s->V[i] - s->m[j] / s->V[j]);
    if (drho > 1e-3 * t->P->rho0) {
        double c2 = dp / drho;
        if (c2 > t->acc[0]) t->acc[0] = c2;
    }
}

static void or_pressure_accelerations(const sph_dims* d, const sph_fluid_params* P, or_state* s,
                                      const or_grid* g) {
    int64_t n = s->n;
    double dt = s->scal[SPH_S_DT];
    for (int64_t i = 0; i < n; ++i) {                 /* compute_non_pressure_forces 1642-1650 */
        for (int k = 0; k < 3; ++k) s->acc[3 * i + k] = P->g[k];
        s->p[i] = or_pressure(P, s->m[i] / s->V[i]);  /* 1514-1522 */
        s->de_visc[i] = 0.0;
    }
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        or_tctx t = { d, P, s, { 0 }, dt };
        or_for_all_neighbors(d, s, g, i, or_press_task, &t);
        for (int k = 0; k < 3; ++k) s->acc[3 * i + k] += t.acc[k];
        s->de_visc[i] = t.acc[3];
    }
}

/* compute_adaptive_dt (DSPH.py:1365-1430) */
static void or_adaptive_dt(const sph_dims* d, const sph_fluid_params* P, or_state* s, const or_grid* g) {
    int64_t n = s->n;
    double* c2 = (double*)malloc(sizeof(double) * n);
    double max_c2 = P->c_signal * P->c_signal, max_a = 1e-10;
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        or_tctx t = { d, P, s, { 0 }, 0.0 };
        t.acc[0] = P->c_signal * P->c_signal;
        if (P->dt_pair) or_for_all_neighbors(d, s, g, i, or_dt_task, &t);
        c2[i] = t.acc[0];
    }
    for (int64_t i = 0; i < n; ++i) {
        if (or_isfinite3(s->x + 3 * i)) {
            const double* a = s->acc + 3 * i;
            double am = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
            if (am > max_a) max_a = am;
        }
        if (c2[i] > max_c2) max_c2 = c2[i];
    }
    free(c2);
    double h = d->h;
    double dt_ac = P->cfl * h / sqrt(max_c2);
    double dt_f = P->cfl * sqrt(h / max_a);
    s->scal[SPH_S_DT] = dt_ac < dt_f ? dt_ac : dt_f;
    s->scal[SPH_S_MAX_C2] = max_c2;
    s->scal[SPH_S_MAX_ACC] = max_a;
}

/* accumulate_step_work (DSPH.py:1714-1763) */
static void or_step_work(const sph_fluid_params* P, or_state* s) {
    double dt = s->scal[SPH_S_DT];
    double g2 = P->g[0] * P->g[0] + P->g[1] * P->g[1] + P->g[2] * P->g[2];
    double wv = 0.0, wl = 0.0;
    for (int64_t i = 0; i < s->n; ++i) {
        double f[3];
        for (int k = 0; k < 3; ++k) f[k] = s->acc[3 * i + k] - P->g[k];
        wl += 0.5 * s->m[i] * dt * dt * (f[0] * f[0] + f[1] * f[1] + f[2] * f[2] - g2);
        if (P->track_visc_work) wv += s->m[i] * s->de_visc[i];
    }
    s->scal[SPH_S_VISC_WORK] += dt * wv;
    s->scal[SPH_S_LEFTOVER] += wl;
}

/* wrap periodic axes into [origin, origin + period) (build-side extension) */
static void or_wrap_pos(const sph_dims* d, double* x) {
    for (int k = 0; k < 3; ++k) {
        if (!(d->flags & (SPH_PERIODIC_X << k))) continue;
        double L = d->period[k], o = d->origin[k];
        x[k] = x[k] - L * floor((x[k] - o) / L);
    }
}

/* displace (particle_system.py:1103-1109) */
static void or_displace(const sph_dims* d, const sph_fluid_params* P, or_state* s, int64_t i,
                        const double* dr) {
    if (P->displacement_positions) {
        for (int k = 0; k < 3; ++k) {
            s->u[3 * i + k] += dr[k];
            s->x[3 * i + k] = s->x0[3 * i + k] + s->u[3 * i + k];
        }
    } else {
        for (int k = 0; k < 3; ++k) s->x[3 * i + k] += dr[k];
    }
    or_wrap_pos(d, s->x + 3 * i);
}

/* advect (DSPH.py:1652-1665); velocity and displacement updates use fma as the
 * device code does, so a step's positions agree to the bit when a does */
static void or_advect(const sph_dims* d, const sph_fluid_params* P, or_state* s) {
    double dt = s->scal[SPH_S_DT];
    for (int64_t i = 0; i < s->n; ++i) {
        double dr[3];
        for (int k = 0; k < 3; ++k) {
            s->v[3 * i + k] = fma(dt, s->acc[3 * i + k], s->v[3 * i + k]);
            dr[k] = dt * s->v[3 * i + k];
        }
        if (P->displacement_positions) {
            for (int k = 0; k < 3; ++k) {
                s->u[3 * i + k] = fma(dt, s->v[3 * i + k], s->u[3 * i + k]);
                s->x[3 * i + k] = s->x0[3 * i + k] + s->u[3 * i + k];
            }
            or_wrap_pos(d, s->x + 3 * i);
        } else {
            or_displace(d, P, s, i, dr);
        }
    }
}

/* pst_pair (DSPH.py:725-769) */
static void or_pst_task(void* c, int64_t i, int64_t j, int mir, const double* r) {
    (void)mir;
    or_tctx* t = (or_tctx*)c;
    const or_state* s = t->s;
    const sph_fluid_params* P = t->P;
    const sph_dims* d = t->d;
    double rn = sqrt(or_dist2(r));
    double gW[3];
    oracle_wendland_grad(d, r, gW);
    double Wij = oracle_wendland(d, rn);
    double Wdx0 = oracle_wendland(d, P->dx0);
    double f = 0.0;
    if (Wdx0 > 1e-10) f = P->pst_R * pow(Wij / Wdx0, (double)P->pst_n);
    if (P->pst_colle) {
        double Vj = or_w(d, P, s, j);
        if (Vj > 1e-30) for (int k = 0; k < 3; ++k) t->acc[k] += (1.0 + f) * gW[k] * Vj;
    } else {
        double mj = or_mw(d, P, s, j);
        if (mj > 1e-10) {
            double rho_i = s->m[i] / s->V[i], rho_j = s->m[j] / s->V[j];
            double ra = 0.5 * (rho_i + rho_j);
            if (ra > 1e-10) for (int k = 0; k < 3; ++k) t->acc[k] += (1.0 + f) * gW[k] * (mj / ra);
        }
    }
}

/* apply_pst (DSPH.py:943-1049) with pst_constrain (888-941) */
static void or_apply_pst(const sph_dims* d, const sph_fluid_params* P, or_state* s, const or_grid* g) {
    int64_t n = s->n;
    double dt = s->scal[SPH_S_DT];
    if (P->pst_ma_mode == 1) {                        /* compute_max_velocity 712-723 */
        double mv = 0.0;
        for (int64_t i = 0; i < n; ++i)
            if (or_isfinite3(s->x + 3 * i)) {
                const double* v = s->v + 3 * i;
                double vn = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                if (vn > mv) mv = vn;
            }
        s->scal[SPH_S_MAX_V] = mv;
    }
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; ++i) {
        or_tctx t = { d, P, s, { 0 }, dt };
        or_for_all_neighbors(d, s, g, i, or_pst_task, &t);
        const double* vi = s->v + 3 * i;
        double vn = sqrt(vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]);
        double Ma;
        if (P->pst_ma_mode == 2) Ma = P->pst_ma_fixed;
        else if (P->pst_ma_mode == 1) Ma = s->scal[SPH_S_MAX_V] / P->c0;
        else Ma = vn / P->c0;
        double h = d->h, dr[3];
        if (P->pst_colle) {
            double dv[3];
            for (int k = 0; k < 3; ++k) dv[k] = -2.0 * h * P->c0 * Ma * t.acc[k];
            if (P->pst_clamp_vel) {
                double cap = P->pst_m_delta * vn;
                double dn = sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
                if (dn > cap) for (int k = 0; k < 3; ++k) dv[k] *= cap / fmax(dn, 1e-30);
            }
            for (int k = 0; k < 3; ++k) dr[k] = dv[k] * dt;
        } else {
            for (int k = 0; k < 3; ++k) dr[k] = -P->cfl * Ma * h * h * t.acc[k];
        }
        if (P->pst_clamp_disp) {
            double mx = P->pst_max_shift * P->dx0;
            double dn = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
            if (dn > mx) for (int k = 0; k < 3; ++k) dr[k] *= mx / dn;
        }
        /* pst_constrain */
        double lam = s->lam[i];
        if (P->pst_surface_none) {
            if (lam < P->pst_lambda_hi) dr[0] = dr[1] = dr[2] = 0.0;
        } else if (lam < P->pst_lambda_hi) {
            if (lam < P->pst_lambda_lo) {
                dr[0] = dr[1] = dr[2] = 0.0;
            } else {
                double ns[3] = { s->lam_grad[3 * i], s->lam_grad[3 * i + 1], s->lam_grad[3 * i + 2] };
                if (d->flags & SPH_TWO_D) ns[2] = 0.0;
                double nn = sqrt(ns[0] * ns[0] + ns[1] * ns[1] + ns[2] * ns[2]);
                if (nn > 1e-10) {
                    double nh[3] = { ns[0] / nn, ns[1] / nn, ns[2] / nn };
                    double dd = dr[0] * nh[0] + dr[1] * nh[1] + dr[2] * nh[2];
                    for (int k = 0; k < 3; ++k) dr[k] -= dd * nh[k];
                }
            }
        }
        if (d->flags & SPH_TWO_D) dr[2] = 0.0;
        for (int k = 0; k < 3; ++k) s->shift[3 * i + k] = dr[k];
    }
    for (int64_t i = 0; i < n; ++i) or_displace(d, P, s, i, s->shift + 3 * i);
}

/* place_component / reflect_component (particle_system.py:1111-1133) */
static void or_place(const sph_fluid_params* P, or_state* s, int64_t i, int k, double val) {
    s->x[3 * i + k] = val;
    if (P->displacement_positions) s->u[3 * i + k] = val - s->x0[3 * i + k];
}
static void or_reflect(const sph_fluid_params* P, or_state* s, int64_t i, int k) {
    s->x[3 * i + k] = -s->x[3 * i + k];
    if (P->displacement_positions) s->u[3 * i + k] = -s->u[3 * i + k] - 2.0 * s->x0[3 * i + k];
}

/* simulate_collisions (sph_base.py:160-221) */
static double or_collide(const sph_fluid_params* P, or_state* s, int64_t i, const double* nrm) {
    double* v = s->v + 3 * i;
    double vn = v[0] * nrm[0] + v[1] * nrm[1] + v[2] * nrm[2];
    double w = 0.0;
    if (vn > 0.0) {
        w = 0.5 * s->m[i] * vn * vn * (1.0 - P->restitution * P->restitution);
        for (int k = 0; k < 3; ++k) v[k] -= (1.0 + P->restitution) * vn * nrm[k];
    }
    return w;
}

/* enforce_boundary_3D / _plane_strain / _axisymmetric (sph_base.py:223-397) */
static double or_walls(const sph_dims* d, const sph_fluid_params* P, or_state* s, int64_t i) {
    int nd = 3;
    if (d->flags & (SPH_TWO_D)) nd = 2;
    double pos[3] = { s->x[3 * i], s->x[3 * i + 1], s->x[3 * i + 2] };
    double nrm[3] = { 0, 0, 0 };
    for (int k = 0; k < nd; ++k) {
        double hi = P->domain_size[k] - P->padding;
        if (pos[k] > hi) { nrm[k] += 1.0; or_place(P, s, i, k, hi); }
        if ((d->flags & SPH_AXISYMMETRIC) && k == 1) continue;   /* the axis is no wall */
        if (pos[k] <= P->padding) { nrm[k] += -1.0; or_place(P, s, i, k, P->padding); }
    }
    double nl = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    if (nl > 1e-6) {
        double nh[3] = { nrm[0] / nl, nrm[1] / nl, nrm[2] / nl };
        return or_collide(P, s, i, nh);
    }
    return 0.0;
}

/* SPHBase.step's boundary sequence (sph_base.py:622-655) */
static void or_boundaries(const sph_dims* d, const sph_fluid_params* P, or_state* s) {
    double ww = 0.0;
    int axi = (d->flags & SPH_AXISYMMETRIC) != 0;
    int pls = (d->flags & SPH_PLANE_STRAIN) != 0;
    for (int64_t i = 0; i < s->n; ++i) {
        if (axi) {                                    /* enforce_axisymmetry_constraint */
            s->v[3 * i + 2] = 0.0; s->acc[3 * i + 2] = 0.0;
            or_place(P, s, i, 2, s->x0[3 * i + 2]);
            if (s->x[3 * i + 1] < 0.0) {
                or_reflect(P, s, i, 1);
                s->v[3 * i + 1] = -s->v[3 * i + 1];
                s->acc[3 * i + 1] = -s->acc[3 * i + 1];
            }
        } else {                                      /* enforce_symmetry_constraint */
            if (!(d->flags & SPH_TWO_D)) {
                if ((d->flags & SPH_MIRROR_Y) && s->x[3 * i + 1] < 0.0) {
                    or_reflect(P, s, i, 1);
                    s->v[3 * i + 1] = -s->v[3 * i + 1];
                    s->acc[3 * i + 1] = -s->acc[3 * i + 1];
                }
                if ((d->flags & SPH_MIRROR_Z) && s->x[3 * i + 2] < 0.0) {
                    or_reflect(P, s, i, 2);
                    s->v[3 * i + 2] = -s->v[3 * i + 2];
                    s->acc[3 * i + 2] = -s->acc[3 * i + 2];
                }
            }
        }
        if (P->walls) ww += or_walls(d, P, s, i);
        if (!axi && pls) {                            /* enforce_plane_strain_constraint */
            s->v[3 * i + 2] = 0.0; s->acc[3 * i + 2] = 0.0;
            or_place(P, s, i, 2, s->x0[3 * i + 2]);
        }
    }
    s->scal[SPH_S_WALL_WORK] += ww;
}

/* ------------------------------------------------------------------------------------ */
/* exported API                                                                         */
/* ------------------------------------------------------------------------------------ */
static or_grid g_grid;   /* the oracle is single-threaded at the API level */

void oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Sort the state and return the neighbour pairs of the sorted state. */
int64_t oracle_sort_and_pairs(const sph_dims* d, or_state* s, int32_t* out, int64_t cap) {
    oracle_sort(d, s, &g_grid);
    return oracle_neighbor_pairs(d, s, &g_grid, out, cap);
}

/* Pairs of the state as it is (grid from the last sort: binned cells, live positions). */
int64_t oracle_pairs_live(const sph_dims* d, const or_state* s, int32_t* out, int64_t cap) {
    return oracle_neighbor_pairs(d, s, &g_grid, out, cap);
}

/* One delta-plus-SPH step: SPHBase.step with DSPHSolver.substep (DSPH.py:1854-1883). */
void oracle_fluid_step(const sph_dims* d, const sph_fluid_params* P, or_state* s) {
    oracle_sort(d, s, &g_grid);
    or_compute_lam(d, P, s, &g_grid);
    or_volume(d, P, s, &g_grid);
    or_pressure_accelerations(d, P, s, &g_grid);
    or_adaptive_dt(d, P, s, &g_grid);
    or_step_work(P, s);
    or_advect(d, P, s);
    if (P->pst_enabled) or_apply_pst(d, P, s, &g_grid);
    or_boundaries(d, P, s);
}

/* f64 diagnostics (sph_base.py:425-620, DSPH.py:1670-1712) */
void oracle_reduce(const or_state* s, int which, const double* g, const sph_fluid_params* P,
                   double* out) {
    out[0] = out[1] = out[2] = 0.0;
    for (int64_t i = 0; i < s->n; ++i) {
        const double* v = s->v + 3 * i;
        double m = s->m[i];
        double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        switch (which) {
        case SPH_R_KINETIC: out[0] += 0.5 * m * v2; break;
        case SPH_R_KINETIC_OFFGRID: if (!or_isfinite3(s->x + 3 * i)) out[0] += 0.5 * m * v2; break;
        case SPH_R_POTENTIAL: out[0] -= m * (g[0] * s->x[3 * i] + g[1] * s->x[3 * i + 1] + g[2] * s->x[3 * i + 2]); break;
        case SPH_R_MASS: out[0] += m; break;
        case SPH_R_MOMENTUM: for (int k = 0; k < 3; ++k) out[k] += m * v[k]; break;
        case SPH_R_GROSS_MOMENTUM: out[0] += m * sqrt(v2); break;
        case SPH_R_TAIT_ENERGY: {
            double rho = m / s->V[i];
            double e = oracle_tait_energy(rho, P->rho0, P->stiffness, P->gamma);
            if (!P->allow_negative_pressure && rho < P->rho0) e = 0.0;
            out[0] += m * e;
        } break;
        default: break;
        }
    }
}
