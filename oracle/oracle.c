/*
 * FP64 oracle for the three workloads.  TEST INFRASTRUCTURE ONLY: this file is compiled into
 * oracle/liboracle.so and used by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline as the checker.  The product (symboliclib_amd) never links, loads or calls it.
 *
 * It evaluates every output value directly from the numeric inputs, the way the reference's
 * examples compute their comparison values, without any symbolic machinery:
 *   A^T A + A           examples/Tutorial/main.cpp:53-54 (Eigen: B = A.transpose() * A + A)
 *   cotangent Laplacian examples/BuildCotan/main.cpp:39-76 (double instantiation of buildCotan)
 *   Baraff-Witkin       examples/BaraffWitkinAD/stretch_shear_forces_impl.hpp:76-234 (the
 *                       hand-derived stretch + shear Hessian, triplets 2 * K9, line 214) and
 *                       the bending stiffness k_bend * Q(e) of
 *                       examples/BaraffWitkin/implementation/bend_quadratic_forces.cpp:54-155
 *                       with the quad-vertex order of examples/BaraffWitkin/adjacency.cpp:131-203.
 * Patterns follow Eigen's semantics: CSC, sorted unique rows per column, structural product
 * (no zero pruning), setFromTriplets summing duplicates.
 *
 * Every value comes with scale = sum of |term| over the terms the value is the sum of (the
 * products A(k,i) A(k,j) and A(i,j); the corner weights; the per-face Hessian entries and
 * bending constants), the denominator of the per-value relative error.
 *
 * Parity pins (tests/test_oracle.py): Tutorial known answers, exact rational A^T A + A on the
 * reference's data/sphere.mtx, scipy structural patterns, closed-form cotangent weights on a
 * regular grid, zero row sums, finite differences of the energy gradient, Hessian symmetry,
 * hand-computed bending stiffness Q on two quads, the contribution scale against finite
 * differences of single-face gradients, and the FP64 values against this same file compiled
 * in 80-bit extended precision (liboracle_ld.so) within 2e-15 of the term scale.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* Scale kinds of the cloth and cotangent values (the A^T A + A terms are the products
 * themselves, so both kinds agree there):
 *   SCALE_TERMS   every term inside each per-element contribution: the closed form evaluated
 *                 with the absolute value of every factor and additive term (the running
 *                 rounding-error scale of evaluating that closed form);
 *   SCALE_CONTRIB |contribution| summed over the per-element contributions of the output (the
 *                 per-face Hessian entries and bending constants, the corner weights). */
enum { SCALE_TERMS = 0, SCALE_CONTRIB = 1 };

/* ---------------------------------------------------------------- assembly --- */

/* Triplet list -> CSC pattern of a square matrix (rows and columns < ncols).  Writes out_cp
 * (ncols + 1) and, when out_ri != NULL, out_ri (nnz) and slot[t] = position of triplet t.
 * Returns nnz.  Two stable counting sorts (by row, then by column) put the triplets in
 * (column, row, triplet) order. */
static long assemble_pattern(long nt, const int* tr, const int* tc, int ncols, int* out_cp, int* out_ri, long* slot) {
    long* cnt = (long*)calloc((size_t)ncols + 1, sizeof(long));
    long* byrow = (long*)malloc(sizeof(long) * (size_t)(nt > 0 ? nt : 1));
    long* order = (long*)malloc(sizeof(long) * (size_t)(nt > 0 ? nt : 1));
    for (long t = 0; t < nt; ++t) cnt[tr[t] + 1]++;
    for (int j = 0; j < ncols; ++j) cnt[j + 1] += cnt[j];
    for (long t = 0; t < nt; ++t) byrow[cnt[tr[t]]++] = t;
    memset(cnt, 0, sizeof(long) * ((size_t)ncols + 1));
    for (long t = 0; t < nt; ++t) cnt[tc[t] + 1]++;
    for (int j = 0; j < ncols; ++j) cnt[j + 1] += cnt[j];
    long* fill = (long*)malloc(sizeof(long) * ((size_t)ncols + 1));
    memcpy(fill, cnt, sizeof(long) * ((size_t)ncols + 1));
    for (long a = 0; a < nt; ++a) {
        const long t = byrow[a];
        order[fill[tc[t]]++] = t;
    }
    long nnz = 0;
    for (int j = 0; j < ncols; ++j) {
        out_cp[j] = (int)nnz;
        int last = -1;
        for (long a = cnt[j]; a < cnt[j + 1]; ++a) {
            const long t = order[a];
            if (tr[t] != last) {
                if (out_ri) out_ri[nnz] = tr[t];
                last = tr[t];
                ++nnz;
            }
            if (slot) slot[t] = nnz - 1;
        }
    }
    out_cp[ncols] = (int)nnz;
    free(cnt);
    free(byrow);
    free(order);
    free(fill);
    return nnz;
}

/* Sum triplet values into their slots; scale += |value|.  Deterministic order (triplet order). */
static void scatter_sum(long nt, const long* slot, const double* tv, long nnz, double* vals, double* scale) {
    memset(vals, 0, sizeof(double) * (size_t)nnz);
    if (scale) memset(scale, 0, sizeof(double) * (size_t)nnz);
    for (long t = 0; t < nt; ++t) {
        vals[slot[t]] += tv[t];
        if (scale) scale[slot[t]] += fabs(tv[t]);
    }
}

/* ---------------------------------------------------------------- A^T A + A --- */

static int cmp_int(const void* a, const void* b) {
    const int x = *(const int*)a, y = *(const int*)b;
    return (x > y) - (x < y);
}

/* Rows i of column j of the structural A^T A + A: every i with A(k, i) and A(k, j) both stored
 * for some k, plus the rows of A's column j; unsorted, written to rows (mark: n ints, the
 * caller's, holding values != j on entry).  Returns their count. */
static int ata_column_rows(int j, const int* cp, const int* ri, const int* rp, const int* ci, int* mark, int* rows) {
    int m = 0;
    for (int p = cp[j]; p < cp[j + 1]; ++p) {
        const int k = ri[p];
        for (int q = rp[k]; q < rp[k + 1]; ++q) {
            const int i = ci[q];
            if (mark[i] != j) {
                mark[i] = j;
                if (rows) rows[m] = i;
                ++m;
            }
        }
    }
    for (int p = cp[j]; p < cp[j + 1]; ++p)
        if (mark[ri[p]] != j) {
            mark[ri[p]] = j;
            if (rows) rows[m] = ri[p];
            ++m;
        }
    return m;
}

long oracle_ata_plus_a_pattern(int n, const int* cp, const int* ri, int* out_cp, int* out_ri) {
    /* CSR of A = rows' column lists */
    int* rp = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int p = 0; p < cp[n]; ++p) rp[ri[p] + 1]++;
    for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
    int* ci = (int*)malloc(sizeof(int) * (size_t)(cp[n] > 0 ? cp[n] : 1));
    int* fill = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    memcpy(fill, rp, sizeof(int) * ((size_t)n + 1));
    for (int j = 0; j < n; ++j)
        for (int p = cp[j]; p < cp[j + 1]; ++p) ci[fill[ri[p]]++] = j;

    long* count = (long*)calloc((size_t)n + 1, sizeof(long));
#pragma omp parallel
    {
        int* mark = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
        for (int i = 0; i < n; ++i) mark[i] = -1;
#pragma omp for schedule(dynamic, 256)
        for (int j = 0; j < n; ++j) count[j + 1] = ata_column_rows(j, cp, ri, rp, ci, mark, NULL);
        free(mark);
    }
    for (int j = 0; j < n; ++j) count[j + 1] += count[j];
    for (int j = 0; j <= n; ++j) out_cp[j] = (int)count[j];
    if (out_ri) {
#pragma omp parallel
        {
            int* mark = (int*)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
            for (int i = 0; i < n; ++i) mark[i] = -1;
#pragma omp for schedule(dynamic, 256)
            for (int j = 0; j < n; ++j) {
                int* rows = out_ri + count[j];
                const int m = ata_column_rows(j, cp, ri, rp, ci, mark, rows);
                qsort(rows, (size_t)m, sizeof(int), cmp_int);
            }
            free(mark);
        }
    }
    const long nnz = count[n];
    free(rp);
    free(ci);
    free(fill);
    free(count);
    return nnz;
}

/* values and scales on the pattern from oracle_ata_plus_a_pattern */
void oracle_ata_plus_a_values(int n, const int* cp, const int* ri, const double* a, const int* bcp, const int* bri,
                              double* bv, double* scale) {
    int* rp = (int*)calloc((size_t)n + 1, sizeof(int));
    for (int p = 0; p < cp[n]; ++p) rp[ri[p] + 1]++;
    for (int i = 0; i < n; ++i) rp[i + 1] += rp[i];
    int* ci = (int*)malloc(sizeof(int) * (size_t)(cp[n] > 0 ? cp[n] : 1));
    double* cv = (double*)malloc(sizeof(double) * (size_t)(cp[n] > 0 ? cp[n] : 1));
    int* fill = (int*)malloc(sizeof(int) * ((size_t)n + 1));
    memcpy(fill, rp, sizeof(int) * ((size_t)n + 1));
    for (int j = 0; j < n; ++j)
        for (int p = cp[j]; p < cp[j + 1]; ++p) {
            ci[fill[ri[p]]] = j;
            cv[fill[ri[p]]++] = a[p];
        }

#pragma omp parallel
    {
        double* acc = (double*)calloc((size_t)n, sizeof(double));
        double* sacc = (double*)calloc((size_t)n, sizeof(double));
#pragma omp for schedule(dynamic, 64)
        for (int j = 0; j < n; ++j) {
            for (int p = cp[j]; p < cp[j + 1]; ++p) {
                const int k = ri[p];
                const double akj = a[p];
                for (int q = rp[k]; q < rp[k + 1]; ++q) {
                    const double t = cv[q] * akj; /* A(k,i) A(k,j) */
                    acc[ci[q]] += t;
                    sacc[ci[q]] += fabs(t);
                }
            }
            for (int p = cp[j]; p < cp[j + 1]; ++p) {
                acc[ri[p]] += a[p];
                sacc[ri[p]] += fabs(a[p]);
            }
            for (int p = bcp[j]; p < bcp[j + 1]; ++p) {
                bv[p] = acc[bri[p]];
                if (scale) scale[p] = sacc[bri[p]];
                acc[bri[p]] = 0.0;
                sacc[bri[p]] = 0.0;
            }
        }
        free(acc);
        free(sacc);
    }
    free(rp);
    free(ci);
    free(cv);
    free(fill);
}

/* ---------------------------------------------------------------- geometry --- */

typedef struct {
    double x, y, z;
} v3;

static inline v3 vget(const double* X, int n, int v) {
    v3 r = {X[v], X[v + n], X[v + 2 * (long)n]};
    return r;
}
static inline v3 vsub(v3 a, v3 b) {
    v3 r = {a.x - b.x, a.y - b.y, a.z - b.z};
    return r;
}
static inline double vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b) {
    v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static inline double vnorm(v3 a) { return sqrt(vdot(a, a)); }
static inline v3 vscale(v3 a, double s) {
    v3 r = {a.x * s, a.y * s, a.z * s};
    return r;
}

/* ---------------------------------------------------------------- cotangent --- */

static void cotan_triplets_rc(int nf, const int* F, int* tr, int* tc) {
    for (long f = 0; f < nf; ++f)
        for (int j = 0; j < 3; ++j) {
            const int j1 = F[3 * f + (j + 1) % 3], j2 = F[3 * f + (j + 2) % 3];
            const long t = 12 * f + 4 * j;
            tr[t] = j1, tc[t] = j1;
            tr[t + 1] = j2, tc[t + 1] = j2;
            tr[t + 2] = j1, tc[t + 2] = j2;
            tr[t + 3] = j2, tc[t + 3] = j1;
        }
}

long oracle_cotan_pattern(int nv, int nf, const int* F, int* out_cp, int* out_ri) {
    const long nt = 12L * nf;
    int* tr = (int*)malloc(sizeof(int) * (size_t)nt);
    int* tc = (int*)malloc(sizeof(int) * (size_t)nt);
    cotan_triplets_rc(nf, F, tr, tc);
    long nnz = assemble_pattern(nt, tr, tc, nv, out_cp, out_ri, NULL);
    free(tr);
    free(tc);
    return nnz;
}

/* corner weights w = (e0 . e1) / |e0 x e1| (BuildCotan/main.cpp:49-70); tm (optional) gets the
 * magnitude of every weight: SCALE_TERMS |e0_k e1_k| summed over k / |e0 x e1|, SCALE_CONTRIB |w| */
static void cotan_triplet_values(int nv, int nf, const int* F, const double* V, double* tv, double* tm, int kind) {
#pragma omp parallel for schedule(static)
    for (long f = 0; f < nf; ++f) {
        const int* fv = F + 3 * f;
        const v3 p0 = vget(V, nv, fv[0]);
        const double a = vnorm(vcross(vsub(p0, vget(V, nv, fv[1])), vsub(p0, vget(V, nv, fv[2]))));
        for (int j = 0; j < 3; ++j) {
            const v3 pj = vget(V, nv, fv[j]);
            const v3 d0 = vsub(pj, vget(V, nv, fv[(j + 1) % 3]));
            const v3 d1 = vsub(pj, vget(V, nv, fv[(j + 2) % 3]));
            const double w = vdot(d0, d1) / a;
            const long t = 12 * f + 4 * j;
            tv[t] = -w;
            tv[t + 1] = -w;
            tv[t + 2] = w;
            tv[t + 3] = w;
            if (tm) {
                const double m = kind == SCALE_CONTRIB ? fabs(w)
                                                       : (fabs(d0.x * d1.x) + fabs(d0.y * d1.y) + fabs(d0.z * d1.z)) / a;
                tm[t] = tm[t + 1] = tm[t + 2] = tm[t + 3] = m;
            }
        }
    }
}

/* Prepared assembly: triplet slots computed once, values summed per call. */
typedef struct {
    long nt, nnz;
    long* slot;
    long* segStart; /* nnz + 1: triplets of slot s are perm[segStart[s] .. segStart[s+1]) */
    long* perm;
    double* tv;
    double* tm; /* optional magnitude of each triplet value (NULL: |tv|) */
} assembly;

static void assembly_init(assembly* A, long nt, const int* tr, const int* tc, int ncols, int* cp, int** ri) {
    A->nt = nt;
    A->slot = (long*)malloc(sizeof(long) * (size_t)(nt > 0 ? nt : 1));
    int* rows = NULL;
    if (ri) *ri = rows = (int*)malloc(sizeof(int) * (size_t)(nt > 0 ? nt : 1)); /* nnz <= nt */
    A->nnz = assemble_pattern(nt, tr, tc, ncols, cp, rows, A->slot);
    A->segStart = (long*)calloc((size_t)A->nnz + 1, sizeof(long));
    A->perm = (long*)malloc(sizeof(long) * (size_t)(nt > 0 ? nt : 1));
    A->tv = (double*)malloc(sizeof(double) * (size_t)(nt > 0 ? nt : 1));
    A->tm = NULL;
    for (long t = 0; t < nt; ++t) A->segStart[A->slot[t] + 1]++;
    for (long s = 0; s < A->nnz; ++s) A->segStart[s + 1] += A->segStart[s];
    long* fill = (long*)malloc(sizeof(long) * ((size_t)A->nnz + 1));
    memcpy(fill, A->segStart, sizeof(long) * ((size_t)A->nnz + 1));
    for (long t = 0; t < nt; ++t) A->perm[fill[A->slot[t]]++] = t;
    free(fill);
}

static void assembly_sum(const assembly* A, double* vals, double* scale) {
#pragma omp parallel for schedule(static)
    for (long s = 0; s < A->nnz; ++s) {
        double v = 0.0, sc = 0.0;
        for (long q = A->segStart[s]; q < A->segStart[s + 1]; ++q) {
            const long k = A->perm[q];
            const double t = A->tv[k];
            v += t;
            sc += A->tm ? A->tm[k] : fabs(t);
        }
        vals[s] = v;
        if (scale) scale[s] = sc;
    }
}

static void assembly_free(assembly* A) {
    free(A->slot);
    free(A->segStart);
    free(A->perm);
    free(A->tv);
    free(A->tm);
}

void oracle_cotan_values(int nv, int nf, const int* F, const double* V, double* vals, double* scale, int kind) {
    const long nt = 12L * nf;
    int* tr = (int*)malloc(sizeof(int) * (size_t)nt);
    int* tc = (int*)malloc(sizeof(int) * (size_t)nt);
    int* cp = (int*)malloc(sizeof(int) * ((size_t)nv + 1));
    cotan_triplets_rc(nf, F, tr, tc);
    assembly A;
    assembly_init(&A, nt, tr, tc, nv, cp, NULL);
    if (scale) A.tm = (double*)malloc(sizeof(double) * (size_t)nt);
    cotan_triplet_values(nv, nf, F, V, A.tv, A.tm, kind);
    assembly_sum(&A, vals, scale);
    assembly_free(&A);
    free(tr);
    free(tc);
    free(cp);
}

/* ---------------------------------------------------------------- Baraff-Witkin --- */

/* rest-shape area and inverse frame (stretch_shear_forces_impl.hpp:23-49): Vinv row-major */
static void bw_rest(int nv, const int* fv, const double* R, double* area, double Vi[4]) {
    const v3 p0 = vget(R, nv, fv[0]);
    const v3 v1 = vsub(vget(R, nv, fv[1]), p0), v2 = vsub(vget(R, nv, fv[2]), p0);
    *area = 0.5 * vnorm(vcross(v1, v2));
    const double l1 = vnorm(v1);
    const v3 u = vscale(v1, 1.0 / l1);
    v3 w = vcross(vcross(u, v2), u);
    w = vscale(w, 1.0 / vnorm(w));
    const double A = l1, B = vdot(v2, u), C = 0.0, D = vdot(v2, w);
    const double inv = 1.0 / (A * D - B * C);
    Vi[0] = D * inv;
    Vi[1] = -B * inv;
    Vi[2] = -C * inv;
    Vi[3] = A * inv;
}

/* Hand-derived 9x9 stretch+shear stiffness K9 of one face (stretch_shear_forces_impl.hpp:99-203),
 * row-major, index 3*vertex + coordinate. */
/* K9 and, when M9 != NULL, the magnitude of every entry: the same closed form evaluated with
 * the absolute value of every factor and every additive term (from the edge vectors
 * D0 = x1 - x0, D1 = x2 - x0 on), i.e. the sum of the absolute magnitudes of the terms the
 * entry is made of.  Entries such as (1 - Jn_x^2) / |J| cancel internally on flat parts of the
 * cloth; their magnitude keeps both terms. */
static void bw_face_K9(int nv, const int* fv, const double* R, const double* X, double ks, double ksh, double K9[81],
                       double* M9) {
    double a, Vi[4];
    bw_rest(nv, fv, R, &a, Vi);
    double B[3][2];
    B[0][0] = -Vi[0] - Vi[2];
    B[0][1] = -Vi[1] - Vi[3];
    B[1][0] = Vi[0];
    B[1][1] = Vi[1];
    B[2][0] = Vi[2];
    B[2][1] = Vi[3];
    /* magnitudes of B: the first row is a sum of two inverse-frame entries, which cancel to
     * exactly zero on some rest triangles (a right isosceles triangle whose first edge is a
     * leg); its rounding then scales with |Vi|, not with |B| */
    const double Bm[3][2] = {{fabs(Vi[0]) + fabs(Vi[2]), fabs(Vi[1]) + fabs(Vi[3])},
                             {fabs(Vi[0]), fabs(Vi[1])},
                             {fabs(Vi[2]), fabs(Vi[3])}};

    const v3 x0 = vget(X, nv, fv[0]);
    const v3 D0 = vsub(vget(X, nv, fv[1]), x0), D1 = vsub(vget(X, nv, fv[2]), x0);
    /* J = D * Vinv: column c = D0 Vi(0,c) + D1 Vi(1,c) */
    const v3 J0 = {D0.x * Vi[0] + D1.x * Vi[2], D0.y * Vi[0] + D1.y * Vi[2], D0.z * Vi[0] + D1.z * Vi[2]};
    const v3 J1 = {D0.x * Vi[1] + D1.x * Vi[3], D0.y * Vi[1] + D1.y * Vi[3], D0.z * Vi[1] + D1.z * Vi[3]};
    const double n0 = vnorm(J0), n1 = vnorm(J1);
    const double Jn0[3] = {J0.x / n0, J0.y / n0, J0.z / n0};
    const double Jn1[3] = {J1.x / n1, J1.y / n1, J1.z / n1};
    const double J0a[3] = {J0.x, J0.y, J0.z}, J1a[3] = {J1.x, J1.y, J1.z};
    /* magnitudes: |D| Vinv with absolute values, norms of those, normalised by the actual norm */
    const double D0a[3] = {D0.x, D0.y, D0.z}, D1a[3] = {D1.x, D1.y, D1.z};
    double J0m[3], J1m[3], Jn0m[3], Jn1m[3];
    for (int r = 0; r < 3; ++r) {
        J0m[r] = fabs(D0a[r] * Vi[0]) + fabs(D1a[r] * Vi[2]);
        J1m[r] = fabs(D0a[r] * Vi[1]) + fabs(D1a[r] * Vi[3]);
        Jn0m[r] = J0m[r] / n0;
        Jn1m[r] = J1m[r] / n1;
    }
    const double n0m = sqrt(J0m[0] * J0m[0] + J0m[1] * J0m[1] + J0m[2] * J0m[2]);
    const double n1m = sqrt(J1m[0] * J1m[0] + J1m[1] * J1m[1] + J1m[2] * J1m[2]);

    /* stretch */
    const double C0 = a * (n0 - 1.0), C1 = a * (n1 - 1.0);
    double dC[3][3][2]; /* dC_dv[i] (3 x 2) */
    for (int i = 0; i < 3; ++i)
        for (int r = 0; r < 3; ++r) {
            dC[i][r][0] = a * B[i][0] * Jn0[r];
            dC[i][r][1] = a * B[i][1] * Jn1[r];
        }
    double IJx[3][3], IJy[3][3];
    for (int r = 0; r < 3; ++r)
        for (int s = 0; s < 3; ++s) {
            const double id = r == s ? 1.0 : 0.0;
            IJx[r][s] = a * (id - Jn0[r] * Jn0[s]) / n0;
            IJy[r][s] = a * (id - Jn1[r] * Jn1[s]) / n1;
        }
    /* shear */
    const double Cs = a * vdot(J0, J1);
    double Cdv[3][3]; /* column i */
    for (int i = 0; i < 3; ++i)
        for (int r = 0; r < 3; ++r) Cdv[r][i] = a * (B[i][0] * J1a[r] + B[i][1] * J0a[r]);

    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            const double ddCs = a * (B[i][0] * B[j][1] + B[i][1] * B[j][0]);
            for (int r = 0; r < 3; ++r)
                for (int s = 0; s < 3; ++s) {
                    const double st = dC[i][r][0] * dC[j][s][0] + dC[i][r][1] * dC[j][s][1] +
                                      B[i][0] * B[j][0] * IJx[r][s] * C0 + B[i][1] * B[j][1] * IJy[r][s] * C1;
                    double sh = Cdv[r][i] * Cdv[s][j];
                    if (r == s) sh += ddCs * Cs;
                    K9[(3 * i + r) * 9 + 3 * j + s] = -ks * st - ksh * sh;
                    if (M9) {
                        const double id = r == s ? 1.0 : 0.0;
                        const double IJxm = a * (id + Jn0m[r] * Jn0m[s]) / n0;
                        const double IJym = a * (id + Jn1m[r] * Jn1m[s]) / n1;
                        const double C0m = a * (n0m + 1.0), C1m = a * (n1m + 1.0);
                        const double stm = a * a * Bm[i][0] * Bm[j][0] * Jn0m[r] * Jn0m[s] +
                                           a * a * Bm[i][1] * Bm[j][1] * Jn1m[r] * Jn1m[s] +
                                           Bm[i][0] * Bm[j][0] * IJxm * C0m + Bm[i][1] * Bm[j][1] * IJym * C1m;
                        const double Cdvm_ri = a * (Bm[i][0] * J1m[r] + Bm[i][1] * J0m[r]);
                        const double Cdvm_sj = a * (Bm[j][0] * J1m[s] + Bm[j][1] * J0m[s]);
                        double shm = Cdvm_ri * Cdvm_sj;
                        if (r == s) {
                            const double Csm = a * (J0m[0] * J1m[0] + J0m[1] * J1m[1] + J0m[2] * J1m[2]);
                            shm += a * (Bm[i][0] * Bm[j][1] + Bm[i][1] * Bm[j][0]) * Csm;
                        }
                        M9[(3 * i + r) * 9 + 3 * j + s] = fabs(ks) * stm + fabs(ksh) * shm;
                    }
                }
        }
}

/* energy of one face (stretch_shear_forces_impl.hpp:51-73) */
static double bw_face_energy(int nv, const int* fv, const double* R, const double* X, double ks, double ksh) {
    double a, Vi[4];
    bw_rest(nv, fv, R, &a, Vi);
    const v3 x0 = vget(X, nv, fv[0]);
    const v3 D0 = vsub(vget(X, nv, fv[1]), x0), D1 = vsub(vget(X, nv, fv[2]), x0);
    const v3 J0 = {D0.x * Vi[0] + D1.x * Vi[2], D0.y * Vi[0] + D1.y * Vi[2], D0.z * Vi[0] + D1.z * Vi[2]};
    const v3 J1 = {D0.x * Vi[1] + D1.x * Vi[3], D0.y * Vi[1] + D1.y * Vi[3], D0.z * Vi[1] + D1.z * Vi[3]};
    const double c0 = a * (vnorm(J0) - 1.0), c1 = a * (vnorm(J1) - 1.0), cs = a * vdot(J0, J1);
    return -ks * (c0 * c0 + c1 * c1) - ksh * cs * cs;
}

double oracle_bw_energy(int nv, int nf, const int* F, const double* R, const double* X, double ks, double ksh) {
    double e = 0.0;
    for (long f = 0; f < nf; ++f) e += bw_face_energy(nv, F + 3 * f, R, X, ks, ksh);
    return e;
}

/* gradient of the stretch+shear energy: -2 (forces) of stretch_shear_forces_impl.hpp:111-172 */
void oracle_bw_gradient(int nv, int nf, const int* F, const double* R, const double* X, double ks, double ksh, double* g) {
    memset(g, 0, sizeof(double) * 3 * (size_t)nv);
    for (long f = 0; f < nf; ++f) {
        const int* fv = F + 3 * f;
        double a, Vi[4];
        bw_rest(nv, fv, R, &a, Vi);
        double B[3][2] = {{-Vi[0] - Vi[2], -Vi[1] - Vi[3]}, {Vi[0], Vi[1]}, {Vi[2], Vi[3]}};
        const v3 x0 = vget(X, nv, fv[0]);
        const v3 D0 = vsub(vget(X, nv, fv[1]), x0), D1 = vsub(vget(X, nv, fv[2]), x0);
        const v3 J0 = {D0.x * Vi[0] + D1.x * Vi[2], D0.y * Vi[0] + D1.y * Vi[2], D0.z * Vi[0] + D1.z * Vi[2]};
        const v3 J1 = {D0.x * Vi[1] + D1.x * Vi[3], D0.y * Vi[1] + D1.y * Vi[3], D0.z * Vi[1] + D1.z * Vi[3]};
        const double n0 = vnorm(J0), n1 = vnorm(J1);
        const double C0 = a * (n0 - 1.0), C1 = a * (n1 - 1.0), Cs = a * vdot(J0, J1);
        const double J0a[3] = {J0.x, J0.y, J0.z}, J1a[3] = {J1.x, J1.y, J1.z};
        for (int i = 0; i < 3; ++i)
            for (int r = 0; r < 3; ++r) {
                const double dC0 = a * B[i][0] * J0a[r] / n0, dC1 = a * B[i][1] * J1a[r] / n1;
                const double dCs = a * (B[i][0] * J1a[r] + B[i][1] * J0a[r]);
                /* dE = -2 ks (C0 dC0 + C1 dC1) - 2 ksh Cs dCs */
                g[fv[i] + (long)r * nv] += -2.0 * ks * (C0 * dC0 + C1 * dC1) - 2.0 * ksh * Cs * dCs;
            }
    }
}

/* interior edges: v0, v1 on the edge (oriented as in the first face), v2 opposite in the first
 * face, v3 opposite in the second; edges in increasing (min, max) order.  Returns count. */
static int cmp_edge(const void* a, const void* b) {
    const long* x = (const long*)a;
    const long* y = (const long*)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    if (x[1] != y[1]) return x[1] < y[1] ? -1 : 1;
    return x[2] < y[2] ? -1 : (x[2] > y[2]);
}

long oracle_interior_edges(int nf, const int* F, int* E4) {
    long* he = (long*)malloc(sizeof(long) * 3 * 3 * (size_t)nf); /* (min, max, face*3+k) */
    for (long f = 0; f < nf; ++f)
        for (int k = 0; k < 3; ++k) {
            const int a = F[3 * f + k], b = F[3 * f + (k + 1) % 3];
            long* h = he + 3 * (3 * f + k);
            h[0] = a < b ? a : b;
            h[1] = a < b ? b : a;
            h[2] = 3 * f + k;
        }
    qsort(he, 3 * (size_t)nf, 3 * sizeof(long), cmp_edge);
    long ne = 0;
    for (long s = 0; s < 3L * nf;) {
        long e = s + 1;
        while (e < 3L * nf && he[3 * e] == he[3 * s] && he[3 * e + 1] == he[3 * s + 1]) ++e;
        if (e - s == 2) {
            const long fa = he[3 * s + 2] / 3, ka = he[3 * s + 2] % 3, fb = he[3 * (s + 1) + 2] / 3;
            const int v0 = F[3 * fa + ka], v1 = F[3 * fa + (ka + 1) % 3], v2 = F[3 * fa + (ka + 2) % 3];
            int v3i = -1;
            for (int k = 0; k < 3; ++k)
                if (F[3 * fb + k] != v0 && F[3 * fb + k] != v1) v3i = F[3 * fb + k];
            if (E4) {
                E4[4 * ne] = v0;
                E4[4 * ne + 1] = v1;
                E4[4 * ne + 2] = v2;
                E4[4 * ne + 3] = v3i;
            }
            ++ne;
        }
        s = e;
    }
    free(he);
    return ne;
}

static double cot_theta(v3 v, v3 w) { return vdot(v, w) / vnorm(vcross(v, w)); }

/* Q(e) = -3 / (A0 + A1) K0 K0^T (bend_quadratic_forces.cpp:54-83), row-major 4x4 */
void oracle_bend_Q(int nv, const int* q, const double* R, double Q[16]) {
    const v3 x0 = vget(R, nv, q[0]), x1 = vget(R, nv, q[1]), x2 = vget(R, nv, q[2]), x3 = vget(R, nv, q[3]);
    const v3 e0 = vsub(x1, x0), e1 = vsub(x2, x0), e2 = vsub(x3, x0), e3 = vsub(x2, x1), e4 = vsub(x3, x1);
    const v3 me0 = vscale(e0, -1.0);
    const double c01 = cot_theta(e0, e1), c02 = cot_theta(e0, e2), c03 = cot_theta(me0, e3), c04 = cot_theta(me0, e4);
    const double K0[4] = {c03 + c04, c01 + c02, -c01 - c03, -c02 - c04};
    const double coef = -3.0 / (vnorm(vcross(e0, e1)) + vnorm(vcross(e0, e2)));
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) Q[4 * i + j] = coef * K0[i] * K0[j];
}

/* triplet (row, col) lists: 81 per face (2 K9, full symmetric), then 3 * 16 per edge (bend) */
static long bw_triplets_rc(int nv, int nf, const int* F, long ne, const int* E4, int* tr, int* tc) {
    long t = 0;
    for (long f = 0; f < nf; ++f)
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q) {
                        tr[t] = F[3 * f + i] + p * nv;
                        tc[t] = F[3 * f + j] + q * nv;
                        ++t;
                    }
    for (long e = 0; e < ne; ++e)
        for (int dim = 0; dim < 3; ++dim)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b) {
                    tr[t] = E4[4 * e + a] + dim * nv;
                    tc[t] = E4[4 * e + b] + dim * nv;
                    ++t;
                }
    return t;
}

long oracle_bw_pattern(int nv, int nf, const int* F, int with_bend, int* out_cp, int* out_ri) {
    const long ne = with_bend ? oracle_interior_edges(nf, F, NULL) : 0;
    int* E4 = (int*)malloc(sizeof(int) * 4 * (size_t)(ne > 0 ? ne : 1));
    if (ne) oracle_interior_edges(nf, F, E4);
    const long nt = 81L * nf + 48L * ne;
    int* tr = (int*)malloc(sizeof(int) * (size_t)nt);
    int* tc = (int*)malloc(sizeof(int) * (size_t)nt);
    bw_triplets_rc(nv, nf, F, ne, E4, tr, tc);
    long nnz = assemble_pattern(nt, tr, tc, 3 * nv, out_cp, out_ri, NULL);
    free(tr);
    free(tc);
    free(E4);
    return nnz;
}

/* Prepared Baraff-Witkin assembly, reusable across value sets (for timing). */
typedef struct {
    int nv, nf;
    long ne;
    const int* F;
    int* E4;
    double* Q; /* 16 per edge, scaled by k_bend */
    const double* R;
    double ks, ksh;
    assembly A;
    int* cp;
    int* ri; /* rows of the pattern (nnz) */
} bw_plan;

void* oracle_bw_plan_create(int nv, int nf, const int* F, const double* R, double ks, double ksh, double kb) {
    bw_plan* P = (bw_plan*)calloc(1, sizeof(bw_plan));
    P->nv = nv;
    P->nf = nf;
    P->F = F;
    P->R = R;
    P->ks = ks;
    P->ksh = ksh;
    P->ne = kb != 0.0 ? oracle_interior_edges(nf, F, NULL) : 0;
    P->E4 = (int*)malloc(sizeof(int) * 4 * (size_t)(P->ne > 0 ? P->ne : 1));
    if (P->ne) oracle_interior_edges(nf, F, P->E4);
    P->Q = (double*)malloc(sizeof(double) * 16 * (size_t)(P->ne > 0 ? P->ne : 1));
    for (long e = 0; e < P->ne; ++e) {
        oracle_bend_Q(nv, P->E4 + 4 * e, R, P->Q + 16 * e);
        for (int k = 0; k < 16; ++k) P->Q[16 * e + k] *= kb;
    }
    const long nt = 81L * nf + 48L * P->ne;
    int* tr = (int*)malloc(sizeof(int) * (size_t)nt);
    int* tc = (int*)malloc(sizeof(int) * (size_t)nt);
    bw_triplets_rc(nv, nf, F, P->ne, P->E4, tr, tc);
    P->cp = (int*)malloc(sizeof(int) * (3 * (size_t)nv + 1));
    assembly_init(&P->A, nt, tr, tc, 3 * nv, P->cp, &P->ri);
    free(tr);
    free(tc);
    /* bending values are constant */
    P->A.tm = (double*)malloc(sizeof(double) * (size_t)(nt > 0 ? nt : 1));
    for (long e = 0; e < P->ne; ++e)
        for (int dim = 0; dim < 3; ++dim)
            for (int k = 0; k < 16; ++k) {
                const long t = 81L * nf + 48 * e + 16 * dim + k;
                P->A.tv[t] = P->Q[16 * e + k];
                P->A.tm[t] = fabs(P->Q[16 * e + k]);
            }
    return P;
}

long oracle_bw_plan_nnz(void* plan) { return ((bw_plan*)plan)->A.nnz; }

/* The plan's CSC pattern (cp: 3 nv + 1, ri: nnz) and, per output, whether a bending constant
 * is among its contributions (bend: nnz bytes). */
void oracle_bw_plan_pattern(void* plan, int* cp, int* ri, unsigned char* bend) {
    const bw_plan* P = (const bw_plan*)plan;
    memcpy(cp, P->cp, sizeof(int) * (3 * (size_t)P->nv + 1));
    memcpy(ri, P->ri, sizeof(int) * (size_t)P->A.nnz);
    const long faceTriplets = 81L * P->nf;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < P->A.nnz; ++s) {
        unsigned char b = 0;
        for (long q = P->A.segStart[s]; q < P->A.segStart[s + 1]; ++q) b |= P->A.perm[q] >= faceTriplets;
        bend[s] = b;
    }
}

void oracle_bw_plan_values(void* plan, const double* X, double* vals, double* scale, int kind) {
    bw_plan* P = (bw_plan*)plan;
#pragma omp parallel for schedule(static)
    for (long f = 0; f < P->nf; ++f) {
        double K9[81], M9[81];
        bw_face_K9(P->nv, P->F + 3 * f, P->R, X, P->ks, P->ksh, K9, scale && kind == SCALE_TERMS ? M9 : NULL);
        double* tv = P->A.tv + 81 * f;
        double* tm = P->A.tm + 81 * f;
        /* triplet (i, j, p, q) -> 2 K9(3i + p, 3j + q) */
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int p = 0; p < 3; ++p)
                    for (int q = 0; q < 3; ++q) {
                        tv[27 * i + 9 * j + 3 * p + q] = 2.0 * K9[(3 * i + p) * 9 + 3 * j + q];
                        if (scale)
                            tm[27 * i + 9 * j + 3 * p + q] = kind == SCALE_TERMS ? 2.0 * M9[(3 * i + p) * 9 + 3 * j + q]
                                                                                  : fabs(tv[27 * i + 9 * j + 3 * p + q]);
                    }
    }
    assembly_sum(&P->A, vals, scale);
}

/* Values with both scales in one pass: st (SCALE_TERMS) and sk (SCALE_CONTRIB). */
void oracle_bw_plan_check(void* plan, const double* X, double* vals, double* st, double* sk) {
    bw_plan* P = (bw_plan*)plan;
    oracle_bw_plan_values(plan, X, vals, st, SCALE_TERMS);
    const assembly* A = &P->A;
#pragma omp parallel for schedule(static)
    for (long s = 0; s < A->nnz; ++s) {
        double m = 0.0;
        for (long q = A->segStart[s]; q < A->segStart[s + 1]; ++q) m += fabs(A->tv[A->perm[q]]);
        sk[s] = m;
    }
}

void oracle_bw_plan_free(void* plan) {
    bw_plan* P = (bw_plan*)plan;
    if (!P) return;
    assembly_free(&P->A);
    free(P->E4);
    free(P->Q);
    free(P->cp);
    free(P->ri);
    free(P);
}

/* cotangent plan, reusable across value sets (for timing) */
typedef struct {
    int nv, nf;
    const int* F;
    assembly A;
    int* cp;
} cotan_plan;

void* oracle_cotan_plan_create(int nv, int nf, const int* F) {
    cotan_plan* P = (cotan_plan*)calloc(1, sizeof(cotan_plan));
    P->nv = nv;
    P->nf = nf;
    P->F = F;
    const long nt = 12L * nf;
    int* tr = (int*)malloc(sizeof(int) * (size_t)nt);
    int* tc = (int*)malloc(sizeof(int) * (size_t)nt);
    cotan_triplets_rc(nf, F, tr, tc);
    P->cp = (int*)malloc(sizeof(int) * ((size_t)nv + 1));
    assembly_init(&P->A, nt, tr, tc, nv, P->cp, NULL);
    free(tr);
    free(tc);
    return P;
}

void oracle_cotan_plan_values(void* plan, const double* V, double* vals, double* scale, int kind) {
    cotan_plan* P = (cotan_plan*)plan;
    if (scale && !P->A.tm) P->A.tm = (double*)malloc(sizeof(double) * (size_t)P->A.nt);
    cotan_triplet_values(P->nv, P->nf, P->F, V, P->A.tv, scale ? P->A.tm : NULL, kind);
    assembly_sum(&P->A, vals, scale);
}

void oracle_cotan_plan_free(void* plan) {
    cotan_plan* P = (cotan_plan*)plan;
    if (!P) return;
    assembly_free(&P->A);
    free(P->cp);
    free(P);
}

int oracle_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
