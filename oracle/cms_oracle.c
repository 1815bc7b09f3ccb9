/*
 * cms_oracle.c — CPU oracle of the Meta-Storms all-pairs path.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: tests, the smoke check
 * and the CPU-baseline leg of the benchmark call it; the product library never
 * links, loads or calls it.
 *
 * Restates, in plain C:
 *   orc_calc_sim_f32   _Comp_Tree::Calc_sim (comp.h:332-375), FP32, same op
 *                      order; built with -ffp-contract=off so every operation
 *                      rounds to float as in the x86-64 reference build.
 *   orc_calc_sim_f64   the same traversal in FP64 (the FP64 oracle).
 *   orc_sharedmin_f64  an independent FP64 closed form of the same value:
 *                      100*sim = sum over child edges e of beta_e *
 *                      min(x_a(e), x_b(e)), where x_s(e) is sample s's mass
 *                      under the child, decayed by w along each edge below it,
 *                      and beta_e = 1 - w_e if the parent's value is consumed,
 *                      1 if the parent is a root. Valid for tree schedules.
 *   orc_load_abd       _Comp_Tree::Load_abd over a table row (comp.h:243-279)
 *                      with the column->leaf map and copy numbers resolved.
 *   orc_all_pairs_f32  Multi_Comp_Table's pair loop (comp_sam.cpp:321-349),
 *                      OpenMP dynamic,1, packed upper triangle.
 *   orc_leaf_sim_f32   comp -M 2..5 (cosine, Euclidean, JSD, Bray-Curtis;
 *                      dist.h:13-159 via comp.h:296-330), FP32 as the
 *                      reference; orc_leaf_sim_f64 the same from their
 *                      definitions in FP64.
 *
 * Pinning: orc_calc_sim_f32 is checked bit-exactly against the compiled
 * reference (golden fixtures); orc_calc_sim_f64 and orc_sharedmin_f64 are two
 * different algorithms that must agree to 1e-12; the 4-leaf tree of the
 * known-answer test fixes all three by hand values.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_REG_SIZE 150
#define ORC_MIN_DIST 0.00001

/* comp.h:332-375 */
float orc_calc_sim_f32(int order_n, const int32_t *o1, const float *d1, const int32_t *o2,
                       const float *d2, const int32_t *od, const float *A, const float *B) {
    float R1[ORC_REG_SIZE], R2[ORC_REG_SIZE];
    float total = 0;
    memset(R1, 0, sizeof R1);
    memset(R2, 0, sizeof R2);
    for (int i = 0; i < order_n; i++) {
        int c1 = o1[i], c2 = o2[i], dst = od[i] + ORC_REG_SIZE;
        float w1 = 1 - d1[i];
        float w2 = 1 - d2[i];
        w1 = (w1 < 0) ? (float)ORC_MIN_DIST : w1;
        w2 = (w2 < 0) ? (float)ORC_MIN_DIST : w2;
        float a1, b1, a2, b2;
        if (c1 >= 0) { a1 = A[c1]; b1 = B[c1]; }
        else { a1 = R1[c1 + ORC_REG_SIZE]; b1 = R2[c1 + ORC_REG_SIZE]; }
        if (c2 >= 0) { a2 = A[c2]; b2 = B[c2]; }
        else { a2 = R1[c2 + ORC_REG_SIZE]; b2 = R2[c2 + ORC_REG_SIZE]; }
        float m1 = (a1 < b1) ? a1 : b1;
        float m2 = (a2 < b2) ? a2 : b2;
        total += m1;
        total += m2;
        R1[dst] = (a1 - m1) * w1 + (a2 - m2) * w2;
        R2[dst] = (b1 - m1) * w1 + (b2 - m2) * w2;
    }
    total /= 100.0;
    total = (total > 1.0) ? 1 : total;
    total = (total < 0.0) ? 0 : total;
    return total;
}

/* The same traversal in FP64 (inputs are the FP32 abundances and weights). */
double orc_calc_sim_f64(int order_n, const int32_t *o1, const float *d1, const int32_t *o2,
                        const float *d2, const int32_t *od, const float *A, const float *B) {
    double R1[ORC_REG_SIZE], R2[ORC_REG_SIZE];
    double total = 0;
    memset(R1, 0, sizeof R1);
    memset(R2, 0, sizeof R2);
    for (int i = 0; i < order_n; i++) {
        int c1 = o1[i], c2 = o2[i], dst = od[i] + ORC_REG_SIZE;
        float w1f = 1 - d1[i], w2f = 1 - d2[i];
        w1f = (w1f < 0) ? (float)ORC_MIN_DIST : w1f;
        w2f = (w2f < 0) ? (float)ORC_MIN_DIST : w2f;
        double w1 = w1f, w2 = w2f;
        double a1, b1, a2, b2;
        if (c1 >= 0) { a1 = A[c1]; b1 = B[c1]; }
        else { a1 = R1[c1 + ORC_REG_SIZE]; b1 = R2[c1 + ORC_REG_SIZE]; }
        if (c2 >= 0) { a2 = A[c2]; b2 = B[c2]; }
        else { a2 = R1[c2 + ORC_REG_SIZE]; b2 = R2[c2 + ORC_REG_SIZE]; }
        double m1 = (a1 < b1) ? a1 : b1;
        double m2 = (a2 < b2) ? a2 : b2;
        total += m1;
        total += m2;
        R1[dst] = (a1 - m1) * w1 + (a2 - m2) * w2;
        R2[dst] = (b1 - m1) * w1 + (b2 - m2) * w2;
    }
    total /= 100.0;
    if (total > 1.0) total = 1.0;
    if (total < 0.0) total = 0.0;
    return total;
}

/*
 * Closed form. Pass 1 finds, for every node, whether its written value is read
 * later (consumed) and, for every register read, which node produced it.
 * Pass 2 computes x for both samples per node and accumulates
 * beta * min(x_a, x_b) per child edge. Returns -1 for schedules that are not
 * trees (a value read twice), which the identity does not cover.
 */
double orc_sharedmin_f64(int order_n, const int32_t *o1, const float *d1, const int32_t *o2,
                         const float *d2, const int32_t *od, const float *A, const float *B) {
    int cur[ORC_REG_SIZE];
    for (int k = 0; k < ORC_REG_SIZE; k++) cur[k] = -1;
    int *reads = (int *)calloc((size_t)order_n + 1, sizeof(int));
    int *prod = (int *)malloc(sizeof(int) * (2 * (size_t)order_n + 1));
    double *xa = (double *)malloc(sizeof(double) * ((size_t)order_n + 1));
    double *xb = (double *)malloc(sizeof(double) * ((size_t)order_n + 1));
    double total = 0;
    int ok = 1;
    for (int i = 0; i < order_n && ok; i++) {
        const int32_t ch[2] = {o1[i], o2[i]};
        for (int j = 0; j < 2; j++) {
            prod[2 * i + j] = -1;
            if (ch[j] < 0) {
                int v = cur[ch[j] + ORC_REG_SIZE];
                if (v < 0) { ok = 0; break; }
                prod[2 * i + j] = v;
                if (++reads[v] > 1) ok = 0;
            }
        }
        cur[od[i] + ORC_REG_SIZE] = i;
    }
    if (ok) {
        for (int i = 0; i < order_n; i++) {
            const int32_t ch[2] = {o1[i], o2[i]};
            const float dd[2] = {d1[i], d2[i]};
            double sa = 0, sb = 0;
            for (int j = 0; j < 2; j++) {
                float wf = 1 - dd[j];
                wf = (wf < 0) ? (float)ORC_MIN_DIST : wf;
                double w = wf;
                double ca, cb;
                if (ch[j] >= 0) { ca = A[ch[j]]; cb = B[ch[j]]; }
                else { ca = xa[prod[2 * i + j]]; cb = xb[prod[2 * i + j]]; }
                double beta = reads[i] > 0 ? 1.0 - w : 1.0;
                total += beta * (ca < cb ? ca : cb);
                sa += w * ca;
                sb += w * cb;
            }
            xa[i] = sa;
            xb[i] = sb;
        }
    }
    free(reads);
    free(prod);
    free(xa);
    free(xb);
    if (!ok) return -1.0;
    total /= 100.0;
    if (total > 1.0) total = 1.0;
    if (total < 0.0) total = 0.0;
    return total;
}

/*
 * comp.h:243-279 for one table row. col_leaf[j] = tree leaf of column j after
 * Check_OTU (or -1), col_cp[j] = copy number used for column j (1 when not
 * corrected or absent). Returns the mapped OTU count.
 */
int orc_load_abd(int n_feat, const float *row, const int32_t *col_leaf, const float *col_cp, int leaf_n,
                 float *abd) {
    memset(abd, 0, sizeof(float) * (size_t)leaf_n);
    char *seen = (char *)calloc((size_t)leaf_n + 1, 1);
    int32_t *touched = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n_feat + 1));
    int nt = 0, mapped = 0;
    float total = 0;
    for (int j = 0; j < n_feat; j++) {
        if (row[j] > 0 && col_leaf[j] >= 0) {
            float x = row[j] / col_cp[j];
            abd[col_leaf[j]] = x;
            touched[nt++] = col_leaf[j];
            total += x;
        }
    }
    total /= 100.0;
    for (int k = 0; k < nt; k++) {
        int32_t l = touched[k];
        if (seen[l]) continue;
        seen[l] = 1;
        abd[l] = abd[l] / total;
        mapped++;
    }
    free(seen);
    free(touched);
    return mapped;
}

/* comp_sam.cpp:321-349: pair list, OpenMP dynamic,1, packed upper triangle.
 * abd is n x leaf_n. pairs_limit > 0 computes only the first pairs_limit pairs
 * of the list (a bounded sample for timing). Returns the pairs computed. */
long orc_all_pairs_f32(int n, int leaf_n, const float *abd, int order_n, const int32_t *o1, const float *d1,
                       const int32_t *o2, const float *d2, const int32_t *od, float *sim_packed, int threads,
                       long pairs_limit) {
    long iter = (long)n * (n - 1) / 2;
    if (pairs_limit > 0 && pairs_limit < iter) iter = pairs_limit;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    (void)threads;
#pragma omp parallel for schedule(dynamic, 1)
    for (long p = 0; p < iter; p++) {
        /* invert p = m*n + k - (m+1)(m+2)/2 */
        long m = 0, base = 0;
        {
            double nn = (double)n;
            double disc = (2 * nn - 1) * (2 * nn - 1) - 8.0 * (double)p;
            m = (long)floor(((2 * nn - 1) - sqrt(disc)) / 2);
            if (m < 0) m = 0;
            while (m > 0 && m * (long)n - m * (m + 1) / 2 > p) m--;
            while ((m + 1) * (long)n - (m + 1) * (m + 2) / 2 <= p) m++;
            base = m * (long)n - m * (m + 1) / 2;
        }
        long k = m + 1 + (p - base);
        sim_packed[p] = orc_calc_sim_f32(order_n, o1, d1, o2, d2, od, abd + (size_t)m * leaf_n,
                                         abd + (size_t)k * leaf_n);
    }
    return iter;
}

/* Same over an explicit pair list (row[i], col[i]). */
void orc_pairs_f32(int leaf_n, const float *abd, int order_n, const int32_t *o1, const float *d1,
                   const int32_t *o2, const float *d2, const int32_t *od, long n_pairs, const int32_t *row,
                   const int32_t *col, float *sim, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    (void)threads;
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < n_pairs; i++)
        sim[i] = orc_calc_sim_f32(order_n, o1, d1, o2, d2, od, abd + (size_t)row[i] * leaf_n,
                                  abd + (size_t)col[i] * leaf_n);
}

void orc_pairs_f64(int leaf_n, const float *abd, int order_n, const int32_t *o1, const float *d1,
                   const int32_t *o2, const float *d2, const int32_t *od, long n_pairs, const int32_t *row,
                   const int32_t *col, double *sim, int closed_form, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    (void)threads;
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < n_pairs; i++) {
        const float *A = abd + (size_t)row[i] * leaf_n, *B = abd + (size_t)col[i] * leaf_n;
        sim[i] = closed_form ? orc_sharedmin_f64(order_n, o1, d1, o2, d2, od, A, B)
                             : orc_calc_sim_f64(order_n, o1, d1, o2, d2, od, A, B);
    }
}

/* comp.h:378-446, unweighted Meta-Storms in FP32, same operation order:
 * presence propagates (a register holds 1 if either child is present);
 * total adds w for every child edge present in both samples, total_count
 * for every child edge present in either; sim = total / total_count. */
float orc_calc_sim_unweight_f32(int order_n, const int32_t *o1, const float *d1, const int32_t *o2,
                                const float *d2, const int32_t *od, const float *A, const float *B) {
    float R1[ORC_REG_SIZE], R2[ORC_REG_SIZE];
    float total = 0, total_count = 0;
    memset(R1, 0, sizeof R1);
    memset(R2, 0, sizeof R2);
    for (int i = 0; i < order_n; i++) {
        int c1 = o1[i], c2 = o2[i], dst = od[i] + ORC_REG_SIZE;
        float w1 = 1 - d1[i];
        float w2 = 1 - d2[i];
        w1 = (w1 < 0) ? (float)ORC_MIN_DIST : w1;
        w2 = (w2 < 0) ? (float)ORC_MIN_DIST : w2;
        float a1, b1, a2, b2;
        if (c1 >= 0) { a1 = (A[c1] > 0) ? 1 : 0; b1 = (B[c1] > 0) ? 1 : 0; }
        else { a1 = R1[c1 + ORC_REG_SIZE]; b1 = R2[c1 + ORC_REG_SIZE]; }
        total += (a1 > 0 && b1 > 0) ? w1 : 0;
        total_count += (a1 > 0 || b1 > 0) ? w1 : 0;
        if (c2 >= 0) { a2 = (A[c2] > 0) ? 1 : 0; b2 = (B[c2] > 0) ? 1 : 0; }
        else { a2 = R1[c2 + ORC_REG_SIZE]; b2 = R2[c2 + ORC_REG_SIZE]; }
        total += (a2 > 0 && b2 > 0) ? w2 : 0;
        total_count += (a2 > 0 || b2 > 0) ? w2 : 0;
        R1[dst] = (a1 > 0 || a2 > 0) ? 1 : 0;
        R2[dst] = (b1 > 0 || b2 > 0) ? 1 : 0;
    }
    total /= total_count;
    total = (total > 1.0) ? 1 : total;
    total = (total < 0.0) ? 0 : total;
    return total;
}

/* The same ratio in FP64 from presence sets: shared / (W_a + W_b - shared),
 * W_s = sum of w over child edges whose subtree holds mass of s. */
double orc_unweight_f64(int order_n, const int32_t *o1, const float *d1, const int32_t *o2, const float *d2,
                        const int32_t *od, const float *A, const float *B) {
    double R1[ORC_REG_SIZE], R2[ORC_REG_SIZE];
    double shared = 0, wa = 0, wb = 0;
    memset(R1, 0, sizeof R1);
    memset(R2, 0, sizeof R2);
    for (int i = 0; i < order_n; i++) {
        const int32_t ch[2] = {o1[i], o2[i]};
        const float dd[2] = {d1[i], d2[i]};
        double pa[2], pb[2];
        for (int j = 0; j < 2; j++) {
            float wf = 1 - dd[j];
            wf = (wf < 0) ? (float)ORC_MIN_DIST : wf;
            if (ch[j] >= 0) { pa[j] = A[ch[j]] > 0; pb[j] = B[ch[j]] > 0; }
            else { pa[j] = R1[ch[j] + ORC_REG_SIZE]; pb[j] = R2[ch[j] + ORC_REG_SIZE]; }
            if (pa[j] > 0 && pb[j] > 0) shared += wf;
            if (pa[j] > 0) wa += wf;
            if (pb[j] > 0) wb += wf;
        }
        R1[od[i] + ORC_REG_SIZE] = (pa[0] > 0 || pa[1] > 0);
        R2[od[i] + ORC_REG_SIZE] = (pb[0] > 0 || pb[1] > 0);
    }
    double t = shared / (wa + wb - shared);
    if (t > 1.0) t = 1.0;
    if (t < 0.0) t = 0.0;
    return t;
}

void orc_pairs_unweight(int leaf_n, const float *abd, int order_n, const int32_t *o1, const float *d1,
                        const int32_t *o2, const float *d2, const int32_t *od, long n_pairs, const int32_t *row,
                        const int32_t *col, float *sim32, double *sim64, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    (void)threads;
#pragma omp parallel for schedule(dynamic, 1)
    for (long i = 0; i < n_pairs; i++) {
        const float *A = abd + (size_t)row[i] * leaf_n, *B = abd + (size_t)col[i] * leaf_n;
        if (sim32) sim32[i] = orc_calc_sim_unweight_f32(order_n, o1, d1, o2, d2, od, A, B);
        if (sim64) sim64[i] = orc_unweight_f64(order_n, o1, d1, o2, d2, od, A, B);
    }
}

/* ------------------------------------------------------------------------
 * Non-tree metrics of comp -M 2..5: dist.h:13-159 (Calc_Dist_Cos,
 * Calc_Dist_E, Calc_Dist_JSD, Calc_Dist_Bray_Curtis) as dispatched by
 * _Comp_Tree::Calc_sim(a, b, mode) (comp.h:296-330): similarity =
 * (float)(1.0 - dist). FP32 with the reference's operation order; where the
 * reference calls pow(float, 2) the square is taken in double and added to a
 * float (C++ promotes pow(float, int) to double); sqrt and log of a float are
 * the float functions (the C++ overloads), the JSD's final sqrt is double.
 */
static float orc_dist_cos(const float *m, const float *n, int dim) {
    float f = 0.0f, fm_sum = 0.0f, fn_sum = 0.0f;
    for (int i = 0; i < dim; i++) {
        const float fm = m[i], fn = n[i];
        f += fm * fn;
        fm_sum += fm * fm;
        fn_sum += fn * fn;
    }
    if (fm_sum * fn_sum == 0.0f) return 1.0f;
    f = f / sqrtf(fm_sum * fn_sum);
    return 1.0f - f;
}

static float orc_dist_e(const float *m, const float *n, int dim) {
    float sm = 0.0f, sn = 0.0f;
    for (int i = 0; i < dim; i++) {
        sm += m[i];
        sn += n[i];
    }
    if (sm <= 0.0f) return 1.0f;
    if (sn <= 0.0f) return 1.0f;
    float f = 0.0f;
    for (int i = 0; i < dim; i++) {
        const float d = m[i] / sm - n[i] / sn;
        f = (float)((double)f + (double)d * (double)d);
    }
    return sqrtf(f);
}

static float orc_dist_jsd(const float *m, const float *n, int dim) {
    float sm = 0.0f, sn = 0.0f;
    for (int i = 0; i < dim; i++) {
        sm += m[i];
        sn += n[i];
    }
    float dkl_m = 0.0f, dkl_n = 0.0f;
    for (int i = 0; i < dim; i++) {
        const float p = m[i] / sm, q = n[i] / sn;
        if ((p == 0.0f) && (q == 0.0f)) continue;
        const float h = (p + q) / 2;
        if (p > 0.0f) dkl_m += p * logf(p / h);
        if (q > 0.0f) dkl_n += q * logf(q / h);
    }
    return (float)sqrt((double)(dkl_m + dkl_n) / 2.0);
}

static float orc_dist_bc(const float *m, const float *n, int dim) {
    float sm = 0.0f, sn = 0.0f;
    for (int i = 0; i < dim; i++) {
        sm += m[i];
        sn += n[i];
    }
    if (sm <= 0.0f) return 1.0f;
    if (sn <= 0.0f) return 1.0f;
    float sum = 0.0f, diff = 0.0f;
    for (int i = 0; i < dim; i++) {
        const float p = m[i] / sm, q = n[i] / sn;
        sum += (p + q);
        float d = p - q;
        if (d < 0.0f) d = d * (-1.0f);
        diff += d;
    }
    if (sum <= 0.0f) return 1.0f;
    return diff / sum;
}

float orc_leaf_sim_f32(int mode, int dim, const float *a, const float *b) {
    float d;
    switch (mode) {
    case 2: d = orc_dist_cos(a, b, dim); break;
    case 3: d = orc_dist_e(a, b, dim); break;
    case 5: d = orc_dist_bc(a, b, dim); break;
    default: d = orc_dist_jsd(a, b, dim); break;
    }
    return (float)(1.0 - (double)d);
}

/* The same four similarities from their definitions in FP64 (normalised
 * p = a / sum a in double). Degenerate samples follow the reference: an
 * all-zero sample gives similarity 0 for Cos, Euclidean and Bray-Curtis; for
 * JSD, 1 when both are all-zero and NaN when exactly one is. */
double orc_leaf_sim_f64(int mode, int dim, const float *a, const float *b) {
    double sa = 0.0, sb = 0.0, dot = 0.0, na = 0.0, nb = 0.0;
    for (int i = 0; i < dim; i++) {
        sa += a[i];
        sb += b[i];
        dot += (double)a[i] * b[i];
        na += (double)a[i] * a[i];
        nb += (double)b[i] * b[i];
    }
    if (mode == 2) return (na * nb == 0.0) ? 0.0 : dot / sqrt(na * nb);
    if (mode == 3 || mode == 5) {
        if (sa <= 0.0 || sb <= 0.0) return 0.0;
        double f = 0.0, sum = 0.0;
        for (int i = 0; i < dim; i++) {
            const double p = a[i] / sa, q = b[i] / sb;
            f += mode == 3 ? (p - q) * (p - q) : fabs(p - q);
            sum += p + q;
        }
        return mode == 3 ? 1.0 - sqrt(f) : 1.0 - f / sum;
    }
    if (sa == 0.0 && sb == 0.0) return 1.0;
    if (sa == 0.0 || sb == 0.0) return NAN;
    double dk = 0.0;
    for (int i = 0; i < dim; i++) {
        const double p = a[i] / sa, q = b[i] / sb;
        if (p == 0.0 && q == 0.0) continue;
        const double h = (p + q) / 2.0;
        if (p > 0.0) dk += p * log(p / h);
        if (q > 0.0) dk += q * log(q / h);
    }
    return 1.0 - sqrt(dk / 2.0);
}

void orc_pairs_leaf(int mode, int leaf_n, const float *abd, long n_pairs, const int32_t *row, const int32_t *col,
                    float *sim32, double *sim64, int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#endif
    (void)threads;
#pragma omp parallel for schedule(dynamic, 16)
    for (long i = 0; i < n_pairs; i++) {
        const float *A = abd + (size_t)row[i] * leaf_n, *B = abd + (size_t)col[i] * leaf_n;
        if (sim32) sim32[i] = orc_leaf_sim_f32(mode, leaf_n, A, B);
        if (sim64) sim64[i] = orc_leaf_sim_f64(mode, leaf_n, A, B);
    }
}
