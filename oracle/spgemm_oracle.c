/*
 * ORACLE -- test infrastructure only.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library, and only as the
 * checker / CPU baseline; the product path never calls it.
 *
 * A CPU restatement of what the reference's `proposal` SpGEMM computes,
 * written as a row-parallel Gustavson multiply (dense accumulator + marker
 * per thread), FP64 accumulation, compiled with -ffp-contract=off.
 *
 * What each function restates (reference paths are relative to the
 * GPUSpGEMM source tree):
 *  - oracle_nip_rows      NIP_i = sum over k in A(i,:) of nnz(B(k,:))
 *                         src/libproposal/include/proposal/setup.cuh:48-91,
 *                         total as src/libutils/include/utils/nip.cuh:35-47.
 *  - oracle_symbolic      nnz_i = |union of B(k,:) columns| (distinct count,
 *                         sym1.cuh:72-98) and row_ptr = exclusive scan of
 *                         nnz_i (binning.cuh:219-224).  Exact-zero sums are
 *                         structural entries: the count depends only on the
 *                         patterns, never on values.
 *  - oracle_numeric       C.col of each row in ascending order (sym2.cuh:104-120)
 *                         and c_ij = sum_k a_ik * b_kj, each product formed
 *                         first and then added to an accumulator that starts at
 *                         0 (num.cuh:97-121; no fused multiply-add across the
 *                         accumulation).  Also |A||B|_ij for error scaling.
 *
 * Pinning (checked by tests/test_oracle.py): brute-force dense products on
 * small random and rectangular inputs, closed forms of the 27-point stencil
 * (NIP, nnz and integer values of the noise-free case), exactly-8-per-row
 * NIP = 64 per row, kept exact-zero sums, empty rows and empty matrices.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <unistd.h>

typedef struct {
    int64_t m, n;
    const int32_t *a_rpt, *a_col;
    const double *a_val;
    const int32_t *b_rpt, *b_col;
    const double *b_val;
    int64_t *c_rpt;     /* symbolic: out nnz per row in c_rpt[i+1]; numeric: in */
    int32_t *c_col;
    double *c_val, *c_abs;
    int64_t *nip;
    int64_t row_lo, row_hi;
    int status;
} job_t;

static int hw_threads(int threads) {
    if (threads > 0) return threads;
    long n = sysconf(_SC_NPROCESSORS_ONLN);
    return n > 0 ? (int)n : 1;
}

static void run_parallel(void *(*fn)(void *), job_t *proto, int threads) {
    threads = hw_threads(threads);
    if (proto->m < 2048) threads = 1;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    int started[256];
    job_t jobs[256];
    /* small chunks handed out round-robin would balance better; contiguous
       chunks keep each thread's marker array hot and are good enough here */
    int64_t chunk = (proto->m + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        jobs[t] = *proto;
        jobs[t].status = 0;
        jobs[t].row_lo = t * chunk < proto->m ? t * chunk : proto->m;
        jobs[t].row_hi = (t + 1) * chunk < proto->m ? (t + 1) * chunk : proto->m;
        started[t] = 0;
        if (t == 0) continue;
        if (pthread_create(&tid[t], NULL, fn, &jobs[t]) == 0)
            started[t] = 1;
        else
            fn(&jobs[t]);          /* run inline if a thread cannot start */
    }
    fn(&jobs[0]);
    for (int t = 1; t < threads; t++)
        if (started[t]) pthread_join(tid[t], NULL);
    proto->status = 0;
    for (int t = 0; t < threads; t++)
        if (jobs[t].status) proto->status = jobs[t].status;
}

/* ---- NIP per row ------------------------------------------------------ */
static void *nip_worker(void *arg) {
    job_t *j = (job_t *)arg;
    for (int64_t i = j->row_lo; i < j->row_hi; i++) {
        int64_t s = 0;
        for (int32_t p = j->a_rpt[i]; p < j->a_rpt[i + 1]; p++) {
            int32_t k = j->a_col[p];
            s += (int64_t)j->b_rpt[k + 1] - j->b_rpt[k];
        }
        j->nip[i] = s;
    }
    return NULL;
}

/* nip[i] for every row of A; returns the total. */
int64_t oracle_nip_rows(int64_t m, const int32_t *a_rpt, const int32_t *a_col,
                        const int32_t *b_rpt, int64_t *nip, int threads) {
    job_t p;
    memset(&p, 0, sizeof p);
    p.m = m; p.a_rpt = a_rpt; p.a_col = a_col; p.b_rpt = b_rpt; p.nip = nip;
    run_parallel(nip_worker, &p, threads);
    int64_t total = 0;
    for (int64_t i = 0; i < m; i++) total += nip[i];
    return total;
}

/* ---- symbolic: distinct columns per row -------------------------------- */
static void *sym_worker(void *arg) {
    job_t *j = (job_t *)arg;
    int64_t *mark = (int64_t *)malloc((size_t)(j->n > 0 ? j->n : 1) * sizeof(int64_t));
    if (!mark) { j->status = 1; return NULL; }
    for (int64_t c = 0; c < j->n; c++) mark[c] = -1;
    for (int64_t i = j->row_lo; i < j->row_hi; i++) {
        int64_t cnt = 0;
        for (int32_t p = j->a_rpt[i]; p < j->a_rpt[i + 1]; p++) {
            int32_t k = j->a_col[p];
            for (int32_t q = j->b_rpt[k]; q < j->b_rpt[k + 1]; q++) {
                int32_t c = j->b_col[q];
                if (mark[c] != i) { mark[c] = i; cnt++; }
            }
        }
        j->c_rpt[i + 1] = cnt;
    }
    free(mark);
    return NULL;
}

/* Fills c_rpt[0..m] (exclusive scan of nnz per row); returns nnz(C) or -1. */
int64_t oracle_symbolic(int64_t m, int64_t n, const int32_t *a_rpt, const int32_t *a_col,
                        const int32_t *b_rpt, const int32_t *b_col, int64_t *c_rpt,
                        int threads) {
    job_t p;
    memset(&p, 0, sizeof p);
    p.m = m; p.n = n; p.a_rpt = a_rpt; p.a_col = a_col; p.b_rpt = b_rpt; p.b_col = b_col;
    p.c_rpt = c_rpt;
    c_rpt[0] = 0;
    run_parallel(sym_worker, &p, threads);
    if (p.status) return -1;
    for (int64_t i = 0; i < m; i++) c_rpt[i + 1] += c_rpt[i];
    return c_rpt[m];
}

/* ---- numeric: sorted columns and values -------------------------------- */
static int cmp_i32(const void *x, const void *y) {
    int32_t a = *(const int32_t *)x, b = *(const int32_t *)y;
    return (a > b) - (a < b);
}

static void *num_worker(void *arg) {
    job_t *j = (job_t *)arg;
    size_t n = (size_t)(j->n > 0 ? j->n : 1);
    int64_t *mark = (int64_t *)malloc(n * sizeof(int64_t));
    double *acc = (double *)malloc(n * sizeof(double));
    double *aacc = (double *)malloc(n * sizeof(double));
    if (!mark || !acc || !aacc) { free(mark); free(acc); free(aacc); j->status = 1; return NULL; }
    for (int64_t c = 0; c < j->n; c++) mark[c] = -1;
    for (int64_t i = j->row_lo; i < j->row_hi; i++) {
        int64_t base = j->c_rpt[i], len = 0, cap = j->c_rpt[i + 1] - base;
        int32_t *cols = j->c_col + base;
        for (int32_t p = j->a_rpt[i]; p < j->a_rpt[i + 1]; p++) {
            int32_t k = j->a_col[p];
            double a = j->a_val[p];
            for (int32_t q = j->b_rpt[k]; q < j->b_rpt[k + 1]; q++) {
                int32_t c = j->b_col[q];
                double prod = a * j->b_val[q];
                if (mark[c] != i) {
                    if (len == cap) { j->status = 2; goto done; }
                    mark[c] = i;
                    acc[c] = 0.0;
                    aacc[c] = 0.0;
                    cols[len++] = c;
                }
                acc[c] = acc[c] + prod;
                aacc[c] = aacc[c] + (prod < 0 ? -prod : prod);
            }
        }
        if (len != cap) { j->status = 2; break; }
        qsort(cols, (size_t)len, sizeof(int32_t), cmp_i32);
        for (int64_t t = 0; t < len; t++) {
            j->c_val[base + t] = acc[cols[t]];
            if (j->c_abs) j->c_abs[base + t] = aacc[cols[t]];
        }
    }
done:
    free(mark); free(acc); free(aacc);
    return NULL;
}

/* Needs c_rpt from oracle_symbolic; fills c_col, c_val and (optional) c_abs. */
int oracle_numeric(int64_t m, int64_t n, const int32_t *a_rpt, const int32_t *a_col,
                   const double *a_val, const int32_t *b_rpt, const int32_t *b_col,
                   const double *b_val, const int64_t *c_rpt, int32_t *c_col,
                   double *c_val, double *c_abs, int threads) {
    job_t p;
    memset(&p, 0, sizeof p);
    p.m = m; p.n = n; p.a_rpt = a_rpt; p.a_col = a_col; p.a_val = a_val;
    p.b_rpt = b_rpt; p.b_col = b_col; p.b_val = b_val;
    p.c_rpt = (int64_t *)c_rpt; p.c_col = c_col; p.c_val = c_val; p.c_abs = c_abs;
    run_parallel(num_worker, &p, threads);
    return p.status;
}

int oracle_hw_threads(void) { return hw_threads(0); }

/* ---- CPU baseline ------------------------------------------------------
 * The product alone, as a caller would time it: the count pass and scan
 * into a fresh row-pointer array, then sorted columns and values into
 * freshly allocated C arrays (no |A||B|, no NIP), `warmups` untimed runs
 * and `runs` timed ones on `threads` threads; wall-clock seconds of each
 * timed run go to times[].  Returns 0, or nonzero when memory runs out. */
#include <time.h>
static double now_seconds(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int oracle_time_product(int64_t m, int64_t n, const int32_t *a_rpt, const int32_t *a_col,
                        const double *a_val, const int32_t *b_rpt, const int32_t *b_col,
                        const double *b_val, int threads, int warmups, int runs, double *times) {
    for (int r = -warmups; r < runs; r++) {
        const double t0 = now_seconds();
        int64_t *c_rpt = (int64_t *)malloc((size_t)(m + 1) * sizeof(int64_t));
        if (!c_rpt) return 1;
        const int64_t nnz = oracle_symbolic(m, n, a_rpt, a_col, b_rpt, b_col, c_rpt, threads);
        int32_t *c_col = (int32_t *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int32_t));
        double *c_val = (double *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
        int st = nnz < 0 || !c_col || !c_val;
        if (!st) st = oracle_numeric(m, n, a_rpt, a_col, a_val, b_rpt, b_col, b_val, c_rpt, c_col, c_val,
                                     NULL, threads);
        free(c_rpt); free(c_col); free(c_val);
        if (st) return 1;
        if (r >= 0) times[r] = now_seconds() - t0;
    }
    return 0;
}
