/*
 * spmv.c -- FP64 oracle SpMV/SpMM and the threaded CPU CSR SpMV baseline.
 * TEST INFRASTRUCTURE ONLY.
 *
 * The reference's SpMV ground truth is cuSPARSE CSR_ALG1 (src/class.cu:182-207),
 * a closed library whose bits cannot be restated; the oracle instead computes
 * y = A x exactly rounded up to compensated-summation error (Neumaier), and
 * absrow[i] = sum_j |a_ij x_j| for the condition-aware parity metric
 * max_i |y_i - yhat_i| / absrow_i.  The reference has no CPU SpMV; the threaded
 * FP32 kernel is the CPU baseline of the benchmark (static row partition
 * balanced by nonzeros, one pthread per core).
 *
 * Pinned by: scipy.sparse csr_matrix @ x on random matrices, including empty
 * rows and columns (tests/test_oracle_basic.py).
 */
#include "oracle.h"
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <time.h>

static void neumaier_add(double* s, double* c, double v)
{
    const double t = *s + v;
    if (fabs(*s) >= fabs(v)) *c += (*s - t) + v;
    else *c += (v - t) + *s;
    *s = t;
}

void oracle_spmv_f64(int32_t n_rows, const int32_t* rp, const int32_t* ci,
                     const double* val, const double* x, double* y, double* absrow)
{
    for (int32_t i = 0; i < n_rows; ++i) {
        double s = 0.0, c = 0.0, a = 0.0;
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
            const double p = val[k] * x[ci[k]];
            /* exact product error via fma keeps the sum faithful */
            const double pe = fma(val[k], x[ci[k]], -p);
            neumaier_add(&s, &c, p);
            c += pe;
            a += fabs(p);
        }
        y[i] = s + c;
        if (absrow) absrow[i] = a;
    }
}

void oracle_spmm_f64(int32_t n_rows, int32_t k, const int32_t* rp, const int32_t* ci,
                     const double* val, const double* x, double* y)
{
    for (int32_t i = 0; i < n_rows; ++i)
        for (int32_t j = 0; j < k; ++j) {
            double s = 0.0, c = 0.0;
            for (int32_t e = rp[i]; e < rp[i + 1]; ++e) {
                const double p = val[e] * x[(int64_t)ci[e] * k + j];
                neumaier_add(&s, &c, p);
                c += fma(val[e], x[(int64_t)ci[e] * k + j], -p);
            }
            y[(int64_t)i * k + j] = s + c;
        }
}

typedef struct {
    const int32_t* rp; const int32_t* ci; const float* val; const float* x; float* y;
    int32_t r0, r1, iters;
} spmv_job;

static void* spmv_worker(void* arg)
{
    spmv_job* j = (spmv_job*)arg;
    for (int32_t it = 0; it < j->iters; ++it)
        for (int32_t i = j->r0; i < j->r1; ++i) {
            float s = 0.0f;
            for (int32_t k = j->rp[i]; k < j->rp[i + 1]; ++k) s += j->val[k] * j->x[j->ci[k]];
            j->y[i] = s;
        }
    return NULL;
}

double oracle_spmv_f32_threads(int32_t n_rows, const int32_t* rp, const int32_t* ci,
                               const float* val, const float* x, float* y,
                               int32_t threads, int32_t iters)
{
    if (threads < 1) threads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    spmv_job* jobs = (spmv_job*)malloc(sizeof(spmv_job) * (size_t)threads);
    const int64_t nnz = rp[n_rows];
    int32_t r = 0;
    for (int32_t t = 0; t < threads; ++t) {
        /* first row whose start reaches the t-th nnz share */
        const int64_t target = nnz * (t + 1) / threads;
        int32_t e = r;
        if (t == threads - 1) e = n_rows;
        else while (e < n_rows && rp[e] < target) ++e;
        jobs[t] = (spmv_job){rp, ci, val, x, y, r, e, iters};
        r = e;
    }
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    for (int32_t t = 0; t < threads; ++t) pthread_create(&th[t], NULL, spmv_worker, &jobs[t]);
    for (int32_t t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &b);
    free(th); free(jobs);
    return (double)(b.tv_sec - a.tv_sec) * 1e3 + (double)(b.tv_nsec - a.tv_nsec) * 1e-6;
}

/* FP64 of the same: the CPU baseline's second precision */
typedef struct {
    const int32_t* rp; const int32_t* ci; const double* val; const double* x; double* y;
    int32_t r0, r1, iters;
} spmv_job64;

static void* spmv_worker64(void* arg)
{
    spmv_job64* j = (spmv_job64*)arg;
    for (int32_t it = 0; it < j->iters; ++it)
        for (int32_t i = j->r0; i < j->r1; ++i) {
            double s = 0.0;
            for (int32_t k = j->rp[i]; k < j->rp[i + 1]; ++k) s += j->val[k] * j->x[j->ci[k]];
            j->y[i] = s;
        }
    return NULL;
}

double oracle_spmv_f64_threads(int32_t n_rows, const int32_t* rp, const int32_t* ci,
                               const double* val, const double* x, double* y,
                               int32_t threads, int32_t iters)
{
    if (threads < 1) threads = 1;
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)threads);
    spmv_job64* jobs = (spmv_job64*)malloc(sizeof(spmv_job64) * (size_t)threads);
    const int64_t nnz = rp[n_rows];
    int32_t r = 0;
    for (int32_t t = 0; t < threads; ++t) {
        const int64_t target = nnz * (t + 1) / threads;
        int32_t e = r;
        if (t == threads - 1) e = n_rows;
        else while (e < n_rows && rp[e] < target) ++e;
        jobs[t] = (spmv_job64){rp, ci, val, x, y, r, e, iters};
        r = e;
    }
    struct timespec a, b;
    clock_gettime(CLOCK_MONOTONIC, &a);
    for (int32_t t = 0; t < threads; ++t) pthread_create(&th[t], NULL, spmv_worker64, &jobs[t]);
    for (int32_t t = 0; t < threads; ++t) pthread_join(th[t], NULL);
    clock_gettime(CLOCK_MONOTONIC, &b);
    free(th); free(jobs);
    return (double)(b.tv_sec - a.tv_sec) * 1e3 + (double)(b.tv_nsec - a.tv_nsec) * 1e-6;
}
