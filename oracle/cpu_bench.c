/* CPU baseline driver (test infrastructure, run only by bench.py's cpu_baseline leg): times the
 * oracle's OpenMP restatements of the reference's CPU kernels on the host, built there with the
 * reference's release flags (-O3 -fopenmp -march=native), and prints one JSON line.
 *
 *   cpu_bench elasticity NZ BUDGET_S   sshex8 linear elasticity, level 16, on a 24 x 24 x NZ slab of
 *                                     the 24^3 macro unit cube (the per-element work of the L=16,
 *                                     24^3 workload): affine_sshex8_linear_elasticity_apply_macro_adjugate
 *                                     (operators/sshex8/sshex8_linear_elasticity.cpp:275-425)
 *   cpu_bench laplacian N BUDGET_S    sshex8 Laplacian, level 8, on an N^3 macro unit cube:
 *                                     affine_sshex8_laplacian_stencil_apply_fff
 *                                     (operators/sshex8/sshex8_laplacian.cpp:698-769)
 *
 * Protocol of drivers/bench/bench_op.exe.cpp:34-56: 2 warm-up applies, then timed applies (at least
 * 5, at most 50, until BUDGET_S seconds), median reported; plus a STREAM triad of the host. */
#define _GNU_SOURCE
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "sfem_oracle.h"

static double now(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int cmp_double(const void *a, const void *b) {
    const double x = *(const double *)a, y = *(const double *)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* xorshift64*: the per-element coefficients (deterministic, seed 1) */
static double uniform01(uint64_t *s) {
    *s ^= *s >> 12;
    *s ^= *s << 25;
    *s ^= *s >> 27;
    return (double)((*s * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
}

static void cpu_model(char *out, size_t n) {
    snprintf(out, n, "unknown");
    FILE *f = fopen("/proc/cpuinfo", "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof line, f)) {
        if (strncmp(line, "model name", 10) == 0) {
            char *p = strchr(line, ':');
            if (p) {
                p++;
                while (*p == ' ') p++;
                p[strcspn(p, "\n")] = 0;
                for (char *q = p; *q; q++)
                    if (*q == '"' || *q == '\\') *q = ' ';
                snprintf(out, n, "%s", p);
            }
            break;
        }
    }
    fclose(f);
}

static double stream_triad_gbs(void) {
    const ptrdiff_t n = (ptrdiff_t)1 << 26; /* 3 x 512 MiB */
    double *a = malloc(sizeof(double) * n), *b = malloc(sizeof(double) * n), *c = malloc(sizeof(double) * n);
    if (!a || !b || !c) return -1;
#pragma omp parallel for schedule(static)
    for (ptrdiff_t i = 0; i < n; i++) {
        a[i] = 0;
        b[i] = 1;
        c[i] = 2;
    }
    double best = 1e30;
    for (int r = 0; r < 5; r++) {
        const double t0 = now();
        oracle_stream_triad(n, a, b, c, 3.0);
        const double t = now() - t0;
        if (t < best) best = t;
    }
    free(a);
    free(b);
    free(c);
    return 3.0 * 8.0 * (double)n / best / 1e9;
}

int main(int argc, char **argv) {
    if (argc < 4) {
        fprintf(stderr, "usage: %s elasticity NZ BUDGET_S | laplacian N BUDGET_S\n", argv[0]);
        return 2;
    }
    const int elasticity = strcmp(argv[1], "elasticity") == 0;
    const int n = atoi(argv[2]);
    const double budget = atof(argv[3]);
    const int L = elasticity ? 16 : 8;
    const int nx = elasticity ? 24 : n, ny = elasticity ? 24 : n, nz = n;
    /* macro mesh: the box of the full workload's spacing (1/24 for the elasticity slab) */
    const double h = elasticity ? 1.0 / 24 : 1.0 / n;
    const ptrdiff_t nn = oracle_box_nnodes(nx + 1, ny + 1, nz + 1), ne = oracle_box_nelements(nx + 1, ny + 1, nz + 1, 0);
    int32_t *mel = malloc(sizeof(int32_t) * 8 * ne);
    float *mpts = malloc(sizeof(float) * 3 * nn);
    oracle_box_mesh(nx + 1, ny + 1, nz + 1, h * nx, h * ny, h * nz, 0, mel, mpts);
    const int nxe = oracle_sshex8_nxe(L);
    const ptrdiff_t nnodes = oracle_sshex8_refine(L, ne, nn, mel, mpts, NULL, NULL);
    int32_t *el = malloc(sizeof(int32_t) * (size_t)nxe * ne);
    float *pts = malloc(sizeof(float) * 3 * (size_t)nnodes);
    oracle_sshex8_refine(L, ne, nn, mel, mpts, el, pts);
    const int corners[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    const int bs = elasticity ? 3 : 1;
    const ptrdiff_t ndofs = nnodes * bs;
    double *x = malloc(sizeof(double) * ndofs), *y = malloc(sizeof(double) * ndofs);
    uint64_t seed = 0x9e3779b97f4a7c15ull;
    for (ptrdiff_t i = 0; i < ndofs; i++) {
        x[i] = 2 * uniform01(&seed) - 1;
        y[i] = 0;
    }
    float *coeff = malloc(sizeof(float) * ne);
    for (ptrdiff_t e = 0; e < ne; e++) coeff[e] = (float)pow(10.0, 2 * uniform01(&seed) - 1); /* [0.1, 10] */
    float *adj = NULL, *det = NULL, *fff = NULL;
    if (elasticity) {
        adj = malloc(sizeof(float) * 9 * ne);
        det = malloc(sizeof(float) * ne);
        oracle_hex8_adjugate_det(ne, mel, corners, nn, mpts, adj, det);
    } else {
        fff = malloc(sizeof(float) * 6 * ne);
        oracle_hex8_fff(ne, mel, corners, nn, mpts, fff);
    }
    const double nu = 0.3, mu = 1.0 / (2 * (1 + nu)), lambda = nu / ((1 + nu) * (1 - 2 * nu));
#define APPLY()                                                                                  \
    do {                                                                                         \
        if (elasticity)                                                                          \
            oracle_sshex8_elasticity_apply(L, ne, el, adj, det, mu, lambda, coeff, x, y);        \
        else                                                                                     \
            oracle_sshex8_laplacian_stencil_apply(L, ne, el, fff, NULL, x, y);                  \
    } while (0)
    for (int w = 0; w < 2; w++) APPLY();
    double times[50];
    int reps = 0;
    const double t_start = now();
    while (reps < 50 && (reps < 5 || now() - t_start < budget)) {
        const double t0 = now();
        APPLY();
        times[reps++] = now() - t0;
    }
    qsort(times, reps, sizeof(double), cmp_double);
    const double med = times[reps / 2];
    double checksum = 0;
    for (ptrdiff_t i = 0; i < ndofs; i++) checksum += y[i];
    char model[256];
    cpu_model(model, sizeof model);
    const double triad = stream_triad_gbs();
    printf("{\"workload\": \"%s\", \"level\": %d, \"macro_elements\": %ld, \"dofs\": %ld, \"median_s\": %.6g, "
           "\"dof_per_s\": %.6g, \"reps\": %d, \"threads\": %d, \"cpu_model\": \"%s\", \"stream_triad_GBs\": %.4g, "
           "\"checksum\": %.6g}\n",
           elasticity ? "sshex8 linear elasticity" : "sshex8 Laplacian (stencil path)", L, (long)ne, (long)ndofs, med,
           (double)ndofs / med, reps, omp_get_max_threads(), model, triad, checksum);
    return 0;
}
