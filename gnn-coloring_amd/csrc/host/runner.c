/*
 * runner.c -- the reference's experiment harness (runner.h) over the GPU
 * path. Timing brackets a complete call (host->HBM copies, ordering,
 * colouring, results back), best of `it`, like runner.c:105-169; the
 * Culberson columns accumulate as in runner.c:123 and 149. Thread-count
 * columns are kept for output compatibility and all time the GPU path.
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "coloring.h"
#include "runner.h"

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

void runner_test_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, const char *name, int k,
                          void (*f)(uint32_t, const uint32_t *, const uint32_t *, double *))
{
    double *p = malloc(sizeof(double) * (N + 8ull));
    int *color = malloc(sizeof(int) * (N + 1ull));
    void *jp = jones_plassmann_setup(N, V[N]);
    printf("%s: ", name);
    fflush(stdout);
    f(N, V, E, p);
    jones_plassmann_internal(N, V, E, jp, p, color);
    printf("%d ", validate_coloring_par(N, V, E, color));
    fflush(stdout);
    for (int i = 1; i < k; i++) {
        for (uint32_t j = 0; j < N; j++)
            p[j] = color[j];
        jones_plassmann_internal(N, V, E, jp, p, color);
        printf("%d ", validate_coloring_par(N, V, E, color));
        fflush(stdout);
    }
    printf("\n");
    jones_plassmann_cleanup(jp);
    free(p);
    free(color);
}

void runner_test_ordering_time(uint32_t N, const uint32_t *V, const uint32_t *E, const char *name,
                               void (*f)(uint32_t, const uint32_t *, const uint32_t *, double *))
{
    double *p = malloc(sizeof(double) * (N + 8ull));
    int *color = malloc(sizeof(int) * (N + 1ull));
    void *gd = greedy_setup(N, V[N]);
    printf("%s: ", name);
    fflush(stdout);
    const double t0 = now_s();
    f(N, V, E, p);
    greedy_internal(N, V, E, gd, p, color);
    const double t1 = now_s();
    printf("%d %.4lf\n", validate_coloring_par(N, V, E, color), t1 - t0);
    greedy_cleanup(gd);
    free(p);
    free(color);
}

/* runner.c:74-83 */
static void check_count(uint32_t N, const uint32_t *V, const uint32_t *E, int *colors, int *C)
{
    const int c = validate_coloring_par(N, V, E, colors);
    if (c == 0)
        printf("Invalid coloring\n");
    else if (*C < 0)
        *C = c;
    else if (*C != c)
        printf("Unexpected number of colors (%d %d)\n", *C, c);
}

void runner_test_full(uint32_t N, const uint32_t *V, const uint32_t *E,
                      void *(*setup)(uint32_t, uint32_t), void (*cleanup)(void *),
                      void (*f_seq)(uint32_t, const uint32_t *, const uint32_t *, void *, void *,
                                    double *, int *),
                      void (*f_par)(uint32_t, const uint32_t *, const uint32_t *, void *, void *,
                                    double *, int *),
                      const char *s, int it, int k, int t_tot, int ntc, const int *nt)
{
    (void)t_tot;
    (void)nt;
    if (it < 1)
        it = 1;
    if (k < 1)
        k = 1;
    const int cols = ntc + 1;
    int *n_colors = malloc(sizeof(int) * (size_t)k);
    double *times = malloc(sizeof(double) * (size_t)it * (size_t)k * (size_t)cols);
#define T(t, i, j) times[((size_t)(t) * (size_t)k + (size_t)(i)) * (size_t)cols + (size_t)(j)]
    for (int i = 0; i < k; i++)
        n_colors[i] = -1;
    int *colors = malloc(sizeof(int) * (N + 1ull));
    double *p = malloc(sizeof(double) * (N + 8ull));
    void *fd = setup(N, V[N]);
    void *jp = jones_plassmann_setup(N, V[N]);
    void *gd = greedy_setup(N, V[N]);
    for (int t = 0; t < it; t++) {
        double t0 = now_s();
        f_seq(N, V, E, fd, gd, p, colors);
        double t1 = now_s();
        T(t, 0, 0) = t1 - t0;
        check_count(N, V, E, colors, &n_colors[0]);
        for (int i = 1; i < k; i++) {
            t0 = now_s();
            for (uint32_t u = 0; u < N; u++)
                p[u] = colors[u];
            greedy_internal(N, V, E, gd, p, colors);
            t1 = now_s();
            T(t, i, 0) = (t1 - t0) + T(t, i - 1, 0);
            check_count(N, V, E, colors, &n_colors[i]);
        }
        for (int c = 0; c < ntc; c++) {
            t0 = now_s();
            f_par(N, V, E, fd, jp, p, colors);
            t1 = now_s();
            T(t, 0, c + 1) = t1 - t0;
            check_count(N, V, E, colors, &n_colors[0]);
            for (int j = 1; j < k; j++) {
                t0 = now_s();
                for (uint32_t u = 0; u < N; u++)
                    p[u] = colors[u];
                jones_plassmann_internal(N, V, E, jp, p, colors);
                t1 = now_s();
                T(t, j, c + 1) = (t1 - t0) + T(t, j - 1, c + 1);
                check_count(N, V, E, colors, &n_colors[j]);
            }
        }
    }
    for (int i = 0; i < k; i++) {
        printf("%s_%d %d ", s, i, n_colors[i]);
        for (int j = 0; j < cols; j++) {
            double best = 9999.9;
            for (int t = 0; t < it; t++)
                if (T(t, i, j) < best)
                    best = T(t, i, j);
            printf("%.6lf ", best);
        }
        printf("\n");
    }
#undef T
    free(colors);
    free(p);
    free(n_colors);
    free(times);
    cleanup(fd);
    jones_plassmann_cleanup(jp);
    greedy_cleanup(gd);
}
