/*
 * runner.c -- the experiment harness of runner.h over the GPU path.
 *
 * Everything here is built on one helper, run_chain(): a "chain" is a first
 * colouring followed by passes-1 Culberson recolourings (priorities = the
 * previous colours, so higher colour classes go first; main.c:32-45). The
 * helper times the first colouring and each recolouring separately and
 * reports, per pass, the cumulative seconds and the validated colour count.
 *
 * runner_test_full keeps a table of those cumulative seconds indexed by
 * (pass, column, trial): column 0 is the sequential entry points (f_seq,
 * then greedy_internal), columns 1..ntc the parallel ones (f_par, then
 * jones_plassmann_internal). The reference's thread counts only label the
 * columns: every column runs the same GPU path. What is printed is the
 * smallest time over the trials, in the reference's line formats
 * (runner.c:15-38, 52-67, 156-169), which existing scripts parse.
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "coloring.h"
#include "runner.h"

typedef void (*ordering_fn)(uint32_t, const uint32_t *, const uint32_t *, double *);
typedef void (*heuristic_fn)(uint32_t, const uint32_t *, const uint32_t *, void *, void *,
                             double *, int *);
typedef void (*recolor_fn)(uint32_t, const uint32_t *, const uint32_t *, void *, const double *,
                           int *);

/* The graph and the caller-visible buffers of one experiment. */
typedef struct {
    uint32_t N;
    const uint32_t *V, *E;
    double *p;  /* N + 8 priorities: the GNN writes rows past N */
    int *color; /* N colours */
} bench_graph;

static double wall_seconds(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

static int bench_graph_open(bench_graph *b, uint32_t N, const uint32_t *V, const uint32_t *E)
{
    b->N = N;
    b->V = V;
    b->E = E;
    b->p = malloc(sizeof(double) * ((size_t)N + 8));
    b->color = malloc(sizeof(int) * ((size_t)N + 1));
    if (!b->p || !b->color) {
        free(b->p);
        free(b->color);
        fprintf(stderr, "runner: out of host memory for %u vertices\n", N);
        return 0;
    }
    return 1;
}

static void bench_graph_close(bench_graph *b)
{
    free(b->p);
    free(b->color);
}

/* The first colouring of a chain: either an ordering plus the colouring
 * entry point (runner_test_ordering*), or a heuristic that does both. */
typedef struct {
    ordering_fn order;  /* when set: order(p), then colour with `again` */
    heuristic_fn both;  /* otherwise: both(state, handle, p, color) */
    void *state;        /* heuristic state from its setup */
    void *handle;       /* colouring handle passed to `both` / `again` */
    recolor_fn again;   /* colouring entry point for the recolourings */
} chain_step;

/* What run_chain reports per pass: i = pass index, colors = the validated
 * count (0 = improper), seconds = cumulative time through pass i. */
typedef void (*pass_report)(void *ctx, int i, int colors, double seconds);

static void run_chain(bench_graph *b, const chain_step *st, int passes, pass_report report,
                      void *ctx)
{
    double t0 = wall_seconds();
    if (st->order) {
        st->order(b->N, b->V, b->E, b->p);
        st->again(b->N, b->V, b->E, st->handle, b->p, b->color);
    } else {
        st->both(b->N, b->V, b->E, st->state, st->handle, b->p, b->color);
    }
    double total = wall_seconds() - t0;
    report(ctx, 0, validate_coloring_par(b->N, b->V, b->E, b->color), total);
    for (int i = 1; i < passes; i++) {
        t0 = wall_seconds();
        for (uint32_t u = 0; u < b->N; u++)
            b->p[u] = (double)b->color[u];
        st->again(b->N, b->V, b->E, st->handle, b->p, b->color);
        total += wall_seconds() - t0;
        report(ctx, i, validate_coloring_par(b->N, b->V, b->E, b->color), total);
    }
}

/* ---------------------------------------------------------------- orderings */

static void print_count(void *ctx, int i, int colors, double seconds)
{
    (void)ctx;
    (void)i;
    (void)seconds;
    printf("%d ", colors);
    fflush(stdout);
}

void runner_test_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, const char *name, int k,
                          void (*f)(uint32_t, const uint32_t *, const uint32_t *, double *))
{
    bench_graph b;
    if (!bench_graph_open(&b, N, V, E))
        return;
    chain_step st = {f, NULL, NULL, jones_plassmann_setup(N, V[N]), jones_plassmann_internal};
    printf("%s: ", name);
    fflush(stdout);
    run_chain(&b, &st, k < 1 ? 1 : k, print_count, NULL);
    printf("\n");
    jones_plassmann_cleanup(st.handle);
    bench_graph_close(&b);
}

/* runner_test_ordering_time reports the first pass only. */
typedef struct {
    int colors;
    double seconds;
} first_pass;

static void keep_first(void *ctx, int i, int colors, double seconds)
{
    if (i == 0) {
        ((first_pass *)ctx)->colors = colors;
        ((first_pass *)ctx)->seconds = seconds;
    }
}

void runner_test_ordering_time(uint32_t N, const uint32_t *V, const uint32_t *E, const char *name,
                               void (*f)(uint32_t, const uint32_t *, const uint32_t *, double *))
{
    bench_graph b;
    if (!bench_graph_open(&b, N, V, E))
        return;
    chain_step st = {f, NULL, NULL, greedy_setup(N, V[N]), greedy_internal};
    printf("%s: ", name);
    fflush(stdout);
    first_pass r = {0, 0.0};
    run_chain(&b, &st, 1, keep_first, &r);
    printf("%d %.4lf\n", r.colors, r.seconds);
    greedy_cleanup(st.handle);
    bench_graph_close(&b);
}

/* -------------------------------------------------------------- full table */

/* Cumulative seconds per (pass, column, trial) and the colour count per
 * pass, which every column and trial must reproduce. */
typedef struct {
    int passes, columns, trials;
    double *sec;
    int *colors; /* -1 until the first validated colouring of the pass */
    int column, trial;
} time_table;

static double *cell(time_table *t, int pass, int column, int trial)
{
    return &t->sec[((size_t)pass * (size_t)t->columns + (size_t)column) * (size_t)t->trials +
                   (size_t)trial];
}

static void record_pass(void *ctx, int i, int colors, double seconds)
{
    time_table *t = (time_table *)ctx;
    *cell(t, i, t->column, t->trial) = seconds;
    if (colors == 0)
        printf("Invalid coloring\n");
    else if (t->colors[i] < 0)
        t->colors[i] = colors;
    else if (t->colors[i] != colors)
        printf("Unexpected number of colors (%d %d)\n", t->colors[i], colors);
}

void runner_test_full(uint32_t N, const uint32_t *V, const uint32_t *E,
                      void *(*setup)(uint32_t, uint32_t), void (*cleanup)(void *),
                      void (*f_seq)(uint32_t, const uint32_t *, const uint32_t *, void *, void *,
                                    double *, int *),
                      void (*f_par)(uint32_t, const uint32_t *, const uint32_t *, void *, void *,
                                    double *, int *),
                      const char *s, int it, int k, int t_tot, int ntc, const int *nt)
{
    (void)t_tot; /* thread counts have no meaning on the GPU path */
    (void)nt;
    time_table t = {k < 1 ? 1 : k, (ntc < 0 ? 0 : ntc) + 1, it < 1 ? 1 : it, NULL, NULL, 0, 0};
    bench_graph b;
    if (!bench_graph_open(&b, N, V, E))
        return;
    t.sec = malloc(sizeof(double) * (size_t)t.passes * (size_t)t.columns * (size_t)t.trials);
    t.colors = malloc(sizeof(int) * (size_t)t.passes);
    if (!t.sec || !t.colors) {
        fprintf(stderr, "runner: out of host memory for the timing table\n");
        free(t.sec);
        free(t.colors);
        bench_graph_close(&b);
        return;
    }
    for (int i = 0; i < t.passes; i++)
        t.colors[i] = -1;
    void *state = setup(N, V[N]);
    chain_step seq = {NULL, f_seq, state, greedy_setup(N, V[N]), greedy_internal};
    chain_step par = {NULL, f_par, state, jones_plassmann_setup(N, V[N]), jones_plassmann_internal};
    for (t.trial = 0; t.trial < t.trials; t.trial++)
        for (t.column = 0; t.column < t.columns; t.column++)
            run_chain(&b, t.column == 0 ? &seq : &par, t.passes, record_pass, &t);
    for (int i = 0; i < t.passes; i++) {
        printf("%s_%d %d ", s, i, t.colors[i]);
        for (int c = 0; c < t.columns; c++) {
            double best = 9999.9; /* printed when nothing faster was measured */
            for (int r = 0; r < t.trials; r++)
                if (*cell(&t, i, c, r) < best)
                    best = *cell(&t, i, c, r);
            printf("%.6lf ", best);
        }
        printf("\n");
    }
    cleanup(state);
    greedy_cleanup(seq.handle);
    jones_plassmann_cleanup(par.handle);
    free(t.sec);
    free(t.colors);
    bench_graph_close(&b);
}
