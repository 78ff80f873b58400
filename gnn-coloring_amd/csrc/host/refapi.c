/*
 * refapi.c -- the reference's C API (coloring.h, fast_gnn.h, gnn.h,
 * first_fit.h, largest_degree_first.h, smallest_degree_last.h) on top of the
 * device-resident gcx API. Host arrays are copied to HBM on every call, the
 * work runs in the HIP kernels, and results are copied back.
 *
 * Like the reference these functions return no status: a GPU failure prints
 * "gcx: <step> failed: <reason>" and exits with status 1 (the reference exits
 * the same way on a missing model file, gnn.c:27-32).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "coloring.h"
#include "fast_gnn.h"
#include "first_fit.h"
#include "gnn.h"
#include "internal.h"
#include "largest_degree_first.h"
#include "runner.h"
#include "smallest_degree_last.h"

static void die(const char *what, int rc)
{
    fprintf(stderr, "gcx: %s failed: %s\n", what, gcx_status_string(rc));
    exit(1);
}

#define CHECK(what, expr)                                                       \
    do {                                                                        \
        int rc_ = (expr);                                                       \
        if (rc_ != GCX_OK)                                                      \
            die(what, rc_);                                                     \
    } while (0)

/* A graph buffer plus a workspace, sized once by a *_setup call. */
typedef struct {
    gcx_graph *g;
    gcx_workspace *ws;
    uint32_t capN;
    uint64_t capM;
} ref_ctx;

static ref_ctx *ctx_create(uint32_t N, uint64_t M)
{
    ref_ctx *c = calloc(1, sizeof(*c));
    if (!c)
        die("allocation", GCX_ERR_NOMEM);
    c->capN = N ? N : 1;
    c->capM = M;
    CHECK("graph allocation", gcx__graph_alloc(c->capN, c->capM, NULL, &c->g));
    CHECK("workspace allocation", gcx_workspace_create(c->capN, c->capM, &c->ws));
    return c;
}

static void ctx_destroy(ref_ctx *c)
{
    if (!c)
        return;
    gcx_graph_free(c->g);
    gcx_workspace_free(c->ws);
    free(c);
}

static void ctx_load(ref_ctx *c, uint32_t N, const uint32_t *V, const uint32_t *E)
{
    if (N > c->capN || (uint64_t)V[N] > c->capM)
        die("graph upload", GCX_ERR_CAPACITY);
    CHECK("graph upload", gcx__graph_fill(c->g, N, V, E, NULL));
}

/* ------------------------------------------------------------ coloring.h */

void *jones_plassmann_setup(uint32_t N, uint32_t M) { return ctx_create(N, M); }

void jones_plassmann_cleanup(void *jp) { ctx_destroy((ref_ctx *)jp); }

void jones_plassmann_internal(uint32_t N, const uint32_t *V, const uint32_t *E, void *jp,
                              const double *p, int *color)
{
    if (N == 0)
        return;
    ref_ctx *c = (ref_ctx *)jp;
    ctx_load(c, N, V, E);
    CHECK("priority upload", gcx_set_priorities(c->ws, c->g, p, NULL));
    CHECK("colouring", gcx_color(c->ws, c->g, NULL));
    CHECK("colour download", gcx_download_colors(c->ws, N, color, NULL));
}

int *jones_plassmann_coloring(uint32_t N, const uint32_t *V, const uint32_t *E, const double *p)
{
    int *color = malloc(sizeof(int) * (N ? N : 1));
    void *jp = jones_plassmann_setup(N, V[N]);
    jones_plassmann_internal(N, V, E, jp, p, color);
    jones_plassmann_cleanup(jp);
    return color;
}

/* The sequential names give the same colours (coloring.c:60-125). */
void *greedy_setup(uint32_t N, uint32_t M) { return ctx_create(N, M); }
void greedy_cleanup(void *gd) { ctx_destroy((ref_ctx *)gd); }

void greedy_internal(uint32_t N, const uint32_t *V, const uint32_t *E, void *gd, const double *p,
                     int *color)
{
    jones_plassmann_internal(N, V, E, gd, p, color);
}

int *greedy_coloring(uint32_t N, const uint32_t *V, const uint32_t *E, const double *p)
{
    return jones_plassmann_coloring(N, V, E, p);
}

int validate_coloring_par(uint32_t N, const uint32_t *V, const uint32_t *E, const int *color)
{
    if (N == 0)
        return 1;
    ref_ctx *c = ctx_create(N, V[N]);
    ctx_load(c, N, V, E);
    CHECK("colour upload", gcx_upload_colors(c->ws, N, color, NULL));
    int C = 0;
    CHECK("validation", gcx_validate(c->ws, c->g, NULL, &C));
    ctx_destroy(c);
    return C;
}

int validate_coloring(uint32_t N, const uint32_t *V, const uint32_t *E, const int *color)
{
    return validate_coloring_par(N, V, E, color);
}

/* ------------------------------------------------------------- orderings */

static void order_into(ref_ctx *c, int heuristic, const gcx_model *m, uint32_t N, double *p)
{
    CHECK("ordering", gcx_order(c->ws, c->g, heuristic, m, GCX_FP32, NULL));
    if (p)
        CHECK("priority download", gcx_download_priorities(c->ws, N, p, NULL));
}

static void ordering_once(int heuristic, const gcx_model *m, uint32_t N, const uint32_t *V,
                          const uint32_t *E, double *p)
{
    if (N == 0)
        return;
    ref_ctx *c = ctx_create(N, V ? V[N] : 0);
    if (heuristic == GCX_FF && !V) {
        /* first_fit.c:8-19 ignores the graph; keys need only N. */
        CHECK("ordering", gcx__order_ff_only(c->ws, N, NULL));
        CHECK("priority download", gcx_download_priorities(c->ws, N, p, NULL));
    } else {
        ctx_load(c, N, V, E);
        order_into(c, heuristic, m, N, p);
    }
    ctx_destroy(c);
}

static int *coloring_once(int heuristic, const gcx_model *m, uint32_t N, const uint32_t *V,
                          const uint32_t *E)
{
    int *color = malloc(sizeof(int) * (N ? N : 1));
    if (N == 0)
        return color;
    ref_ctx *c = ctx_create(N, V[N]);
    ctx_load(c, N, V, E);
    order_into(c, heuristic, m, N, NULL);
    CHECK("colouring", gcx_color(c->ws, c->g, NULL));
    CHECK("colour download", gcx_download_colors(c->ws, N, color, NULL));
    ctx_destroy(c);
    return color;
}

/* Runner plug-in body: ordering + colouring on the handle's graph; p and
 * colour are written back for the Culberson chain that follows. */
static void heuristic_run(int heuristic, const gcx_model *m, uint32_t N, const uint32_t *V,
                          const uint32_t *E, void *jp, double *p, int *color)
{
    if (N == 0)
        return;
    ref_ctx *c = (ref_ctx *)jp;
    ctx_load(c, N, V, E);
    order_into(c, heuristic, m, N, p);
    CHECK("colouring", gcx_color(c->ws, c->g, NULL));
    CHECK("colour download", gcx_download_colors(c->ws, N, color, NULL));
}

static void *no_setup(uint32_t N, uint32_t M)
{
    (void)N;
    (void)M;
    return NULL;
}

static void no_cleanup(void *p) { (void)p; }

void first_fit_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p)
{
    ordering_once(GCX_FF, NULL, N, V, E, p);
}

void first_fit_ordering_par(uint32_t N, const uint32_t *V, const uint32_t *E, double *p)
{
    ordering_once(GCX_FF, NULL, N, V, E, p);
}

int *first_fit_coloring(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    return coloring_once(GCX_FF, NULL, N, V, E);
}

int *first_fit_coloring_par(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    return coloring_once(GCX_FF, NULL, N, V, E);
}

static void ff_run(uint32_t N, const uint32_t *V, const uint32_t *E, void *fd, void *jp, double *p,
                   int *color)
{
    (void)fd;
    heuristic_run(GCX_FF, NULL, N, V, E, jp, p, color);
}

void first_fit_runner(uint32_t N, const uint32_t *V, const uint32_t *E, int it, int k, int t_tot,
                      int ntc, const int *nt)
{
    runner_test_full(N, V, E, no_setup, no_cleanup, ff_run, ff_run, "FF", it, k, t_tot, ntc, nt);
}

void largest_degree_first_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p)
{
    ordering_once(GCX_LF, NULL, N, V, E, p);
}

int *largest_degree_first_coloring(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    return coloring_once(GCX_LF, NULL, N, V, E);
}

int *largest_degree_first_coloring_par(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    return coloring_once(GCX_LF, NULL, N, V, E);
}

static void lf_run(uint32_t N, const uint32_t *V, const uint32_t *E, void *fd, void *jp, double *p,
                   int *color)
{
    (void)fd;
    heuristic_run(GCX_LF, NULL, N, V, E, jp, p, color);
}

void largest_degree_first_runner(uint32_t N, const uint32_t *V, const uint32_t *E, int it, int k,
                                 int t_tot, int ntc, const int *nt)
{
    runner_test_full(N, V, E, no_setup, no_cleanup, lf_run, lf_run, "LF", it, k, t_tot, ntc, nt);
}

void largest_log_degree_first_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p)
{
    ordering_once(GCX_LLF, NULL, N, V, E, p);
}

int *largest_log_degree_first_coloring(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    return coloring_once(GCX_LLF, NULL, N, V, E);
}

int *largest_log_degree_first_coloring_par(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    /* The reference passes a JP handle to greedy_internal here
     * (largest_degree_first.c:140-151); the colours it means are these. */
    return coloring_once(GCX_LLF, NULL, N, V, E);
}

static void llf_run(uint32_t N, const uint32_t *V, const uint32_t *E, void *fd, void *jp,
                    double *p, int *color)
{
    (void)fd;
    heuristic_run(GCX_LLF, NULL, N, V, E, jp, p, color);
}

void largest_log_degree_first_runner(uint32_t N, const uint32_t *V, const uint32_t *E, int it,
                                     int k, int t_tot, int ntc, const int *nt)
{
    runner_test_full(N, V, E, no_setup, no_cleanup, llf_run, llf_run, "LLF", it, k, t_tot, ntc, nt);
}

void smallest_degree_last_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p)
{
    ordering_once(GCX_SL, NULL, N, V, E, p);
}

int *smallest_degree_last_coloring(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    return coloring_once(GCX_SL, NULL, N, V, E);
}

int *smallest_degree_last_coloring_par(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    return coloring_once(GCX_SL, NULL, N, V, E);
}

static void sl_run(uint32_t N, const uint32_t *V, const uint32_t *E, void *fd, void *jp, double *p,
                   int *color)
{
    (void)fd;
    heuristic_run(GCX_SL, NULL, N, V, E, jp, p, color);
}

void smallest_degree_last_runner(uint32_t N, const uint32_t *V, const uint32_t *E, int it, int k,
                                 int t_tot, int ntc, const int *nt)
{
    runner_test_full(N, V, E, no_setup, no_cleanup, sl_run, sl_run, "SL", it, k, t_tot, ntc, nt);
}

void smallest_log_degree_last_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p)
{
    ordering_once(GCX_SLL, NULL, N, V, E, p);
}

int *smallest_log_degree_last_coloring(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    return coloring_once(GCX_SLL, NULL, N, V, E);
}

int *smallest_log_degree_last_coloring_par(uint32_t N, const uint32_t *V, const uint32_t *E)
{
    return coloring_once(GCX_SLL, NULL, N, V, E);
}

static void sll_run(uint32_t N, const uint32_t *V, const uint32_t *E, void *fd, void *jp,
                    double *p, int *color)
{
    (void)fd;
    heuristic_run(GCX_SLL, NULL, N, V, E, jp, p, color);
}

void smallest_log_degree_last_runner(uint32_t N, const uint32_t *V, const uint32_t *E, int it,
                                     int k, int t_tot, int ntc, const int *nt)
{
    runner_test_full(N, V, E, no_setup, no_cleanup, sll_run, sll_run, "SLL", it, k, t_tot, ntc,
                     nt);
}

/* ------------------------------------------------------------------- GNN */

void cgnn_graph_sage_16_par(int params, const float *param, uint32_t N, const uint32_t *V,
                            const uint32_t *E, int md, float *x, float *y, double *p)
{
    (void)x;
    (void)y;
    if (N == 0)
        return;
    gcx_model *m = NULL;
    CHECK("model upload", gcx_model_from_params(params, param, md, &m));
    ordering_once(GCX_GNN, m, N, V, E, p);
    gcx_model_free(m);
}

void cgnn_graph_sage_16(int params, const float *param, uint32_t N, const uint32_t *V,
                        const uint32_t *E, int md, float *x, float *y, double *p)
{
    cgnn_graph_sage_16_par(params, param, N, V, E, md, x, y, p);
}

/* gnn.c:18: the layer count used by the next setup (not reentrant there
 * either). */
static int g_layers = 2;

static gcx_model *load_layers(int l)
{
    char path[4096];
    const char *dir = getenv("GCX_MODEL_DIR");
    snprintf(path, sizeof(path), "%s/m%d.gnn", dir && *dir ? dir : "models", l);
    gcx_model *m = NULL;
    const int rc = gcx_model_load(path, &m);
    if (rc == GCX_ERR_IO) {
        fprintf(stderr, "Failed to open model file (%d layers)\n", l);
        exit(1);
    }
    CHECK("model load", rc);
    return m;
}

void gnn_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p, int l)
{
    g_layers = l;
    gcx_model *m = load_layers(l);
    ordering_once(GCX_GNN, m, N, V, E, p);
    gcx_model_free(m);
}

static void *gnn_setup_cb(uint32_t N, uint32_t M)
{
    (void)N;
    (void)M;
    return load_layers(g_layers);
}

static void gnn_cleanup_cb(void *m) { gcx_model_free((gcx_model *)m); }

static void gnn_run(uint32_t N, const uint32_t *V, const uint32_t *E, void *fd, void *jp,
                    double *p, int *color)
{
    heuristic_run(GCX_GNN, (const gcx_model *)fd, N, V, E, jp, p, color);
}

void gnn_runner(uint32_t N, const uint32_t *V, const uint32_t *E, int l, int it, int k, int t_tot,
                int ntc, const int *nt)
{
    char name[64];
    g_layers = l;
    snprintf(name, sizeof(name), "GNN%d", l);
    runner_test_full(N, V, E, gnn_setup_cb, gnn_cleanup_cb, gnn_run, gnn_run, name, it, k, t_tot,
                     ntc, nt);
}
