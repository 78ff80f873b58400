/*
 * cli.c -- `coloring`, a command-line driver compatible with the reference's
 * ./coloring (main.c:17-30 usage, 47-214 flow) over the GPU library.
 *
 *   coloring -g graph.mtx [-v] [-q] [-f] [-c K] [-t T] tc1 .. tcn
 *
 * -q prints the colour counts "FF LF LLF SL SLL ID SDA SD GNN2 GNN3 GNN4";
 * under -f the LLF, SLL, ID, SDA and SD columns print 0, as in the
 * reference. The ID, SDA and SD orderings are sequential walks on the host
 * (incidence_degree.h); their priorities are coloured on the GPU like the
 * others.
 */
#include <stdio.h>
#include <stdlib.h>
#include <time.h>
#include <unistd.h>

#include "coloring.h"
#include "fast_io.h"
#include "first_fit.h"
#include "gnn.h"
#include "incidence_degree.h"
#include "largest_degree_first.h"
#include "smallest_degree_last.h"

static const char *usage =
    "GNN-ordered greedy graph colouring on the GPU\n"
    "\nUsage: coloring -g <graph.mtx> [options] tc1 tc2 ... tcn\n"
    "\nWithout -q, one line per heuristic:\n"
    "  heuristic_#culberson #colours seconds_sequential seconds_tc1 ... seconds_tcn\n"
    "\nOptions:\n"
    "  -h      this message\n"
    "  -v      verbose\n"
    "  -g path input graph (Matrix Market), required\n"
    "  -q      only print the colours used by each heuristic\n"
    "  -f      only FF, LF, SL and GNN (skip the log variants)\n"
    "  -c K    Culberson recolourings per heuristic\n"
    "  -t T    accepted for compatibility\n";

static double now_s(void)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* main.c:32-45: colour, then K-1 Culberson passes, then validate. */
static int chain(uint32_t N, const uint32_t *V, const uint32_t *E, void *jp, double *p, int *color,
                 int K)
{
    jones_plassmann_internal(N, V, E, jp, p, color);
    for (int i = 1; i < K; i++) {
        for (uint32_t j = 0; j < N; j++)
            p[j] = color[j];
        jones_plassmann_internal(N, V, E, jp, p, color);
    }
    return validate_coloring_par(N, V, E, color);
}

int main(int argc, char **argv)
{
    const char *path = NULL;
    int quick = 0, fast = 0, culb = 1, threads = 1, verbose = 0, opt;
    opterr = 0;
    while ((opt = getopt(argc, argv, "hvfqg:c:t:")) != -1) {
        switch (opt) {
        case 'g': path = optarg; break;
        case 'q': quick = 1; break;
        case 'f': fast = 1; break;
        case 'c': culb = atoi(optarg); break;
        case 't': threads = atoi(optarg); break;
        case 'v': verbose = 1; break;
        case 'h': printf("%s", usage); return 0;
        default:
            fprintf(stderr, "Unknown option or missing argument: -%c\n", optopt);
            return 1;
        }
    }
    const int ntc = argc - optind;
    int *nt = malloc(sizeof(int) * (ntc ? ntc : 1));
    for (int i = 0; i < ntc; i++) {
        nt[i] = atoi(argv[optind + i]);
        if (nt[i] > threads)
            threads = nt[i];
    }
    if (culb < 1)
        culb = 1;
    FILE *f = path ? fopen(path, "r") : NULL;
    if (!f) {
        printf("Unable to open file %s\n", path ? path : "");
        return 1;
    }
    uint32_t N, *V, *E;
    double t0 = now_s();
    mtx_parse_par(f, &N, &V, &E);
    fclose(f);
    if (verbose)
        printf("Parsing + GPU CSR build took: %.4lfs, |V|=%u, |E|=%u\n", now_s() - t0, N, V[N]);
    if (quick) {
        void *jp = jones_plassmann_setup(N, V[N]);
        double *p = malloc(sizeof(double) * (N + 8ull));
        int *color = malloc(sizeof(int) * (N + 1ull));
        int ff, lf, llf = 0, sl, sll = 0, id = 0, sda = 0, sd = 0, gnn2, gnn3, gnn4;
        first_fit_ordering_par(N, V, E, p);
        ff = chain(N, V, E, jp, p, color, culb);
        largest_degree_first_ordering(N, V, E, p);
        lf = chain(N, V, E, jp, p, color, culb);
        smallest_degree_last_ordering(N, V, E, p);
        sl = chain(N, V, E, jp, p, color, culb);
        gnn_ordering(N, V, E, p, 2);
        gnn2 = chain(N, V, E, jp, p, color, culb);
        gnn_ordering(N, V, E, p, 3);
        gnn3 = chain(N, V, E, jp, p, color, culb);
        gnn_ordering(N, V, E, p, 4);
        gnn4 = chain(N, V, E, jp, p, color, culb);
        if (!fast) {
            largest_log_degree_first_ordering(N, V, E, p);
            llf = chain(N, V, E, jp, p, color, culb);
            smallest_log_degree_last_ordering(N, V, E, p);
            sll = chain(N, V, E, jp, p, color, culb);
            incidence_degree_ordering(N, V, E, p);
            id = chain(N, V, E, jp, p, color, culb);
            saturation_degree_ordering_alt(N, V, E, p);
            sda = chain(N, V, E, jp, p, color, culb);
            saturation_degree_ordering(N, V, E, p);
            sd = chain(N, V, E, jp, p, color, culb);
        }
        if (verbose)
            printf("FF LF LLF SL SLL ID SDA SD GNN2 GNN3 GNN4\n");
        printf("%d %d %d %d %d %d %d %d %d %d %d\n", ff, lf, llf, sl, sll, id, sda, sd, gnn2,
               gnn3, gnn4);
        jones_plassmann_cleanup(jp);
        free(p);
        free(color);
    } else {
        if (verbose) {
            printf("name colors seq ");
            for (int i = 0; i < ntc; i++)
                printf("%d ", nt[i]);
            printf("\n");
        }
        first_fit_runner(N, V, E, 5, culb, threads, ntc, nt);
        largest_degree_first_runner(N, V, E, 5, culb, threads, ntc, nt);
        if (!fast)
            largest_log_degree_first_runner(N, V, E, 5, culb, threads, ntc, nt);
        smallest_degree_last_runner(N, V, E, 5, culb, threads, ntc, nt);
        if (!fast)
            smallest_log_degree_last_runner(N, V, E, 5, culb, threads, ntc, nt);
        gnn_runner(N, V, E, 2, 5, culb, threads, ntc, nt);
        gnn_runner(N, V, E, 3, 5, culb, threads, ntc, nt);
        gnn_runner(N, V, E, 4, 5, culb, threads, ntc, nt);
    }
    free(V);
    free(E);
    free(nt);
    return 0;
}
