/*
 * seqorder.c -- incidence-degree (ID) and saturation-degree (SD, SDA)
 * orderings, reference incidence_degree.c:107-360.
 *
 * All three are one sequential walk. Each step takes the vertex at the top of
 * a binary max-heap keyed by (count, degree, id), gives it the next priority
 * N, N-1, ..., and raises the count of its unordered neighbours: for ID the
 * count is the number of ordered neighbours; for SD and SDA it is the number
 * of distinct colours among the coloured neighbours, each picked vertex
 * taking the smallest colour its coloured neighbours leave free. Components
 * are walked one after another in the order of their smallest vertex, each
 * from its largest-degree vertex (the smallest id among equal degrees); SDA
 * instead starts with every vertex of maximum degree in its component on the
 * heap. ID and SD leave an isolated last vertex (N > 1) at p = -1, as the
 * reference does. These orderings run on the host: no parallel walk has the same
 * output, and they are not part of the GPU ordering + colouring path.
 *
 * The heap reproduces the reference's mechanics, because the order depends
 * on them: taking an entry out of an interior slot moves the last entry into
 * that slot and only sifts it down (incidence_degree.c:62-76). The heap can
 * then hold an entry above a smaller parent, and a later pop need not return
 * the largest key; a textbook priority queue gives another order on about
 * one random small graph in ten. Insertion appends and sifts up while the
 * entry ranks strictly above its parent (17-37); sift-down moves an entry to
 * its larger child, left checked first (39-60).
 *
 * Heap entries carry their key. That is exact: a vertex's count only changes
 * immediately before it is taken out of the heap and put back, so every entry
 * in the heap always holds its vertex's current key.
 *
 * SD's test "does v already have a coloured neighbour other than u with u's
 * colour" (incidence_degree.c:228-236, a scan of v's row) is answered by a
 * hash set of (v, colour) pairs, filled as each neighbour of v is coloured.
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "incidence_degree.h"

typedef struct {
    uint64_t key; /* count << 32 | degree */
    uint32_t v;
} hent;

typedef struct {
    uint32_t n;   /* entries in slots 1..n */
    hent *slot;   /* N + 1 slots */
    uint32_t *at; /* vertex -> slot, 0 when absent */
} vheap;

static void *xalloc(size_t n, size_t sz, int zero)
{
    void *p = zero ? calloc(n ? n : 1, sz) : malloc((n ? n : 1) * sz);
    if (!p) {
        fprintf(stderr, "seqorder: out of host memory (%zu x %zu bytes)\n", n, sz);
        exit(1);
    }
    return p;
}

/* a ranks strictly above b: larger count, then larger degree, then larger id */
static inline int above(hent a, hent b)
{
    return a.key != b.key ? a.key > b.key : a.v > b.v;
}

static inline void place(vheap *h, uint32_t s, hent e)
{
    h->slot[s] = e;
    h->at[e.v] = s;
}

static void push(vheap *h, hent e)
{
    uint32_t s = ++h->n;
    while (s > 1 && above(e, h->slot[s >> 1])) {
        place(h, s, h->slot[s >> 1]);
        s >>= 1;
    }
    place(h, s, e);
}

static void sink(vheap *h, uint32_t s, hent e)
{
    for (;;) {
        const uint32_t l = 2u * s, r = l + 1u;
        uint32_t best = s;
        hent b = e;
        if (l <= h->n && above(h->slot[l], b)) {
            best = l;
            b = h->slot[l];
        }
        if (r <= h->n && above(h->slot[r], b)) {
            best = r;
            b = h->slot[r];
        }
        if (best == s)
            break;
        place(h, s, b);
        s = best;
    }
    place(h, s, e);
}

/* Take out the entry in slot s (no-op for s = 0): the last entry fills the
 * hole and is only sifted down. */
static void take(vheap *h, uint32_t s)
{
    if (!s)
        return;
    h->at[h->slot[s].v] = 0;
    const hent last = h->slot[h->n--];
    if (s <= h->n)
        sink(h, s, last);
}

static uint32_t pop(vheap *h)
{
    const uint32_t u = h->slot[1].v;
    take(h, 1);
    return u;
}

/* ------------------------------------------------------------ (v, colour) set */

typedef struct {
    uint64_t *t; /* 0 = empty; a pair is never 0 because colours start at 1 */
    uint64_t mask, used;
} pairset;

static inline uint64_t mix(uint64_t z)
{
    z ^= z >> 33;
    z *= 0xff51afd7ed558ccdull;
    z ^= z >> 33;
    z *= 0xc4ceb9fe1a85ec53ull;
    return z ^ (z >> 33);
}

static void ps_init(pairset *s, uint64_t expect)
{
    uint64_t cap = 64;
    while (cap < 2 * expect)
        cap <<= 1;
    s->t = (uint64_t *)xalloc(cap, sizeof(uint64_t), 1);
    s->mask = cap - 1;
    s->used = 0;
}

static void ps_grow(pairset *s)
{
    const uint64_t old = s->mask + 1;
    uint64_t *t = s->t;
    s->t = (uint64_t *)xalloc(2 * old, sizeof(uint64_t), 1);
    s->mask = 2 * old - 1;
    for (uint64_t i = 0; i < old; i++) {
        if (!t[i])
            continue;
        uint64_t j = mix(t[i]) & s->mask;
        while (s->t[j])
            j = (j + 1) & s->mask;
        s->t[j] = t[i];
    }
    free(t);
}

/* Adds (v, c); returns 1 when it was already present. */
static int ps_add(pairset *s, uint32_t v, uint32_t c)
{
    const uint64_t k = (uint64_t)v << 32 | c;
    uint64_t j = mix(k) & s->mask;
    while (s->t[j]) {
        if (s->t[j] == k)
            return 1;
        j = (j + 1) & s->mask;
    }
    s->t[j] = k;
    if (++s->used * 2 > s->mask + 1)
        ps_grow(s);
    return 0;
}

/* ------------------------------------------------------------------- the walk */

enum { WALK_ID, WALK_SD, WALK_SDA };

typedef struct {
    uint32_t N;
    const uint32_t *V, *E;
    double *p;
    vheap h;
    uint32_t *cnt;  /* ordered neighbours (ID) or saturation (SD, SDA) */
    uint8_t *done;  /* priority assigned */
    uint32_t *col;  /* SD, SDA: colour given at pick time */
    uint32_t *seen; /* SD, SDA: mex stamps by colour */
    pairset ps;
} walk;

static inline uint32_t deg(const walk *w, uint32_t u) { return w->V[u + 1] - w->V[u]; }

static inline hent entry(const walk *w, uint32_t v)
{
    hent e = {(uint64_t)w->cnt[v] << 32 | deg(w, v), v};
    return e;
}

/* Raise v's count and move it to its new place (or onto the heap). */
static inline void bump(walk *w, uint32_t v)
{
    w->cnt[v]++;
    take(&w->h, w->h.at[v]);
    push(&w->h, entry(w, v));
}

static void id_pick(walk *w, uint32_t u)
{
    for (uint32_t i = w->V[u]; i < w->V[u + 1]; i++) {
        const uint32_t v = w->E[i];
        if (!w->done[v])
            bump(w, v);
    }
}

/* Colour u with the smallest colour its coloured neighbours leave free, then
 * raise the saturation of each unordered neighbour that had no neighbour of
 * that colour yet. */
static void sd_pick(walk *w, uint32_t u)
{
    const uint32_t a = w->V[u], b = w->V[u + 1], stamp = u + 1u;
    for (uint32_t i = a; i < b; i++) {
        const uint32_t c = w->col[w->E[i]];
        if (c && c <= b - a + 1u)
            w->seen[c] = stamp;
    }
    uint32_t c = 1;
    while (w->seen[c] == stamp)
        c++;
    w->col[u] = c;
    for (uint32_t i = a; i < b; i++) {
        const uint32_t v = w->E[i];
        if (!w->done[v] && !ps_add(&w->ps, v, c))
            bump(w, v);
    }
}

/* Component labels in the order of each component's smallest vertex; per
 * component its first vertex of maximum degree (start) and that degree. */
static uint32_t components(const walk *w, uint32_t *lab, uint32_t **start_out,
                           uint32_t **maxdeg_out)
{
    const uint32_t N = w->N;
    uint32_t *queue = (uint32_t *)xalloc(N, sizeof(uint32_t), 0);
    uint32_t *start = (uint32_t *)xalloc(N, sizeof(uint32_t), 0);
    uint32_t *maxdeg = (uint32_t *)xalloc(N, sizeof(uint32_t), 0);
    for (uint32_t u = 0; u < N; u++)
        lab[u] = UINT32_MAX;
    uint32_t nc = 0;
    for (uint32_t k = 0; k < N; k++) {
        if (lab[k] != UINT32_MAX)
            continue;
        uint32_t head = 0, tail = 0, best = k;
        lab[k] = nc;
        queue[tail++] = k;
        while (head < tail) {
            const uint32_t u = queue[head++];
            const uint32_t du = deg(w, u), db = deg(w, best);
            if (du > db || (du == db && u < best))
                best = u;
            for (uint32_t i = w->V[u]; i < w->V[u + 1]; i++) {
                const uint32_t v = w->E[i];
                if (lab[v] == UINT32_MAX) {
                    lab[v] = nc;
                    queue[tail++] = v;
                }
            }
        }
        start[nc] = best;
        maxdeg[nc] = deg(w, best);
        nc++;
    }
    free(queue);
    *start_out = start;
    *maxdeg_out = maxdeg;
    return nc;
}

static void run(int kind, uint32_t N, const uint32_t *V, const uint32_t *E, double *p)
{
    if (N == 0)
        return;
    walk w;
    memset(&w, 0, sizeof(w));
    w.N = N;
    w.V = V;
    w.E = E;
    w.p = p;
    w.h.slot = (hent *)xalloc((size_t)N + 1, sizeof(hent), 0);
    w.h.at = (uint32_t *)xalloc(N, sizeof(uint32_t), 1);
    w.cnt = (uint32_t *)xalloc(N, sizeof(uint32_t), 1);
    w.done = (uint8_t *)xalloc(N, 1, 1);
    uint32_t maxd = 0;
    for (uint32_t u = 0; u < N; u++)
        if (V[u + 1] - V[u] > maxd)
            maxd = V[u + 1] - V[u];
    if (kind != WALK_ID) {
        w.col = (uint32_t *)xalloc(N, sizeof(uint32_t), 1);
        w.seen = (uint32_t *)xalloc((size_t)maxd + 2, sizeof(uint32_t), 1);
        ps_init(&w.ps, V[N] / 2);
    }
    uint32_t *lab = (uint32_t *)xalloc(N, sizeof(uint32_t), 0);
    uint32_t *start, *maxdeg;
    components(&w, lab, &start, &maxdeg);

    double t = (double)N;
    if (kind == WALK_SDA) {
        uint32_t *batch = (uint32_t *)xalloc(N, sizeof(uint32_t), 0);
        for (uint32_t u = 0; u < N; u++)
            if (V[u + 1] - V[u] == maxdeg[lab[u]])
                push(&w.h, entry(&w, u));
        while (w.h.n) {
            const uint64_t top = w.h.slot[1].key;
            uint32_t nb = 0;
            while (w.h.n && w.h.slot[1].key == top) {
                const uint32_t u = pop(&w.h);
                p[u] = t;
                w.done[u] = 1;
                batch[nb++] = u;
            }
            t -= 1.0;
            for (uint32_t i = 0; i < nb; i++)
                sd_pick(&w, batch[i]);
        }
        free(batch);
    } else {
        for (uint32_t k = 0; k < N; k++) {
            if (w.done[k])
                continue;
            if (k > 0 && k == N - 1) {
                /* the reference puts vertex 0's component on the heap first
                 * and every later one after draining the previous one
                 * (incidence_degree.c:133-164), so a component that starts
                 * at the last vertex, i.e. an isolated last vertex other
                 * than vertex 0, is never walked and keeps p = -1 */
                p[k] = -1.0;
                break;
            }
            push(&w.h, entry(&w, start[lab[k]]));
            while (w.h.n) {
                const uint32_t u = pop(&w.h);
                p[u] = t;
                t -= 1.0;
                w.done[u] = 1;
                if (kind == WALK_ID)
                    id_pick(&w, u);
                else
                    sd_pick(&w, u);
            }
        }
    }
    free(lab);
    free(start);
    free(maxdeg);
    free(w.h.slot);
    free(w.h.at);
    free(w.cnt);
    free(w.done);
    free(w.col);
    free(w.seen);
    free(w.ps.t);
}

void incidence_degree_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p)
{
    run(WALK_ID, N, V, E, p);
}

void saturation_degree_ordering(uint32_t N, const uint32_t *V, const uint32_t *E, double *p)
{
    run(WALK_SD, N, V, E, p);
}

void saturation_degree_ordering_alt(uint32_t N, const uint32_t *V, const uint32_t *E, double *p)
{
    run(WALK_SDA, N, V, E, p);
}
