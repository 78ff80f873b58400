"""Incidence-degree (ID) and saturation-degree (SD, SDA) orderings in pure
Python, restating incidence_degree.c:1-360 step for step.

TEST INFRASTRUCTURE ONLY: the checker that the library's orderings are
compared with. Pure-Python loops, meant for graphs of up to a few thousand
vertices.

The heap is the reference's own: 1-based slots, a slot table per vertex
(0 = absent), insertion that sifts up while strictly above the parent
(incidence_degree.c:17-37), sift-down to the larger child, left first
(39-60), and removal from any slot that moves the last entry there and only
sifts down (62-76). The order the walk produces depends on that last rule.

Pinning: golden vectors written by the compiled reference
(tests/golden/tiny_*.npz, rmat10.npz, seq_*.npz) that include graphs on
which a textbook priority queue gives another order; hand-derived known
answers (star, clique, path, isolated vertices); the compiled reference on
random small graphs when oracle/_ref is built.
"""
from __future__ import annotations

import numpy as np


class _Heap:
    def __init__(self, n, rank):
        self.slot = [0] * (n + 1)
        self.at = [0] * n
        self.n = 0
        self.rank = rank  # rank(x) -> tuple; larger ranks higher

    def _gt(self, a, b):
        return self.rank(a) > self.rank(b)

    def insert(self, x):  # incidence_degree.c:17-37
        self.n += 1
        i = self.n
        self.at[x] = i
        self.slot[i] = x
        par = i >> 1
        while i != 1 and self._gt(x, self.slot[par]):
            y = self.slot[par]
            self.at[y] = i
            self.slot[i] = y
            i = par
            self.at[x] = i
            self.slot[i] = x
            par = i >> 1

    def down(self, i):  # incidence_degree.c:39-60
        while True:
            left, right, big = 2 * i, 2 * i + 1, i
            if left <= self.n and self._gt(self.slot[left], self.slot[big]):
                big = left
            if right <= self.n and self._gt(self.slot[right], self.slot[big]):
                big = right
            if big == i:
                return
            x, y = self.slot[i], self.slot[big]
            self.at[y] = i
            self.at[x] = big
            self.slot[i] = y
            self.slot[big] = x
            i = big

    def remove(self, i):  # incidence_degree.c:62-76
        if i == 0:
            return
        x, y = self.slot[i], self.slot[self.n]
        self.slot[i] = y
        self.at[y] = i
        self.at[x] = 0
        self.n -= 1
        self.down(i)


def _components(V, E, N):
    """incidence_degree.c:78-105 (a DFS labelling; the labels only group
    vertices, so an explicit stack replaces the recursion)."""
    cc = [-1] * N
    t = 0
    for s in range(N):
        if cc[s] >= 0:
            continue
        cc[s] = t
        stack = [s]
        while stack:
            u = stack.pop()
            for v in E[V[u]:V[u + 1]]:
                if cc[v] < 0:
                    cc[v] = t
                    stack.append(v)
        t += 1
    return cc, t


def _walk(V, E, kind):
    V = [int(x) for x in V]
    E = [int(x) for x in E]
    N = len(V) - 1
    p = [-1.0] * N
    if N == 0:
        return np.zeros(0)
    deg = [V[u + 1] - V[u] for u in range(N)]
    cnt = [0] * N
    h = _Heap(N, lambda x: (cnt[x], deg[x], x))  # compare, incidence_degree.c:7-14
    cc, ncc = _components(V, E, N)
    colors = [0] * N

    def colour(u):  # incidence_degree.c:214-221 (mex, without the bound quirk)
        used = {colors[v] for v in E[V[u]:V[u + 1]]}
        c = 1
        while c in used:
            c += 1
        colors[u] = c

    def saturate(u):  # incidence_degree.c:223-243
        for v in E[V[u]:V[u + 1]]:
            if p[v] >= 0.0:
                continue
            found = any(w != u and colors[w] == colors[u] for w in E[V[v]:V[v + 1]])
            if not found:
                cnt[v] += 1
                h.remove(h.at[v])
                h.insert(v)

    if kind == "SDA":  # incidence_degree.c:259-360
        md = [0] * ncc
        for u in range(N):
            md[cc[u]] = max(md[cc[u]], deg[u])
        for u in range(N):
            if deg[u] == md[cc[u]]:
                h.insert(u)
        t = N
        while h.n:
            top = h.slot[1]
            best = (cnt[top], deg[top])
            batch = []
            while h.n and (cnt[h.slot[1]], deg[h.slot[1]]) == best:
                u = h.slot[1]
                h.remove(1)
                p[u] = float(t)
                batch.append(u)
            t -= 1
            for u in batch:
                colour(u)
                saturate(u)
        return np.array(p)

    md = [-1] * ncc  # incidence_degree.c:124-131, 189-199
    for u in range(N):
        if md[cc[u]] < 0 or deg[md[cc[u]]] < deg[u]:
            md[cc[u]] = u
    t = N
    h.insert(md[cc[0]])
    for k in range(N):  # incidence_degree.c:136-164, 205-250
        while h.n:
            u = h.slot[1]
            h.remove(1)
            p[u] = float(t)
            t -= 1
            if kind == "ID":
                for v in E[V[u]:V[u + 1]]:
                    if p[v] >= 0.0:
                        continue
                    cnt[v] += 1
                    h.remove(h.at[v])
                    h.insert(v)
            else:
                colour(u)
                saturate(u)
        if p[k] < 0.0:
            h.insert(md[cc[k]])
    return np.array(p)


def incidence_degree(V, E):
    """incidence_degree.c:107-170"""
    return _walk(V, E, "ID")


def saturation_degree(V, E):
    """incidence_degree.c:172-257"""
    return _walk(V, E, "SD")


def saturation_degree_alt(V, E):
    """incidence_degree.c:259-360"""
    return _walk(V, E, "SDA")


def textbook_incidence_degree(V, E):
    """ID with a correct priority queue (always the largest key next): the
    plausible restatement that golden vectors must tell apart."""
    V = [int(x) for x in V]
    E = [int(x) for x in E]
    N = len(V) - 1
    deg = [V[u + 1] - V[u] for u in range(N)]
    cc, ncc = _components(V, E, N)
    md = [-1] * ncc
    for u in range(N):
        if md[cc[u]] < 0 or deg[md[cc[u]]] < deg[u]:
            md[cc[u]] = u
    cnt = [0] * N
    p = [-1.0] * N
    t = N
    for k in range(N - 1):
        if p[k] >= 0.0:
            continue
        live = {md[cc[k]]}
        while live:
            u = max(live, key=lambda x: (cnt[x], deg[x], x))
            live.discard(u)
            p[u] = float(t)
            t -= 1
            for v in E[V[u]:V[u + 1]]:
                if p[v] < 0.0:
                    cnt[v] += 1
                    live.add(v)
    return np.array(p)
