"""Independent brute-force definitions for tiny graphs (pure Python).

Shortest-path counts by explicit BFS layering and pair dependencies by the
textbook definition delta_s(v) = sum_t sigma_st(v) / sigma_st with
sigma_st(v) = sigma_sv * sigma_vt when d(s,v) + d(v,t) = d(s,t).  Nothing
here follows the oracle's code path (no dependency recursion).
"""
from __future__ import annotations

from collections import deque


def adjacency(N, col, row):
    return [list(row[col[u]:col[u + 1]]) for u in range(N)]


def bfs(adj, s):
    """dist (-1 unreachable) and exact path counts (Python ints)."""
    n = len(adj)
    dist = [-1] * n
    cnt = [0] * n
    dist[s] = 0
    cnt[s] = 1
    q = deque([s])
    while q:
        u = q.popleft()
        for w in adj[u]:
            if dist[w] == -1:
                dist[w] = dist[u] + 1
                q.append(w)
            if dist[w] == dist[u] + 1:
                cnt[w] += cnt[u]
    return dist, cnt


def dependencies(adj, s):
    """delta_s(v) for all v != s by the pair definition (exact counts)."""
    n = len(adj)
    ds, cs = bfs(adj, s)
    per = [bfs(adj, v) for v in range(n)]
    out = [0.0] * n
    for v in range(n):
        if v == s or ds[v] < 0:
            continue
        dv, cv = per[v]
        tot = 0.0
        for t in range(n):
            if t in (s, v) or ds[t] < 0 or dv[t] < 0:
                continue
            if ds[v] + dv[t] == ds[t]:
                tot += cs[v] * cv[t] / cs[t]
        out[v] = tot
    return out
