"""Small helpers shared by the tests: brute force, closed-form graphs."""
import itertools
import random


def brute_force(n, adj, k, q):
    """All maximal k-plexes of size >= q of a tiny graph (adj: list of sets)."""
    nbm = [sum(1 << v for v in adj[u]) for u in range(n)]

    def is_kplex(mask):
        s = bin(mask).count("1")
        m = mask
        while m:
            v = (m & -m).bit_length() - 1
            m &= m - 1
            if bin(nbm[v] & mask).count("1") + k < s:
                return False
        return True

    plexes = []
    for mask in range(1, 1 << n):
        if bin(mask).count("1") < q or not is_kplex(mask):
            continue
        if any(not (mask >> u) & 1 and is_kplex(mask | (1 << u)) for u in range(n)):
            continue
        plexes.append(tuple(v for v in range(n) if (mask >> v) & 1))
    return sorted(plexes)


def random_graph(n, p, seed):
    rng = random.Random(seed)
    adj = [set() for _ in range(n)]
    for u in range(n):
        for v in range(u + 1, n):
            if rng.random() < p:
                adj[u].add(v)
                adj[v].add(u)
    return adj


def multipartite(parts, t):
    """Complete multipartite graph K_{t,...,t} with `parts` parts."""
    n = parts * t
    adj = [set() for _ in range(n)]
    for u in range(n):
        for v in range(n):
            if u != v and u // t != v // t:
                adj[u].add(v)
    return adj


def disjoint_cliques(sizes):
    adj, base = [], 0
    for s in sizes:
        for i in range(s):
            adj.append({base + j for j in range(s) if j != i})
        base += s
    return adj


def csr(adj):
    deg = [len(a) for a in adj]
    nbr = [v for a in adj for v in sorted(a)]
    return deg, nbr


def edges(adj):
    return [(u, v) for u in range(len(adj)) for v in adj[u] if u < v]


def clique_plus_one(n):
    """A clique on 0..n-1 plus vertex n adjacent to 0..n-2 (not to n-1).  For
    any k with 2 <= k <= n and q <= n+1 the one maximal k-plex of size >= q is
    the whole graph (every vertex misses at most one other), and the root that
    finds it (vertex n, first in the degeneracy order) has vertex n-1 as its
    only 2-hop vertex: the k-search must take it."""
    adj = [set(range(n)) - {u} for u in range(n)]
    for u in range(n - 1):
        adj[u].add(n)
    adj.append(set(range(n - 1)))
    return adj
