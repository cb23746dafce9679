"""Hines matrix, elimination and walkTree numbering, pinned without the
reference build:
  - the 9 hand-drawn trees of hsolve/HinesMatrix.cpp:503-689 (children in Hines
    indices, Ra=15+3i, Rm=45+15i, Cm=500+200i^2, dt=1) against a dense matrix
    built as hsolve/testHSolve.cpp:37-108 does, |error| < 1e-17;
  - forward elimination / back-substitution for two passes against dense
    Gaussian elimination (hsolve/HSolvePassive.cpp:852-960), relative 4 ulp;
  - walkTree: reverse DFS pre-order from a terminal, checked on hand-derived
    orders and on structural invariants of random trees."""
import numpy as np
import pytest

from moose_core_amd import workloads
from moose_core_amd.cellspec import CellSpec, passive_compartments

TREES = [
    {0: [], 1: [0], 2: [1], 3: [], 4: [3], 5: [], 6: [5], 7: [4, 6], 8: [7], 9: [8], 10: [],
     11: [10], 12: [], 13: [12], 14: [11, 13], 15: [14, 16], 16: [17], 17: [2, 9, 18], 18: [19], 19: []},
    {0: [], 1: [], 2: [0, 1, 3], 3: []},
    {0: [2], 1: [], 2: [1, 3], 3: []},
    {0: [], 1: [], 2: [0, 1], 3: [2]},
    {0: [], 1: [2], 2: [0, 4], 3: [], 4: [3, 5], 5: []},
    {0: [], 1: [], 2: [], 3: [4], 4: [0, 1, 2, 5, 6], 5: [], 6: []},
    {0: []},
    {0: [], 1: [0, 2], 2: []},
    {0: [1], 1: [2], 2: []},
]


def full_matrix(children, Ra, Rm, Cm, dt):
    """Dense Hines matrix as makeFullMatrix builds it (testHSolve.cpp:37-108)."""
    n = len(children)
    CmByDt = [Cm[i] / (dt / 2.0) for i in range(n)]
    Ga = [2.0 / Ra[i] for i in range(n)]
    coupled = [list(children[i]) + [i] for i in range(n) if len(children[i]) >= 1]
    M = np.zeros((n, n))
    for i in range(n):
        M[i, i] = CmByDt[i] + 1.0 / Rm[i]
    for g in coupled:
        gsum = sum(Ga[c] for c in g)
        for c in g:
            M[c, c] += Ga[c] * (1.0 - Ga[c] / gsum)
    for g in coupled:
        gsum = sum(Ga[c] for c in g)
        for a in range(len(g) - 1):
            for b in range(a + 1, len(g)):
                gij = Ga[g[a]] * Ga[g[b]] / gsum
                M[g[a], g[b]] = -gij
                M[g[b], g[a]] = -gij
    return M


def params(n):
    i = np.arange(n, dtype=np.float64)
    return 15.0 + 3.0 * i, 45.0 + 15.0 * i, 500.0 + 200.0 * i * i


@pytest.mark.parametrize("k", range(len(TREES)))
def test_hines_matrix_against_dense(oracle_mod, k):
    t = TREES[k]
    n = len(t)
    children = [t[i] for i in range(n)]
    Ra, Rm, Cm = params(n)
    m = oracle_mod.OracleMatrix(children, Ra, Rm, Cm, np.full(n, -0.06), np.full(n, -0.06), 1.0)
    M = full_matrix(children, Ra, Rm, Cm, 1.0)
    assert np.abs(m.A() - M).max() < 1e-17


def close(a, b, ulps):
    """isClose of hsolve/HSolvePassive.cpp:430-446: |(a-b)/a| and |(a-b)/b| < ulps*eps."""
    eps = ulps * np.finfo(float).eps
    a, b = np.asarray(a), np.asarray(b)
    d = np.abs(a - b)
    with np.errstate(divide="ignore", invalid="ignore"):
        ok = (d == 0) | ((d / np.abs(a) < eps) & (d / np.abs(b) < eps))
    return bool(np.all(ok))


@pytest.mark.parametrize("k", range(len(TREES)))
def test_elimination_against_dense_gauss(oracle_mod, k):
    t = TREES[k]
    n = len(t)
    children = [t[i] for i in range(n)]
    Ra, Rm, Cm = params(n)
    Em = np.full(n, -0.06)
    V = -0.06 + 0.01 * np.arange(n)
    dt = 1.0
    m = oracle_mod.OracleMatrix(children, Ra, Rm, Cm, Em, V.copy(), dt)
    M0 = full_matrix(children, Ra, Rm, Cm, dt)
    Vloc = V.copy()
    for _ in range(2):
        m.update()
        M = M0.copy()
        B = Vloc * Cm / (dt / 2.0) + Em / Rm
        assert close(m.B(), B, 4)
        m.forward()
        for i in range(n - 1):
            for j in range(i + 1, n):
                div = M[j, i] / M[i, i]
                M[j, :] -= div * M[i, :]
                B[j] -= div * B[i]
        assert close(np.triu(m.A()), np.triu(M), 4)
        m.backward()
        VMid = np.zeros(n)
        for i in range(n - 1, -1, -1):
            VMid[i] = B[i]
            for j in range(n - 1, i, -1):
                VMid[i] -= VMid[j] * M[i, j]
            VMid[i] /= M[i, i]
            Vloc[i] = 2 * VMid[i] - Vloc[i]
        assert close(m.VMid(), VMid, 4)
        assert close(m.V(), Vloc, 4)


def chain_cell(axial, n, seed=0):
    comp, _ = passive_compartments(np.full(n, 10e-6), np.full(n, 1e-6))
    return CellSpec(comp, np.array(axial, dtype=np.int32).reshape(-1, 2), seed=seed)


def test_walktree_hand_derived(oracle_mod):
    # hines_order()[c] = Hines index of compartment c.
    # soma 0 with children 1 (first message) and 2: the terminal found from the
    # soma is 1 (first neighbour); DFS from 1 visits 1, 0, 2; reversed: 2, 0, 1
    cell = oracle_mod.OracleCell(chain_cell([(0, 1), (0, 2)], 3), 1e-4)
    assert list(cell.hines_order()) == [1, 2, 0]
    # a 4-chain 0-1-2-3 seeded at 0 (a terminal already): DFS 0,1,2,3 reversed
    cell = oracle_mod.OracleCell(chain_cell([(0, 1), (1, 2), (2, 3)], 4), 1e-4)
    assert list(cell.hines_order()) == [3, 2, 1, 0]
    # seeded in the middle: follow adjacent[0] (child 2) to the terminal 3
    cell = oracle_mod.OracleCell(chain_cell([(0, 1), (1, 2), (2, 3)], 4, seed=1), 1e-4)
    assert list(cell.hines_order()) == [0, 1, 2, 3]
    # star: soma 0 with children 1, 2, 3 (in message order); terminal = 1;
    # DFS 1, 0, then 0's other neighbours popped last-first: 3, then 2
    cell = oracle_mod.OracleCell(chain_cell([(0, 1), (0, 2), (0, 3)], 4), 1e-4)
    # visit order 1, 0, 3, 2 -> Hines: 2->0, 3->1, 0->2, 1->3
    assert list(cell.hines_order()) == [2, 3, 0, 1]


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_walktree_invariants(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    spec = workloads.random_ca_cell(rng, 200)
    order = oracle_mod.OracleCell(spec, 1e-4).hines_order()
    n = spec.n
    assert sorted(order) == list(range(n))
    # reverse pre-order: every node's subtree (in the tree rooted at the Hines
    # root) is a contiguous block ending at the node
    adj = [[] for _ in range(n)]
    for p, c in spec.axial:
        adj[p].append(c)
        adj[c].append(p)
    h2d = np.argsort(order)
    root = h2d[n - 1]
    assert len(adj[root]) <= 1  # the Hines root is a terminal
    size = np.ones(n, dtype=int)
    parent = -np.ones(n, dtype=int)
    stack, seen, post = [root], {root}, []
    while stack:
        x = stack.pop()
        post.append(x)
        for y in adj[x]:
            if y not in seen:
                seen.add(y)
                parent[y] = x
                stack.append(y)
    for x in reversed(post):
        if parent[x] >= 0:
            size[parent[x]] += size[x]
    for x in range(n):
        if parent[x] >= 0:
            assert order[parent[x]] > order[x]
        lo = order[x] - size[x] + 1
        members = {h2d[h] for h in range(lo, order[x] + 1)}
        assert len(members) == size[x]
