"""Host symbolic analysis: pinned by the reference's AMD fixtures, its published
nested-dissection figures, and brute-force symbolic elimination.

CPU only (the analysis never touches the GPU)."""
import json
import os

import numpy as np
import pytest

import oracle as O
import rslab_amd as R
from rslab_amd import matgen

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---- pattern generators of crates/rslab-amd/tests/oracle_match.rs -------------
def csc_from_pairs(n, pairs):
    s = set()
    for i, j in pairs:
        s.add((i, j))
        s.add((j, i))
    cols = [[] for _ in range(n)]
    for r, c in s:
        cols[c].append(r)
    colp = [0]
    rows = []
    for c in cols:
        rows += sorted(c)
        colp.append(len(rows))
    return np.array(colp, dtype=np.int64), np.array(rows, dtype=np.int64)


def band(n, b):
    return csc_from_pairs(n, [(i, i) for i in range(n)] + [(i, i + k) for i in range(n) for k in range(1, b + 1) if i + k < n])


def arrow(n):
    return csc_from_pairs(n, [(i, i) for i in range(n)] + [(0, i) for i in range(1, n)])


def grid2(m, n):
    t = []
    for r in range(m):
        for c in range(n):
            k = r * n + c
            t.append((k, k))
            if r + 1 < m:
                t.append((k, (r + 1) * n + c))
            if c + 1 < n:
                t.append((k, r * n + c + 1))
    return csc_from_pairs(m * n, t)


FIXTURES = {
    "diag_4": lambda: band(4, 0),
    "tridiag_10": lambda: band(10, 1),
    "arrow_5": lambda: arrow(5),
    "arrow_200": lambda: arrow(200),
    "band_20_3": lambda: band(20, 3),
    "grid_7x7": lambda: grid2(7, 7),
    "amd_demo_24": lambda: grid2(6, 4),
}


def parse_fixture(name):
    d = {}
    for line in open(os.path.join(GOLD, "amd_oracle", name + ".txt")):
        line = line.strip()
        if not line or line.startswith("#") or ":" not in line:
            continue
        k, v = line.split(":", 1)
        d[k.strip()] = v.strip()
    return d


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_amd_matches_suitesparse_fixture(name):
    colp, rows = FIXTURES[name]()
    fx = parse_fixture(name)
    n = int(fx["n"])
    perm, st = R.amd_order(n, colp, rows)
    assert perm.tolist() == [int(x) for x in fx["perm"].split()]
    assert st["ncmpa"] == int(fx["ncmpa"])
    assert st["ndense"] == int(fx["n_dense"])
    for k in ("lnz", "ndiv", "nms_ldl", "nms_lu"):
        assert st[k] == int(fx[k]), k


def lower_from_full(colp, rows):
    n = len(colp) - 1
    lc, lr = [0], []
    for j in range(n):
        for k in range(colp[j], colp[j + 1]):
            if rows[k] >= j:
                lr.append(rows[k])
        lc.append(len(lr))
    return R.CscMatrix(n, np.array(lc), np.array(lr), np.ones(len(lr)))


def random_symmetric_lower(n, density, seed):
    rng = np.random.default_rng(seed)
    pairs = [(i, i) for i in range(n)]
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < density:
                pairs.append((i, j))
    return lower_from_full(*csc_from_pairs(n, pairs))


def true_structure(e):
    """Row structure of each column of L of the permuted pattern (brute force)."""
    n = e.n
    cols = [set() for _ in range(n)]
    for j in range(n):
        for k in range(e.pp_colp[j], e.pp_colp[j + 1]):
            i = int(e.pp_rowi[k])
            if i > j:
                cols[j].add(i)
    out = []
    for j in range(n):
        rows = sorted(cols[j])
        out.append(rows)
        if rows:
            cols[rows[0]].update(rows[1:])
    return out


@pytest.mark.parametrize("seed,n,dens,ordering", [
    (1, 40, 0.08, R.OrderingMethod.Amd),
    (2, 60, 0.05, R.OrderingMethod.Amd),
    (3, 80, 0.04, R.OrderingMethod.Natural),
    (4, 300, 0.01, R.OrderingMethod.MetisND),
])
def test_symbolic_invariants_brute_force(seed, n, dens, ordering):
    a = random_symmetric_lower(n, dens, seed)
    sym = R.LdltSymbolic.analyze_with(a, R.SolverSettings().with_ordering(ordering).with_nemin(4))
    e = sym.export()
    info = sym.info()
    # permutation is a bijection and its inverse is consistent
    assert sorted(e.perm.tolist()) == list(range(n))
    assert np.all(e.perm_inv[e.perm] == np.arange(n))
    # exact column counts (Gilbert-Ng-Peyton) == brute-force elimination
    cc = O.brute_force_col_counts(n, e.pp_colp, e.pp_rowi)
    assert info["colcount_nnz"] == int(cc.sum())
    # postordered: every etree subtree is a contiguous column range ending at its root
    parent = O.brute_force_etree(n, e.pp_colp, e.pp_rowi)
    first = np.arange(n)
    for j in range(n):
        p = parent[j]
        if p >= 0:
            assert p > j
            first[p] = min(first[p], first[j])
    size = np.ones(n, dtype=int)
    for j in range(n):
        if parent[j] >= 0:
            size[parent[j]] += size[j]
    assert np.all(first == np.arange(n) - size + 1)
    # supernodes: contiguous cover, children precede parents, rows cover L's structure
    ns = len(e.sn_first)
    assert e.sn_first[0] == 0 and np.all(e.sn_first[1:] == np.cumsum(e.sn_ncol)[:-1])
    assert int(e.sn_ncol.sum()) == n
    struct = true_structure(e)
    for s in range(ns):
        ch = e.ch_idx[e.ch_ptr[s]:e.ch_ptr[s + 1]]
        assert np.all(ch < s)
        f, w, h = int(e.sn_first[s]), int(e.sn_ncol[s]), int(e.sn_nrow[s])
        # row structure of s per the row-structure rule, compared with L's
        need = set()
        for j in range(f, f + w):
            need.update(r for r in struct[j] if r >= f + w)
        assert h == w + len(need)


def test_nd_reproduces_published_fill():
    pub = json.load(open(os.path.join(GOLD, "metis_published.json")))
    for m in (24, 40, 56):
        a = matgen.grid_laplacian((m, m, m))
        sym = R.LdltSymbolic.analyze_with(a, R.SolverSettings().with_ordering(R.OrderingMethod.MetisND))
        got = sym.info()["colcount_nnz"]
        want = pub["nnz_l"][str(m)]
        assert abs(got - want) / want < pub["rel_tol_nnz_l"], (m, got, want)
        if m == 40:
            e = sym.export()
            inf = sym.info()
            g = pub["grid40"]
            assert abs(inf["geom_flops"] - g["geom_flops"]) / g["geom_flops"] < g["rel_tol"]
            h = e.sn_nrow.astype(float)
            w = e.sn_ncol.astype(float)
            crit = np.zeros(len(h))
            for s in range(len(h)):
                ch = e.ch_idx[e.ch_ptr[s]:e.ch_ptr[s + 1]]
                crit[s] = h[s] * h[s] * w[s] + (crit[ch].max() if len(ch) else 0.0)
            assert abs(crit.max() - g["critical_path"]) / g["critical_path"] < g["rel_tol"]


def test_nd_permutation_valid_deterministic_and_disconnected():
    colp, rows = grid2(30, 30)
    p1 = R.nd_order(900, colp, rows)
    p2 = R.nd_order(900, colp, rows)
    assert sorted(p1.tolist()) == list(range(900))
    assert np.array_equal(p1, p2)
    # two disconnected grids (node_nd.rs component split)
    c1, r1 = grid2(16, 16)
    colp2 = np.concatenate([c1, c1[1:] + c1[-1]])
    rows2 = np.concatenate([r1, r1 + 256])
    p3 = R.nd_order(512, colp2, rows2)
    assert sorted(p3.tolist()) == list(range(512))
    # different seeds give different (valid) orderings
    p4 = R.nd_order(900, colp, rows, seed=2)
    assert sorted(p4.tolist()) == list(range(900)) and not np.array_equal(p1, p4)


def test_amalgamation_relax_widens_supernodes():
    a = matgen.grid_laplacian((40, 40))
    s1 = R.LdltSymbolic.analyze_with(a, R.SolverSettings()).info()
    s0 = R.LdltSymbolic.analyze_with(a, R.SolverSettings().with_relax(None)).info()
    assert s1["nsuper"] < s0["nsuper"]
    assert s1["colcount_nnz"] == s0["colcount_nnz"]  # counts are value-free and relax-free


def test_config_pattern_sizes():
    a1, b1, x1 = matgen.config1()
    assert (a1.n, a1.nnz) == (10000, 29800)
    a3, b3 = matgen.config3(m=60, nrhs=8)
    assert (a3.n, a3.nnz, b3.shape) == (216000, 853200, (216000, 8))
    a2, b2 = matgen.config2(m=40, nrhs=16)
    assert (a2.n, a2.nnz, b2.shape) == (64000, 251200, (64000, 16))
    assert np.iscomplexobj(a2.values)


def test_batch_symbolic_is_per_system_symbolic():
    """rsl_analyze_batch lays the systems' own analyses side by side."""
    systems = matgen.config4(count=6)
    mats = [a for a, _, _ in systems]
    eb = R.LdltSymbolic.analyze_batch(mats).export()
    off = soff = 0
    for m in mats:
        ei = R.LdltSymbolic.analyze(m).export()
        ns = len(ei.sn_first)
        assert np.array_equal(eb.perm[off:off + m.n] - off, ei.perm)
        assert np.array_equal(eb.sn_first[soff:soff + ns] - off, ei.sn_first)
        assert np.array_equal(eb.sn_ncol[soff:soff + ns], ei.sn_ncol)
        assert np.array_equal(eb.sn_nrow[soff:soff + ns], ei.sn_nrow)
        off += m.n
        soff += ns
    assert off == eb.n and soff == len(eb.sn_first)


def _export_arrays(sym):
    e = sym.export()
    return e, e.sn_parent()


@pytest.mark.parametrize("ordering", [R.OrderingMethod.Amd, R.OrderingMethod.MetisND])
def test_plan_from_caller_symbolic_round_trips(ordering):
    """rsl_plan_create on the permutation, supernode partition and tree that
    an analysis exported gives back the same symbolic: permutation, partition,
    panel heights, children, permuted pattern and work figures."""
    a, _ = matgen.config3(m=12, nrhs=1)
    st = R.SolverSettings().with_ordering(ordering)
    sym = R.LdltSymbolic.analyze_with(a, st)
    e, par = _export_arrays(sym)
    sym2 = R.LdltSymbolic.from_symbolic(a, e.perm, e.sn_first, e.sn_ncol, par, st)
    e2 = sym2.export()
    for k in ("perm", "perm_inv", "sn_first", "sn_ncol", "sn_nrow", "ch_ptr", "ch_idx", "pp_colp", "pp_rowi"):
        assert np.array_equal(getattr(e, k), getattr(e2, k)), k
    i1, i2 = sym.info(), sym2.info()
    for k in ("nsuper", "panel_entries", "levels", "nnzl_padded", "factor_flops", "colcount_nnz"):
        assert i1[k] == i2[k], k


def test_plan_from_caller_symbolic_without_amalgamation():
    """A different valid partition (every column its own supernode, the
    column elimination tree as the supernode tree) is accepted and gives the
    scalar structure: panel heights equal the column counts."""
    a = matgen.grid_laplacian((9, 9), shift=-2.0)
    sym = R.LdltSymbolic.analyze(a)
    e = sym.export()
    n = a.n
    # column etree of the permuted pattern
    parent = O.brute_force_etree(n, e.pp_colp, e.pp_rowi)
    sym2 = R.LdltSymbolic.from_symbolic(a, e.perm, np.arange(n), np.ones(n), parent)
    cc = O.brute_force_col_counts(n, e.pp_colp, e.pp_rowi)
    assert np.array_equal(sym2.export().sn_nrow, cc)


@pytest.mark.parametrize("case", ["perm_dup", "perm_len", "gap", "parent_earlier", "not_ancestor"])
def test_plan_from_caller_symbolic_rejects_invalid(case):
    a, _ = matgen.config3(m=8, nrhs=1)
    sym = R.LdltSymbolic.analyze_with(a, R.SolverSettings().with_ordering(R.OrderingMethod.MetisND))
    e, par = _export_arrays(sym)
    perm, first, ncol = e.perm.copy(), e.sn_first.copy(), e.sn_ncol.copy()
    if case == "perm_dup":
        perm[1] = perm[0]
    elif case == "perm_len":
        perm = perm[:-1]
    elif case == "gap":
        first[1] += 1
    elif case == "parent_earlier":
        par = par.copy()
        par[-1] = 0
    else:
        # re-hang a leaf supernode with off-diagonal rows under another root
        # branch: its rows then reach supernodes outside its new ancestors
        par = par.copy()
        leaf = next(s for s in range(len(first)) if e.sn_nrow[s] > e.sn_ncol[s] and par[s] != len(first) - 1)
        par[leaf] = len(first) - 1
    with pytest.raises(R.InvalidInput):
        R.LdltSymbolic.from_symbolic(a, perm, first, ncol, par)
