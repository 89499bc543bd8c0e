"""Second, independent restatement of the reference's partition rules in plain
Python, written from the definitions rather than from the bisection loops:

* a search is expressed through bisect (largest / first index satisfying the
  order predicate) plus the reference's final "exact hit on a[r]" rule;
* the double formulas are evaluated with Python floats (IEEE binary64) and the
  int32 formulas with Python ints under an explicit overflow check.

TEST INFRASTRUCTURE ONLY.  Agreement between this module and oracle.c pins the
double rounding of the v1_numa formulas (src/dspmv_mgpu_v1_numa.cu:210-214,
:432-436) and the search semantics (src/spmv_helper.cu:18-44,646-685).
"""
from __future__ import annotations

import bisect
import math

INT_MAX = 2**31 - 1
UNDEFINED = "undefined"


def grfi(n, a, idx):
    """get_row_from_index: largest l in [0, n) with a[l] <= idx (0 if none);
    when l == n-1, a[l] != idx and a[n] == idx the reference returns n."""
    if n <= 0:
        return 0
    l = max(bisect.bisect_right(a, idx, 0, n) - 1, 0)
    if a[l] == idx:
        return l
    if l == n - 1 and n < len(a) and a[n] == idx:
        return n
    return l


def ffs(a, n, key, alen=None):
    """findFirstInSorted: first position of key, else the last index < key."""
    alen = len(a) if alen is None else alen
    if n <= 0:
        return 0
    r = bisect.bisect_left(a, key, 0, n)
    l = max(r - 1, 0)
    if a[l] == key:
        return l
    if r < alen and a[r] == key:
        return r
    return l


def fls(a, n, key, alen=None):
    """findLastInSorted: last position of key, else the last index <= key."""
    alen = len(a) if alen is None else alen
    if n <= 0:
        return 0
    l = max(bisect.bisect_right(a, key, 0, n) - 1, 0)
    if a[l] == key:
        return l
    if l + 1 < alen and l + 1 == n and a[l + 1] == key:
        return l + 1
    return l


def _chunk(si, ei, sr, er, sf, ef, mc, nnzc):
    return dict(start_idx=si, end_idx=ei, start_row=sr, end_row=er,
                start_flag=int(bool(sf)), end_flag=int(bool(ef)), m_c=mc, nnz_c=nnzc)


def _numa(mapping, C):
    mp = [0] * C if mapping is None else list(mapping)
    nodes = max(mp) + 1
    gpus = [mp.count(k) for k in range(nodes)]
    before = [sum(gpus[:k]) for k in range(nodes + 1)]
    lids = [mp[:d].count(mp[d]) for d in range(C)]
    return mp, nodes, gpus, before, lids


def part_v1(m, rp, nnz, C):
    if C * nnz > INT_MAX:
        return UNDEFINED
    out = []
    for i in range(C):
        si = math.floor(float(i * nnz) / float(C))
        ei = math.floor(float((i + 1) * nnz) / float(C)) - 1
        sr = grfi(m, rp, si)
        er = grfi(m, rp, ei)
        out.append(_chunk(si, ei, sr, er, si > rp[sr], ei < rp[er + 1] - 1,
                          er - sr + 1, ei - si + 1))
    return out


def part_v2(m, rp, nnz, nb, q):
    cap = int(0.8 * 16 * 1e9 / 16.0) // q
    nb = min(nb, cap)
    if nnz + nb - 1 > INT_MAX:
        return UNDEFINED
    T = (nnz + nb - 1) // nb
    if T * nnz > INT_MAX:
        return UNDEFINED
    out = []
    for t in range(T):
        si = (t * nnz) // T
        ei = ((t + 1) * nnz) // T - 1
        sr = grfi(m, rp, si)
        er = grfi(m, rp, ei)
        out.append(_chunk(si, ei, sr, er, si > rp[sr], ei < rp[er + 1] - 1,
                          er - sr + 1, ei - si + 1))
    return out


def part_rowblock(m, rp, C):
    if C * m > INT_MAX:
        return UNDEFINED
    out = []
    for d in range(C):
        sr = (d * m) // C
        er = ((d + 1) * m) // C - 1
        nnzc = rp[er + 1] - rp[sr]
        out.append(_chunk(rp[sr], rp[sr] + nnzc - 1, sr, er, 0, 0, er - sr + 1, nnzc))
    return out


def part_baseline_coo(m, row_idx, C):
    nnz = len(row_idx)
    out = []
    for d in range(C):
        sr = (d * m) // C
        er = ((d + 1) * m) // C - 1
        si = ffs(row_idx, nnz, sr)
        ei = fls(row_idx, nnz, er)
        out.append(_chunk(si, ei, sr, er, 0, 0, er - sr + 1, ei - si + 1))
    return out


def part_v1_numa(m, rp, nnz, C, mapping=None):
    mp, nodes, gpus, before, lids = _numa(mapping, C)
    t = float(nnz) / float(C)
    out = [None] * C
    node_out = [None] * nodes
    for k in range(nodes):
        if gpus[k] == 0:
            continue
        S = math.floor(before[k] * t)
        E = math.floor(before[k + 1] * t) - 1
        SR, ER = grfi(m + 1, rp, S), grfi(m + 1, rp, E)
        SF, EF = S > rp[SR], E < rp[ER + 1] - 1
        M, N1 = ER - SR + 1, E - S + 1
        node_out[k] = _chunk(S, E, SR, ER, SF, EF, M, N1)
        L = [0] + [rp[SR + j] - S for j in range(1, M)] + [N1]
        t2 = float(N1) / float(gpus[k])
        for d in range(C):
            if mp[d] != k:
                continue
            j = lids[d]
            si = math.floor(j * t2)
            ei = math.floor((j + 1) * t2) - 1
            sr, er = grfi(M, L, si), grfi(M, L, ei)
            sf = SF if j == 0 else si > L[sr]
            ef = EF if j + 1 == gpus[k] else ei < L[er + 1] - 1
            out[d] = _chunk(S + si, S + ei, SR + sr, SR + er, sf, ef, er - sr + 1, ei - si + 1)
    return out, node_out


def part_v1_numa_csc(n, cp, nnz, C, mapping=None):
    mp, nodes, gpus, before, lids = _numa(mapping, C)
    t = float(nnz) / float(C)
    out = [None] * C
    node_out = [None] * nodes
    for k in range(nodes):
        if gpus[k] == 0:
            continue
        S = math.floor(before[k] * t)
        E = math.floor(before[k + 1] * t) - 1
        SC, EC = grfi(n + 1, cp, S), grfi(n + 1, cp, E)
        SF, EF = S > cp[SC], E < cp[EC + 1] - 1
        N, N1 = EC - SC + 1, E - S + 1
        node_out[k] = _chunk(S, E, SC, EC, SF, EF, N, N1)
        L = [0] + [cp[SC + j] - S for j in range(1, N)] + [N1]
        t2 = float(N1) / float(gpus[k])
        for d in range(C):
            if mp[d] != k:
                continue
            j = lids[d]
            si = math.floor(j * t2)
            ei = math.floor((j + 1) * t2) - 1
            sc, ec = grfi(N, L, si), grfi(N, L, ei)
            out[d] = _chunk(S + si, S + ei, SC + sc, SC + ec, si > L[sc], ei < L[ec + 1] - 1,
                            ec - sc + 1, ei - si + 1)
    return out, node_out


def part_v1_numa_coo(m, row_idx, C, mapping=None):
    mp, nodes, gpus, before, lids = _numa(mapping, C)
    nnz = len(row_idx)
    t = float(nnz) / float(C)
    out = [None] * C
    node_out = [None] * nodes
    for k in range(nodes):
        if gpus[k] == 0:
            continue
        S = math.floor(before[k] * t)
        E = math.floor(before[k + 1] * t) - 1
        SR, ER = row_idx[S], row_idx[E]
        SF = S > ffs(row_idx, nnz, SR)
        EF = E < fls(row_idx, nnz, ER)
        M, N1 = ER - SR + 1, E - S + 1
        node_out[k] = _chunk(S, E, SR, ER, SF, EF, M, N1)
        R = [v - SR for v in row_idx[S:E + 1]]
        t2 = float(N1) / float(gpus[k])
        for d in range(C):
            if mp[d] != k:
                continue
            j = lids[d]
            si = math.floor(j * t2)
            ei = math.floor((j + 1) * t2) - 1
            sr, er = R[si], R[ei]
            nnzc = ei - si + 1
            sf = SF if j == 0 else si > ffs(R, nnzc, sr, alen=N1)
            ef = EF if j + 1 == gpus[k] else ei < fls(R, nnzc, er, alen=N1)
            out[d] = _chunk(S + si, S + ei, SR + sr, SR + er, sf, ef, er - sr + 1, nnzc)
    return out, node_out
