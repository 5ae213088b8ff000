"""The SpMV formats' internal row order (panel.hip internal_order), run on the
CPU through the host mirror.  The order is a permutation of the basis rows
keyed by the orbit, not by the member the reference picked: every row's
state rotated to its smallest lower-layer word x (ties: smallest upper word
y), then (x, necklace class of y, position of y in its necklace).  What it is
for is checked on the oracle's Hamiltonian: in the internal order more lanes
of a 64-row step-major slice hold an entry at each step and a step gathers
from fewer x cache lines than in the reference order."""
import numpy as np
import pytest

import blqh_amd
from tests import hostmirror


def _rot(x, r, nphi):
    m = (1 << nphi) - 1
    return ((x << r) | (x >> (nphi - r))) & m if r else x


def _key(w, S, nphi):
    """(x, class(y), position) of word w, by plain Python loops."""
    m = (1 << nphi) - 1
    sp = [(w >> (s * nphi)) & m for s in range(S)]
    half = S // 2

    def grp(lo, hi, r):
        return sum(_rot(sp[s], r, nphi) << ((s - lo) * nphi) for s in range(lo, hi))

    xs = [(grp(half, S, r), grp(0, half, r)) for r in range(nphi)]
    bx, by = min(xs)
    ys = [grp(0, half, r) for r in range(nphi)]
    cls = min(ys)
    r0 = ys.index(cls)
    ridx = next(k for k in range(nphi) if ys[(r0 + k) % nphi] == by)
    return (bx, cls, ridx)


@pytest.mark.parametrize("case", [(2, 8, 8, -1, -1, 0, 0), (2, 10, 10, 5, -1, 3, 2), (4, 4, 4, 2, 2, 1, 3)],
                         ids=["alldens_K0", "split_complex", "spinful"])
def test_order_is_the_keyed_permutation(case):
    sec = blqh_amd.Sector(*case)
    ids, _ = hostmirror.basis(sec)
    order = hostmirror.internal_order(sec, ids)
    assert order is not None
    assert np.array_equal(np.sort(order), np.arange(len(ids)))
    S, nphi = case[0], case[1]
    keys = [_key(int(w), S, nphi) for w in ids]
    inv = np.empty_like(order)
    inv[order] = np.arange(len(order))
    seq = [keys[i] for i in inv]
    assert all(a < b for a, b in zip(seq, seq[1:]))  # strictly increasing: one key per orbit


def test_no_translation_reduction_keeps_the_reference_order():
    sec = blqh_amd.Sector(2, 6, 6, -1, -1, 0, -1)
    ids, _ = hostmirror.basis(sec)
    assert hostmirror.internal_order(sec, ids) is None


def _slice_stats(rowptr, col, order, width):
    """Step-major runs of 64-row slices and `width`-column panels, columns
    ascending: mean lanes holding an entry per step, and distinct 8-column x
    lines per entry."""
    n = len(rowptr) - 1
    inv = np.empty(n, np.int64)
    inv[order] = np.arange(n)
    steps = ent = lines = 0
    for s0 in range(0, n, 64):
        rows = [np.sort(order[col[rowptr[o]:rowptr[o + 1]]]) for o in inv[s0:s0 + 64]]
        for p0 in range(0, n, width):
            run = [r[(r >= p0) & (r < p0 + width)] for r in rows]
            mx = max(len(r) for r in run)
            steps += mx
            for t in range(mx):
                c = [r[t] for r in run if len(r) > t]
                ent += len(c)
                lines += len(set(x >> 3 for x in c))
    return ent / steps, lines / ent


def test_order_packs_slices_on_the_oracle_hamiltonian(oracle_mod):
    O = oracle_mod
    case = (2, 10, 10, -1, -1, 0, 0)  # 6,166-state all-density sector, real
    sec = blqh_amd.Sector(*case)
    ids, _ = hostmirror.basis(sec)
    OB = O.Basis(O.sector(*case))
    assert np.array_equal(OB.id, ids)
    OH = O.Hamiltonian(OB, O.params(10, d=1.0, dsas=0.05))
    order = hostmirror.internal_order(sec, ids)
    # 24 column panels: about the panels per row of the 2.35M-state sector (36)
    w = 256
    lanes_ref, lines_ref = _slice_stats(OH.rowptr, OH.col, np.arange(len(ids)), w)
    lanes_int, lines_int = _slice_stats(OH.rowptr, OH.col, order, w)
    print(f"active lanes {lanes_ref:.1f} -> {lanes_int:.1f}, lines per entry {lines_ref:.3f} -> {lines_int:.3f}")
    # measured: 29.8 -> 36.0 lanes, 0.374 -> 0.189 lines per entry (2.35M states,
    # 65,536-column panels, offline: 23.6 -> 54.1 lanes, 0.85 -> 0.41)
    assert lanes_int > 1.1 * lanes_ref
    assert lines_int < 0.6 * lines_ref
