"""Checkpoint plans of the memory-bounded gradient (fwi3d_grad_plan, host only).

Properties that fix the plan: the top segments cover the Nt steps exactly; at
every depth a segment splits into at most k pieces of the next depth's length;
k is the smallest integer with k^(levels+1) >= Nt; the stored-state count is
the top checkpoints, k slots per inner depth, the leaf buffer and the adjoint
state; and walking the segments depth-first in reverse visits every step once,
from the last to the first (the order in which the adjoint consumes them)."""
import math

import numpy as np
import pytest

from mittet_main_amd.adjoint import grad_plan

NTS = [1, 2, 7, 64, 199, 1000, 3839, 7999, 11519]


def leaves_in_reverse(nt, seglen, nseg):
    out = []

    def walk(depth, n0, L):
        if depth == len(seglen) - 1:
            out.append((n0, L))
            return
        P = seglen[depth + 1]
        npieces = -(-L // P)
        for p in reversed(range(npieces)):
            walk(depth + 1, n0 + p * P, min(P, L - p * P))

    for q in reversed(range(nseg)):
        n0 = q * seglen[0]
        walk(0, n0, min(seglen[0], nt - n0))
    return out


@pytest.mark.parametrize("nt", NTS)
def test_single_level_is_the_sqrt_scheme(nt):
    sl, nseg, states = grad_plan(nt, 1)
    assert sl == [math.ceil(math.sqrt(nt))]
    assert nseg == -(-nt // sl[0])
    assert states == nseg + sl[0] + 2


@pytest.mark.parametrize("levels", [2, 3, 4, 5])
@pytest.mark.parametrize("nt", NTS)
def test_plan_covers_every_step_once(nt, levels):
    sl, nseg, states = grad_plan(nt, levels)
    k = 1
    while k ** (levels + 1) < nt:
        k += 1
    assert sl == [min(k ** (levels - l), nt) for l in range(levels)]
    assert nseg * sl[0] >= nt > (nseg - 1) * sl[0]
    inner = sum(-(-sl[l - 1] // sl[l]) for l in range(1, levels))
    assert states == nseg + inner + sl[-1] + 2
    assert states <= (levels + 1) * k + 2
    # leaves, taken depth-first in reverse, tile [0, nt) from the end
    steps = []
    for n0, L in leaves_in_reverse(nt, sl, nseg):
        assert 1 <= L <= sl[-1]
        steps += list(range(n0 + L - 1, n0 - 1, -1))
    assert steps == list(range(nt - 1, -1, -1))


def test_more_levels_store_fewer_states():
    for nt in (3839, 11519):
        counts = [grad_plan(nt, lv)[2] for lv in range(1, 6)]
        assert counts == sorted(counts, reverse=True) and counts[0] > 3 * counts[3]


def test_plan_rejects_bad_requests():
    from mittet_main_amd._lib import Fwi3dError
    for nt, lv in ((0, 1), (10, 0), (10, 9)):
        with pytest.raises(Fwi3dError):
            grad_plan(nt, lv)
