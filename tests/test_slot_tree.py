"""The row-end reduction tree of the register kernels, replayed on the CPU.

slot_tree (asgard_amd/csrc/kron_impl.hpp) sums the per-lane accumulators of a
wavefront over its column slots by recursive halving: at each lane bit from
32 down to the lane-group size G, a lane keeps half of the entries it still
owns and adds its partner's copy.  This replays those exact index steps for
every R (entries per lane) and power-of-two G the kernels use and checks the
partition property the GPU code relies on: every (sub, entry) pair is
emitted by exactly one lane, with the sum over the slots.  A wrong half
length, a lost padding slot or a mis-tracked end index fails here.
"""
import numpy as np
import pytest


def slot_tree_emit(V, G):
    lanes, R = V.shape
    vals = [list(V[l]) for l in range(lanes)]
    e = [0] * lanes
    end = [R] * lanes
    mask, LEN = lanes // 2, R
    while mask >= G:
        H = (LEN + 1) // 2
        new = []
        for l in range(lanes):
            up = (l & mask) != 0
            p = l ^ mask
            row = []
            for i in range(H):
                lo, hi = vals[l][i], (vals[l][H + i] if H + i < LEN else 0.0)
                plo, phi = vals[p][i], (vals[p][H + i] if H + i < LEN else 0.0)
                recv = plo if (p & mask) else phi  # what the partner sends
                row.append((hi if up else lo) + recv)
            new.append(row)
        e, end = ([e[l] + H if l & mask else e[l] for l in range(lanes)],
                  [end[l] if l & mask else min(end[l], e[l] + H) for l in range(lanes)])
        vals, LEN, mask = new, H, mask // 2
    out = {}
    for l in range(lanes):
        for i in range(LEN):
            if e[l] + i < end[l]:
                key = (l % G, e[l] + i)
                assert key not in out, key
                out[key] = vals[l][i]
    return out


@pytest.mark.parametrize("G", [1, 2, 4, 8])
def test_every_entry_once_with_its_slot_sum(G):
    rng = np.random.default_rng(G)
    for R in range(1, 65):
        V = rng.standard_normal((64, R))
        out = slot_tree_emit(V, G)
        assert len(out) == G * R
        for sub in range(G):
            for r in range(R):
                want = V[sub::G, r].sum()
                assert abs(out[(sub, r)] - want) <= 1e-12 * max(1.0, np.abs(V[sub::G, r]).sum())
