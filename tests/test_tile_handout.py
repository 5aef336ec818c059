"""Forward progress of the onesweep pass's tile hand-out under limited
residency (CPU model, no GPU).

A tile's decoupled look-back waits until every lower tile has published its
prefix, so tiles finish in tile order.  Workgroups start in blockIdx order
whenever one of R resident slots is free, and take their tile at start:
  - one ticket counter (group 0): tile = arrival order;
  - grouped hand-out (group G): workgroup b belongs to class b % 8 (its XCD)
    and takes class_tile(class, j, G, tiles) for its class's j-th ticket
    (the mapping of gpusort_amd/csrc/radix_onesweep.hip, class_tile).
The model reports a deadlock when every resident workgroup waits on a lower
tile that no started workgroup holds.  The arrival-order ticket never
deadlocks; the grouped one needs about 8*G resident workgroups, which is why
the library uses it only on the null stream (a sort that has the GPU to
itself) and the arrival-order ticket on every other stream.
"""
from __future__ import annotations

import pytest


def class_tile(c: int, j: int, G: int, tiles: int) -> int:
    full = tiles // (8 * G)
    if j < full * G:
        return (j // G) * 8 * G + c * G + (j % G)
    return full * 8 * G + (j - full * G) * 8 + c


def simulate(tiles: int, resident: int, group: int) -> bool:
    """True when every tile finishes, False on a deadlock."""
    next_block = 0          # next workgroup to dispatch (blockIdx order)
    next_done = 0           # tiles [0, next_done) have finished
    held = set()            # tiles held by resident workgroups
    ticket = 0
    class_ticket = [0] * 8
    while next_done < tiles:
        progressed = False
        while len(held) < resident and next_block < tiles:
            b = next_block
            next_block += 1
            if group == 0:
                t = ticket
                ticket += 1
            else:
                c = b % 8
                t = class_tile(c, class_ticket[c], group, tiles)
                class_ticket[c] += 1
            held.add(t)
            progressed = True
        while next_done in held:  # look-back closes in tile order
            held.remove(next_done)
            next_done += 1
            progressed = True
        if not progressed:
            return False
    return True


def test_class_tile_is_a_bijection():
    for tiles in (1, 7, 128, 129, 1000, 13797):
        for G in (1, 4, 16):
            per_class = [(tiles - c + 7) // 8 for c in range(8)]  # blocks of each class in a grid of `tiles`
            got = sorted(class_tile(c, j, G, tiles) for c in range(8) for j in range(per_class[c]))
            assert got == list(range(tiles)), (tiles, G)


@pytest.mark.parametrize("resident", [1, 2, 3, 17, 256])
def test_arrival_order_ticket_always_progresses(resident):
    for tiles in (1, 9, 1000, 13797):
        assert simulate(tiles, resident, 0)


def test_grouped_handout_needs_resident_workgroups():
    tiles = 13797  # 2^28 u32 keys with u32 payloads at 19 K keys per tile
    # One workgroup per CU on a GPU of 256 CUs, G = 16: progresses.
    assert simulate(tiles, 256, 16)
    assert simulate(tiles, 8 * 16, 16)
    # A sort that shares the GPU can get far fewer slots: grouped hand-out
    # then deadlocks where the arrival-order ticket does not.
    assert not simulate(tiles, 64, 16)
    assert not simulate(tiles, 8, 4)
    assert simulate(tiles, 64, 0) and simulate(tiles, 8, 0)
