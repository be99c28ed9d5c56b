"""Exhaustive interleaving model of the global-slot hand-off of
k_path_swath_bands with one wave per band (odba_dtw_kernels.hip), on the
CPU.

Each claim q sweeps swath q of one open-end job and runs these steps, in
this order, as the kernel does:
  read   skip = stop <= q                          (pp.cancelled())
  take   took = has_out and not skip; a slot or -2 is announced in slot_of[q]
  wait   (q > 0) wait for slot_of[q-1]; -2 turns skip on (had_slot = False)
  sweep  (not skip) may decide the stop: stop = min(stop, q + 1)
  mark   the slot's done mark (rule under test)
  back   (had_slot) wait for the left slot's done mark, then return it
Claims run concurrently (separate workgroups), so a later claim may read
the stop word before an earlier one. Every interleaving of the steps of
three or four claims is explored; a state where some claim is unfinished
and none can move is a hang (in the kernel: a wait that times out and fails
the round).

A second model adds the sweep's row wait and the hand-off across the
two-phase claim boundary of open-end swath bands (through a job's own
boundary slot, with no announcement and no pool slot) at every position.
"""
import pytest

INF = 10**9


def run_all(nclaims, stop_decider, mark_rule):
    """Explore every interleaving; return the number of hung end states."""
    hangs = 0
    seen = set()

    def step(state, q):
        pcs, skip, took, had, stop, slot_of, done = state
        pcs, skip, took, had = list(pcs), list(skip), list(took), list(had)
        slot_of, done = list(slot_of), list(done)
        pc = pcs[q]
        has_out = q < nclaims - 1
        if pc == 0:  # read
            skip[q] = stop <= q
        elif pc == 1:  # take
            took[q] = has_out and not skip[q]
            if has_out:
                slot_of[q] = q if took[q] else -2  # slot q: a distinct slot per producer
        elif pc == 2:  # wait for the left announcement
            if q > 0:
                if slot_of[q - 1] == -1:
                    return None  # blocked
                had[q] = slot_of[q - 1] >= 0
                if not had[q]:
                    skip[q] = True
        elif pc == 3:  # sweep
            if not skip[q] and q == stop_decider:
                stop = min(stop, q + 1)
        elif pc == 4:  # done mark
            if mark_rule(took[q], has_out, skip[q]):
                done[q] = True
        elif pc == 5:  # wait for the left slot's done mark
            if had[q] and not done[q - 1]:
                return None
        pcs[q] = pc + 1
        return (tuple(pcs), tuple(skip), tuple(took), tuple(had), stop, tuple(slot_of), tuple(done))

    init = ((0,) * nclaims, (False,) * nclaims, (False,) * nclaims, (False,) * nclaims, INF,
            (-1,) * nclaims, (False,) * nclaims)
    stack = [init]
    while stack:
        st = stack.pop()
        if st in seen:
            continue
        seen.add(st)
        moved = False
        for q in range(nclaims):
            if st[0][q] < 6:
                nxt = step(st, q)
                if nxt is not None:
                    moved = True
                    stack.append(nxt)
        if not moved and any(pc < 6 for pc in st[0]):
            hangs += 1
    return hangs


def old_rule(took, has_out, skip):
    return has_out and not skip  # the done mark only for a strip that swept


def new_rule(took, has_out, skip):
    return took  # the done mark for every announced slot


@pytest.mark.parametrize("nclaims,decider", [(4, 0), (5, 0), (5, 1)])
def test_announced_slot_always_gets_its_done_mark(nclaims, decider):
    assert run_all(nclaims, decider, old_rule) > 0  # the hang the old rule allowed
    assert run_all(nclaims, decider, new_rule) == 0


def run_all_phases(nclaims, stop_decider, boundary):
    """The same protocol with the sweep's row wait made explicit and the
    second-phase boundary of k_path_swath_bands: claim `boundary` - 1 hands
    its edge to claim `boundary` through the job's own boundary slot, so the
    producer takes no pool slot and announces nothing, and the consumer
    neither waits for an announcement nor returns a slot. A consumer's sweep
    ends when its producer has swept, or when the stop word has reached its
    own swath (the wait then returns: BandPipe::wait polls cancelled()).
    Returns the number of hung end states."""
    hangs = 0
    seen = set()

    def step(state, q):
        pcs, skip, took, had, stop, slot_of, done, swept = state
        pcs, skip, took, had = list(pcs), list(skip), list(took), list(had)
        slot_of, done, swept = list(slot_of), list(done), list(swept)
        pc = pcs[q]
        has_out = q < nclaims - 1
        bnd_out = q == boundary - 1
        bnd_in = q == boundary
        if pc == 0:  # read
            skip[q] = stop <= q
        elif pc == 1:  # take
            pool_out = has_out and not bnd_out
            took[q] = pool_out and not skip[q]
            if pool_out:
                slot_of[q] = q if took[q] else -2
        elif pc == 2:  # wait for the left announcement (pool consumers only)
            if q > 0 and not bnd_in:
                if slot_of[q - 1] == -1:
                    return None
                had[q] = slot_of[q - 1] >= 0
                if not had[q]:
                    skip[q] = True
        elif pc == 3:  # sweep: waits for the producer's rows unless cancelled
            if not skip[q]:
                if q > 0 and not swept[q - 1]:
                    if stop <= q:
                        skip[q] = True  # cancelled while waiting
                    else:
                        return None
                if not skip[q]:
                    swept[q] = True
                    if q == stop_decider:
                        stop = min(stop, q + 1)
        elif pc == 4:  # done mark (pool producers that announced a slot)
            if took[q]:
                done[q] = True
        elif pc == 5:  # wait for the left slot's done mark
            if had[q] and not done[q - 1]:
                return None
        pcs[q] = pc + 1
        return (tuple(pcs), tuple(skip), tuple(took), tuple(had), stop, tuple(slot_of), tuple(done), tuple(swept))

    init = ((0,) * nclaims, (False,) * nclaims, (False,) * nclaims, (False,) * nclaims, INF,
            (-1,) * nclaims, (False,) * nclaims, (False,) * nclaims)
    stack = [init]
    while stack:
        st = stack.pop()
        if st in seen:
            continue
        seen.add(st)
        moved = False
        for q in range(nclaims):
            if st[0][q] < 6:
                nxt = step(st, q)
                if nxt is not None:
                    moved = True
                    stack.append(nxt)
        if not moved and any(pc < 6 for pc in st[0]):
            hangs += 1
    return hangs


@pytest.mark.parametrize("nclaims", [3, 4])
def test_phase_boundary_handoff_never_hangs(nclaims):
    """Every boundary position, every claim deciding the stop (or none), and
    every interleaving of the claims' steps: nothing hangs."""
    for boundary in range(1, nclaims):
        for decider in list(range(nclaims)) + [None]:
            assert run_all_phases(nclaims, decider, boundary) == 0, (boundary, decider)
    # without a boundary (no claim at that position) the row-wait model
    # agrees with the slot model above
    for decider in range(nclaims):
        assert run_all_phases(nclaims, decider, -5) == 0
