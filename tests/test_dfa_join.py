"""The reference DFA (classify.rs:147-518) is a join on per-side label sets.

Every occurrence observes an in-label and an out-label in canonical
orientation (a base, or "none" for a missing neighbour / forced MIMO). The
state reached by the reference's branch transitions from any sequence of
observations equals the state encoded from the union of the observed labels:
SISO(f,b) when both sides hold exactly one base, MISO/SIMO when one side does,
MIMO otherwise (src/state.rs:92-113 codes). The GPU classifies by OR-ing
10-bit label masks, so this is the equivalence its parity rests on; checked
here exhaustively over all pairs of observations, on random triples and
random long sequences, against the oracle's branch restatement.
"""
import itertools
import random

from oracle import lib

NONE = 4


def obs_labels(type_, fwd, pb, nb):
    if type_ in (0, 4):
        return NONE, NONE
    if type_ == 5:
        type_ = 2
    if type_ == 1:
        return (NONE, nb) if fwd else (3 - nb, NONE)
    if type_ == 2:
        return (pb, NONE) if fwd else (NONE, 3 - pb)
    return (pb, nb) if fwd else (3 - nb, 3 - pb)


def state_of_sets(I, O):
    if not I and not O:
        return 0
    si = len(I) == 1 and NONE not in I
    so = len(O) == 1 and NONE not in O
    if si and so:
        return 16 + 4 * next(iter(I)) + next(iter(O))
    if so:
        return 4 + next(iter(O))
    if si:
        return 8 + next(iter(I))
    return 3


ALL_OBS = [(t, f, p, n) for t in range(6) for f in (0, 1) for p in range(4) for n in range(4)]


def fold(seq):
    s = 0
    for o in seq:
        s = lib().ref_dfa_step(s, *o)
    return s


def test_join_equivalence_exhaustive_up_to_three():
    for L in (1, 2):
        for seq in itertools.product(ALL_OBS, repeat=L):
            I, O = set(), set()
            for o in seq:
                a, b = obs_labels(*o)
                I.add(a)
                O.add(b)
            assert fold(seq) == state_of_sets(I, O), seq
    r = random.Random(0)
    for _ in range(20000):
        seq = [r.choice(ALL_OBS) for _ in range(3)]
        I = {obs_labels(*o)[0] for o in seq}
        O = {obs_labels(*o)[1] for o in seq}
        assert fold(seq) == state_of_sets(I, O)


def test_order_invariance_long_sequences():
    r = random.Random(1)
    for _ in range(3000):
        seq = [r.choice(ALL_OBS[:96]) for _ in range(r.randint(1, 12))]
        a = fold(seq)
        r.shuffle(seq)
        assert fold(seq) == a
