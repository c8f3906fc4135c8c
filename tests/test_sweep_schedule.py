"""The triplet schedules of the adjoint sweeps cover every unordered triplet
of an atom exactly once.

The lane-to-partner schedules of k_adjoint (meamzilla_amd/csrc/mz_adjoint.hip,
phase C) are restated here index for index and enumerated for every
neighbour count up to 130 (the workloads have 24 to 58): the union of the
pairs {j, k} the lanes visit must be all C(n, 2) of them, each once, which is
what makes the sweep's f-value and g-coefficient adjoints equal the sum over
triplets of the reference's triplet loop (src/pots/pot_meam_spline/
pot_meam_spline.cpp:240-269).  A schedule that drops or repeats one pair
fails here before any GPU run.
"""
import itertools

import pytest


def single_pair_schedule(n):
    """Lane l (pair l of the atom) visits l+1 .. l+(n-1)//2 (mod n), plus
    l+n/2 for the first n/2 lanes when n is even."""
    visits = []
    for l in range(n):
        trips = (n - 1) // 2 + (1 if n % 2 == 0 and l < n // 2 else 0)
        for m in range(trips):
            visits.append((l, (l + 1 + m) % n))
    return visits


def couples_schedule(n):
    """Couples c = local pairs (2c, 2c+1); both lanes hold both pairs.  The
    even lane first takes its couple's own triplet; with n odd the odd lane
    first visits the last pair; then lane e of couple c visits pair 2c'+e of
    couples c' = c+1 .. c+(mc-1)//2 (mod mc), plus c+mc/2 for the first mc/2
    couples when mc is even.  Each visit of partner k serves both own pairs."""
    visits = []
    mc = n // 2
    for l in range(2 * mc):
        c, e = l >> 1, l & 1
        if e == 0:
            visits.append((2 * c, 2 * c + 1))
        odd = e == 1 and n % 2 == 1
        trips = (mc - 1) // 2 + (1 if mc % 2 == 0 and c < mc // 2 else 0) + (1 if odd else 0)
        cn = c if odd else (0 if c + 1 == mc else c + 1)
        k = n - 1 if odd else 2 * cn + e
        for _ in range(trips):
            visits.append((2 * c, k))
            visits.append((2 * c + 1, k))
            cn = 0 if cn + 1 == mc else cn + 1
            k = 2 * cn + e
    return visits


@pytest.mark.parametrize("schedule", [single_pair_schedule, couples_schedule])
def test_every_triplet_once(schedule):
    for n in range(0, 131):
        seen = [frozenset(v) for v in schedule(n)]
        assert all(len(s) == 2 for s in seen), (n, "a pair visits itself")
        assert len(seen) == n * (n - 1) // 2, (n, len(seen))
        assert set(seen) == {frozenset(p) for p in itertools.combinations(range(n), 2)}, n


def test_couples_trip_balance():
    """Within an atom the lanes' step counts differ by at most two (the extra
    round-robin visit of the first half of the couples when their number is
    even, and the even lane's own-couple triplet), and a step serves two
    triplets, so the couples schedule takes about half the single-pair
    steps."""
    for n in range(4, 131):
        mc = n // 2
        steps = []
        for l in range(2 * mc):
            c, e = l >> 1, l & 1
            odd = e == 1 and n % 2 == 1
            steps.append((mc - 1) // 2 + (1 if mc % 2 == 0 and c < mc // 2 else 0) + (1 if odd or e == 0 else 0))
        assert max(steps) - min(steps) <= 2, n
        single = (n - 1) // 2 + (1 if n % 2 == 0 else 0)
        assert max(steps) <= single // 2 + 2, (n, max(steps), single)
