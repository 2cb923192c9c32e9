"""The nearest-active-station rule of k_near_list / k_nearest, restated on
the CPU: the first active station of a cell's list of its K nearest
stations (ordered by distance, then by the rank in which a step compacts
its stations) is the station a scan of the step's compacted stations
picks (strict <, so ties go to the lower compacted index), whenever the
list holds an active station; and the kernel's unrolled insertion update
keeps the K smallest keys in stable order.  Integer coordinates make
exact distance ties common."""
import numpy as np


def _insert_update(vals, K):
    """k_near_list's insertion: slot k takes slot k-1's entry when the new
    key goes above it, else the new key when it goes above slot k."""
    d = [np.inf] * K
    ids = [-1] * K
    for g, v in enumerate(vals):
        if not v < d[K - 1]:
            continue
        for k in range(K - 1, 0, -1):
            dk, dkm, ik, ikm = d[k], d[k - 1], ids[k], ids[k - 1]
            up = v < dkm
            d[k] = dkm if up else (v if v < dk else dk)
            ids[k] = ikm if up else (g if v < dk else ik)
        if v < d[0]:
            d[0], ids[0] = v, g
    return ids


def test_insert_update_is_a_stable_partial_sort():
    rng = np.random.default_rng(1)
    for K in (1, 4, 8):
        for _ in range(1500):
            n = int(rng.integers(1, 40))
            vals = rng.integers(0, 10, n).astype(float)
            ref = sorted(range(n), key=lambda g: (vals[g], g))[:K]
            assert _insert_update(vals, K) == ref + [-1] * (K - len(ref))


def test_first_active_listed_station_is_the_scan_choice():
    rng = np.random.default_rng(0)
    listed = 0
    for _ in range(2000):
        N = int(rng.integers(5, 80))
        sx = rng.integers(-5, 6, N).astype(float)
        sy = rng.integers(-5, 6, N).astype(float)
        sorder = rng.permutation(N)
        x, y = (float(v) for v in rng.integers(-5, 6, 2))
        d2 = (x - sx) ** 2 + (y - sy) ** 2
        # stations staged in rank order: the list's keys are (d2, rank)
        lst = [int(sorder[i]) for i in
               _insert_update(d2[sorder], min(8, N)) if i >= 0]
        for _ in range(5):
            act = rng.random(N) < rng.uniform(0.05, 1.0)
            if not act.any():
                continue
            idx = [int(g) for g in sorder if act[g]]   # k_compact's order
            pos = {g: q for q, g in enumerate(idx)}
            bd, bj = np.inf, -1
            for q, g in enumerate(idx):
                if d2[g] < bd:
                    bd, bj = d2[g], q
            got = next((pos[g] for g in lst if act[g]), None)
            if got is not None:
                listed += 1
                assert got == bj
    assert listed > 5000
