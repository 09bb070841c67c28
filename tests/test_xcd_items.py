"""The stage engine's item lists (engine.cpp, xcd_items / xcd_items_bound): each instance's chunks go to the XCD
queue that is shortest at that point, and the padded list of any subset of a list's instances - the running ones,
rewritten as instances finish and join - stays within the region reserved for the list, 8 (ceil(T / 8) + c) items
for T chunks in all and c the most of one instance.  (Assigning the running instances round-robin instead can
outgrow the full list's own padded length; the second test keeps the counterexample.)"""
import random


def greedy_len(chunks):
    q = [0] * 8
    for z in chunks:
        x = min(range(8), key=lambda y: (q[y], y))
        q[x] += z
    return 8 * max(q) if chunks else 0


def round_robin_len(chunks):
    q = [0] * 8
    for g, z in enumerate(chunks):
        q[g % 8] += z
    return 8 * max(q) if chunks else 0


def bound(chunks):
    return 8 * ((sum(chunks) + 7) // 8 + max(chunks)) if chunks else 0


def test_greedy_subsets_stay_within_the_bound():
    rng = random.Random(1)
    for _ in range(3000):
        chunks = [rng.choice([1, 2, 3, 8]) for _ in range(rng.randint(1, 200))]
        B = bound(chunks)
        assert greedy_len(chunks) <= B
        for _ in range(4):
            keep = rng.random()
            sub = [z for z in chunks if rng.random() < keep]
            assert greedy_len(sub) <= B


def test_equal_chunks_are_round_robin():
    for n in range(1, 100):
        assert greedy_len([8] * n) == round_robin_len([8] * n)


def test_round_robin_subset_can_outgrow_the_full_list():
    full = [2, 2, 2, 2, 8, 8, 2, 2, 1, 2, 8, 1, 1, 1, 1, 8, 2, 2, 2, 1, 1, 1, 2, 2, 2, 1, 2, 8, 1]
    keep = [0, 4, 5, 6, 7, 8, 9, 10, 15, 16, 17, 18, 20, 22, 23, 24, 26, 27, 28]
    sub = [full[k] for k in keep]
    assert round_robin_len(sub) > round_robin_len(full)
    assert greedy_len(sub) <= bound(full)
