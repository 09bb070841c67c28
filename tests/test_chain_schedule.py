"""The pipelined prior chain's schedule (CPU emulation of k_init_chain's indexing, slip_smc.hip): with L lanes per
slip component, lag = max(band + 1, ceil(ndof / (L - 1))) and a ring of W >= lag + 1 values per lane (a power of
two), every (particle, coordinate) update runs exactly once, in round k lag + d on lane k mod L, after every value
it reads was written and before the ring slot holding it is reused; particle 0 reads zeros from the ring of lane
L - 1.  Covers every lane count the library may choose (2 .. 32) on meshes from 12 to 625 DOF per component."""
import pytest


def lag_for(ndof, band, L):
    return max(band + 1, (ndof + L - 2) // (L - 1))


def emulate(ndof, band, P, L):
    lag = lag_for(ndof, band, L)
    W = 1
    while W < lag + 1:
        W *= 2
    rounds = (P - 1) * lag + ndof
    ring = {}  # (lane, slot) -> (particle, coordinate) last written
    done = set()
    for t in range(rounds):
        writes = []
        for l in range(L):
            kmax = t // lag
            k = kmax - (((kmax - l) % L) + L) % L
            d = t - k * lag
            if not (0 <= k < P and d < ndof):
                continue
            assert k % L == l and (k, d) not in done
            for dd in range(max(0, d - band), d):  # this particle's earlier coordinates (own ring)
                assert ring.get((l, dd % W)) == (k, dd)
            prev = (l + L - 1) % L
            for dd in range(d + 1, min(ndof, d + band + 1)):  # the previous particle's later coordinates
                assert ring.get((prev, dd % W)) == (None if k == 0 else (k - 1, dd))
            done.add((k, d))
            writes.append(((l, d % W), (k, d)))
        for key, val in writes:  # the round's writes land before the next round (barrier)
            ring[key] = val
    assert len(done) == P * ndof
    return lag, W


@pytest.mark.parametrize("ndof,band", [(12, 5), (25, 12), (40, 39), (81, 18), (240, 32), (561, 68), (625, 52)])
@pytest.mark.parametrize("L", [2, 4, 8, 16, 32])
def test_chain_schedule(ndof, band, L):
    lag, W = emulate(ndof, band, 2 * L + 3, L)
    assert (L - 1) * lag >= ndof and lag > band and W >= lag + 1
