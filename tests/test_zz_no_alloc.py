"""smcf_find_next_gamma_noalloc (the allocation-free form of smcf_find_next_gamma) against the allocating entry
point, through the C ABI: identical gamma, weights and evidence on two calls in a row, the second with fewer
particles than the first, so it also checks that the shared scratch left by a larger call does not leak into a
smaller one."""
import numpy as np
import pytest

import smc_inversion_fort_amd as S


def _likelihoods(P, seed):
    rng = np.random.default_rng(seed)
    # spread of a stage's likelihoods: a bulk plus a few far particles, as late stages have
    L = 50.0 + 30.0 * rng.standard_normal(P)
    L[rng.integers(0, P, size=max(1, P // 50))] += 400.0
    return L


@pytest.mark.gpu
def test_noalloc_matches_allocating_entry_point_twice():
    ctx = S.Context(0)
    try:
        for P, gprev, cv, tol, seed in ((3584, 0.0, 2.0, 1e-10, 11), (700, 0.37, 0.2, 1e-8, 12)):
            L = _likelihoods(P, seed)
            g_ref, w_ref, ev_ref = ctx.find_next_gamma(gprev, L, cv=cv, tol=tol)
            g, w, ev = ctx.find_next_gamma_noalloc(gprev, L, cv=cv, tol=tol)
            print("P=%d gamma %.17g (allocating %.17g) evidence %.17g (%.17g)" % (P, g, g_ref, ev, ev_ref),
                  flush=True)
            assert g == g_ref
            assert ev == ev_ref
            assert np.array_equal(w, w_ref)
            assert gprev < g <= 1.0
    finally:
        ctx.close()


@pytest.mark.gpu
def test_noalloc_rejects_sizes_beyond_its_scratch():
    ctx = S.Context(0)
    try:
        with pytest.raises(S.SmcfError, match="scratch maximum"):
            ctx.find_next_gamma_noalloc(0.0, np.zeros(3585))
    finally:
        ctx.close()
