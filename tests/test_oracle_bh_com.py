"""The oracle's octree centres of mass (oracle/steps_oracle.c): with an
opening angle so large that every row takes the root, a row's force is
M_tot (C - x_i) / |C - x_i|^3 with C the centre of mass of all particles,
computed here in extended precision.  Both the reference's running averages
(forces.cc:164-170) and the sums from the children (com_from_children) give
it to rounding; the sums closer, as sums are."""
import numpy as np
import pytest

import oracle_lib as oracle


@pytest.mark.parametrize("N,seed", [(3000, 1), (20000, 2)])
def test_root_only_force_is_the_centre_of_mass_closed_form(N, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(N, 3)) * np.array([1.0, 2.0, 0.5]) + np.array([30.0, -20.0, 10.0])
    M = rng.uniform(0.5, 3.0, N)
    h = np.full(N, 1e-9)
    xl, Ml = x.astype(np.longdouble), M.astype(np.longdouble)
    C = (Ml[:, None] * xl).sum(axis=0) / Ml.sum()
    d = C[None, :] - xl
    r = np.sqrt((d * d).sum(axis=1))
    want = (Ml.sum() * d / r[:, None] ** 3).astype(np.float64)
    p = oracle.params(0)
    err = {}
    for sums, tol in ((False, 2e-11), (True, 5e-13)):
        F = oracle.forces_bh_ex(p, 1e9, 1e-9, x, M, h, com_from_children=sums)
        e = np.linalg.norm(F - want, axis=1) / np.linalg.norm(want, axis=1)
        assert e.max() <= tol, (sums, e.max())
        err[sums] = e.max()
    # the running averages' rounding grows with the insertions (N = 20000: 5e-12
    # against 1.3e-13 for the sums)
    assert err[True] * 5 < err[False]


def test_children_sums_keep_the_tree_and_the_walk():
    # at theta 0.5 on clustered particles the two differ only by rounding,
    # and at a tiny theta (no node taken) not at all
    rng = np.random.default_rng(4)
    x = rng.normal(size=(4000, 3)) * rng.uniform(0.1, 3.0, size=(4000, 1))
    M = rng.uniform(0.5, 2.0, 4000)
    h = np.full(4000, 1e-3)
    p = oracle.params(0)
    a = oracle.forces_bh_ex(p, 0.5, 1e-3, x, M, h)
    b = oracle.forces_bh_ex(p, 0.5, 1e-3, x, M, h, com_from_children=True)
    e = np.linalg.norm(a - b, axis=1) / np.linalg.norm(a, axis=1)
    assert 0 < e.max() <= 1e-12
    a = oracle.forces_bh_ex(p, 1e-12, 1e-3, x, M, h)
    b = oracle.forces_bh_ex(p, 1e-12, 1e-3, x, M, h, com_from_children=True)
    assert np.array_equal(a, b)
