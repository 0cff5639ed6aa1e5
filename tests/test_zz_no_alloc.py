"""steps_force_eval_noalloc, the entry point that allocates no device memory,
against steps_force_eval: the same kernels, plan and reduction order, so the
forces are bitwise equal; two calls in a row, the second with a smaller input
(the static scratch keeps the first call's contents), for every kind of
device table it stages; an evaluation larger than its scratch is refused."""
import ctypes as C

import numpy as np
import pytest

from steps_amd import _abi
from steps_amd import workloads as wl

pytestmark = pytest.mark.gpu


def noalloc(steps, d, x, M, soft, id_min=0, id_max=None):
    lib = steps.load_library()
    x, M, soft = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, M, soft))
    N = M.size
    id_max = N - 1 if id_max is None else id_max
    nrows = max(0, id_max - id_min + 1)
    F = np.zeros(3 * max(nrows, 1))
    rc = lib.steps_force_eval_noalloc(C.byref(d), N, _abi._ptr(x.reshape(-1)), _abi._ptr(M), _abi._ptr(soft),
                                      _abi._ptr(F), id_min, id_max)
    if rc != 0:
        raise _abi.StepsError(rc, lib.steps_last_error().decode())
    return F[: 3 * nrows].reshape(nrows, 3)


def cases(steps, oracle):
    r3a, r3b = wl.r3_compactified(N=20000, seed=7), wl.r3_compactified(N=5000, seed=8)
    kw = dict(cosmology=1, comoving=1, mass_in_unit_sphere=r3a.mass_in_unit_sphere)
    out = []
    for prec in (steps.PREC_FP64, steps.PREC_MIXED):
        d, keep = steps.make_desc(steps.TOPO_R3, precision=prec, **kw)
        out.append((f"r3 prec={prec}", d, keep, r3a, r3b))
    d, keep = steps.make_desc(steps.TOPO_R3, bh_theta=0.5, bh_particle_radius=r3a.particle_radius, **kw)
    out.append(("r3 barnes-hut", d, keep, r3a, r3b))
    rtab = np.linspace(-1.0, 1.0, 32) * 1e-3 * r3a.mass_in_unit_sphere
    d, keep = steps.make_desc(steps.TOPO_R3, bh_theta=0.5, bh_particle_radius=r3a.particle_radius, Rsim=50.0,
                              bh_random=[0.3, 0.6, 0.2, 0.9, 0.4, 0.7], radial_bh_table=rtab, **kw)
    out.append(("r3 barnes-hut randomised, radial correction", d, keep, r3a, r3b))
    ta, tb = wl.t3_lattice(n_side=16, seed=1001), wl.t3_lattice(n_side=10, seed=1002)
    T3 = steps.t3_table(2, ta.L)
    d, keep = steps.make_desc(steps.TOPO_T3, is_periodic=2, L=ta.L, t3_table=T3)
    out.append(("t3 ewald", d, keep, ta, tb))
    ca = wl.s1r2_cylinder(n_omega=16, n_phi=8, n_z=16, seed=6)
    cb = wl.s1r2_cylinder(n_omega=8, n_phi=8, n_z=8, seed=9)
    Ts, shape = steps.s1r2_table(2, ca.L, ca.Rsim)
    base = dict(L=ca.L, Rsim=ca.Rsim, cosmology=1, comoving=1, mass_in_unit_sphere=ca.mass_in_unit_sphere)
    d, keep = steps.make_desc(steps.TOPO_S1R2, is_periodic=2, s1r2_lookup=1, interp_order=4, s1r2_table=Ts,
                              s1r2_shape=shape, **base)
    out.append(("s1r2 tsc", d, keep, ca, cb))
    radial = steps.radial_table(ca.Rsim, 0.5 * ca.L, 1000, 1000)
    d, keep = steps.make_desc(steps.TOPO_S1R2, is_periodic=1, radial_table=radial, **base)
    out.append(("s1r2 nearest radial", d, keep, ca, cb))
    return out


def test_noalloc_matches_force_eval_twice(steps, oracle):
    for name, d, keep, wa, wb in cases(steps, oracle):
        print(f"case noalloc {name} N={wa.N} then N={wb.N}", flush=True)
        for w in (wa, wb):
            Fa = noalloc(steps, d, w.x, w.M, w.soft)
            Fr = steps.force_eval(d, w.x, w.M, w.soft)
            assert np.array_equal(Fa, Fr), name
        Fs = noalloc(steps, d, wa.x, wa.M, wa.soft, id_min=3, id_max=wa.N // 2)
        assert np.array_equal(Fs, steps.force_eval(d, wa.x, wa.M, wa.soft, id_min=3, id_max=wa.N // 2)), name


def test_noalloc_refuses_what_does_not_fit(steps):
    print("case noalloc too large", flush=True)
    n = 1 << 22
    d, keep = steps.make_desc(steps.TOPO_R3)
    x = np.random.default_rng(1).normal(size=(n, 3))
    with pytest.raises(_abi.StepsError) as ei:
        noalloc(steps, d, x, np.ones(n), np.full(n, 0.01))
    assert ei.value.code == _abi.STEPS_ERR_ARG and "bytes of device scratch" in str(ei.value)
    w = wl.r3_compactified(N=3000, seed=3)                     # and it still works afterwards
    assert np.array_equal(noalloc(steps, d, w.x, w.M, w.soft), steps.force_eval(d, w.x, w.M, w.soft))
