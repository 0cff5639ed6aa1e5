"""Barnes-Hut octree forces on the device (the reference's USE_BH engine,
forces.cc:97-187, 373-506 R3, 582-772 T3, 881-1217 S1R2, CPU-only there)
against the oracle's restatement of the reference's pointer octree.  The device builds the same tree (cubes
holding two or more particles, children in octant order) and makes the same
accept / open decision per (row, node), so the forces agree to the rounding
of the centres of mass and of the sums: per-row relative error <= 1e-12,
for several opening angles, row ranges, the external-term modes and
particle counts from 1 up; and the octree with an opening angle so small
that no node is accepted is the direct sum.  The randomised octree
(RANDOMIZE_BH: rotated / shifted frame, moved root, forces turned back) and
the radial correction table are held to the same bound against the oracle,
in one-shot and resident evaluations with new draws per evaluation."""
import numpy as np
import pytest

from conftest import rel_err, report
from steps_amd import _abi
from steps_amd import workloads as wl

pytestmark = pytest.mark.gpu


def _desc(steps, w, theta, mode="comoving"):
    kw = {}
    if mode == "comoving":
        kw = dict(cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere)
    elif mode == "dark_energy":
        kw = dict(cosmology=1, comoving=0, H0=wl.H0, Omega_lambda=wl.OMEGA_L)
    return steps.make_desc(steps.TOPO_R3, bh_theta=theta, bh_particle_radius=w.particle_radius, **kw), kw


@pytest.mark.parametrize("N,seed,theta,mode", [
    (4096, 11, 0.5, "comoving"),
    (4096, 12, 0.8, "none"),
    (6000, 13, 0.25, "dark_energy"),
    (65536, 14, 0.5, "comoving"),
])
def test_bh_matches_oracle(steps, oracle, N, seed, theta, mode):
    print(f"case bh N={N} seed={seed} theta={theta} {mode}", flush=True)
    w = wl.r3_compactified(N=N, seed=seed)
    (d, keep), kw = _desc(steps, w, theta, mode)
    p = oracle.params(0, **kw)
    Fo = oracle.forces_bh(p, theta, w.particle_radius, w.x, w.M, w.soft)
    F = steps.force_eval(d, w.x, w.M, w.soft)
    e = rel_err(F, Fo)
    report(f"bh N={N} theta={theta}", e)
    assert e.max() <= 1e-12
    lo, hi = N // 3, N // 3 + 777  # a row range: rows outside are not written
    Fr = steps.force_eval(d, w.x, w.M, w.soft, id_min=lo, id_max=hi)
    assert np.array_equal(Fr, F[lo:hi + 1])


def test_bh_tiny_theta_is_the_direct_sum(steps):
    print("case bh theta 1e-12 vs direct sum", flush=True)
    w = wl.r3_compactified(N=3000, seed=15)
    (d, keep), kw = _desc(steps, w, 1e-12)
    dd, keep2 = steps.make_desc(steps.TOPO_R3, **kw)
    e = rel_err(steps.force_eval(d, w.x, w.M, w.soft), steps.force_eval(dd, w.x, w.M, w.soft))
    report("bh tiny theta vs direct", e)
    assert e.max() <= 1e-11


def test_bh_close_pairs_resolve_below_level_21(steps, oracle):
    """150 pairs placed 10^-12 .. 10^-6.5 of the root side apart (below the
    2^-21 cube a single 63-bit key resolves; 2^-40 at the closest): the device
    splits them level by level as the reference's pointer tree does, so every
    row agrees with the oracle."""
    print("case bh N=4096 with 150 close pairs", flush=True)
    w = wl.r3_compactified(N=4096, seed=18)
    x = w.x.copy()
    rng = np.random.default_rng(18)
    far = int(np.argmax(np.abs(x).max(axis=1)))  # keeps the root cube
    pick = rng.choice(np.delete(np.arange(w.N), far), 300, replace=False)
    a, b = pick[:150], pick[150:]
    root = 2.00002 * np.abs(x).max()
    sep = root * 10.0 ** rng.uniform(-12.0, -6.5, 150)
    u = rng.normal(size=(150, 3))
    x[b] = x[a] + u / np.linalg.norm(u, axis=1, keepdims=True) * sep[:, None]
    assert 2.00002 * np.abs(x).max() == root
    (d, keep), kw = _desc(steps, w, 0.5)
    Fo = oracle.forces_bh(oracle.params(0, **kw), 0.5, w.particle_radius, x, w.M, w.soft)
    F = steps.force_eval(d, x, w.M, w.soft)
    e = rel_err(F, Fo)
    report("bh close pairs", e)
    assert e.max() <= 1e-12


def test_bh_coincident_particles(steps):
    """Exact duplicates (the reference's insertion would recurse without end;
    the device keeps particles that share all 42 digits as leaves of one
    level-42 node): forces stay finite, and with an opening angle too small
    to accept any node they are the direct sum, whose softened kernel gives
    a coincident pair no force."""
    print("case bh N=3000 with coincident particles", flush=True)
    w = wl.r3_compactified(N=3000, seed=19)
    x = w.x.copy()
    x[1500:1540] = x[100:140]  # 40 duplicated particles
    x[1600:1610] = x[200]  # 11 particles at one point
    (d, keep), kw = _desc(steps, w, 1e-12)
    dd, keep2 = steps.make_desc(steps.TOPO_R3, **kw)
    F, Fd = steps.force_eval(d, x, w.M, w.soft), steps.force_eval(dd, x, w.M, w.soft)
    assert np.isfinite(F).all() and np.isfinite(Fd).all()
    e = rel_err(F, Fd)
    report("bh coincident, tiny theta vs direct", e)
    assert e.max() <= 1e-11
    (d5, keep5), kw5 = _desc(steps, w, 0.5)
    assert np.isfinite(steps.force_eval(d5, x, w.M, w.soft)).all()


def test_bh_small_counts(steps, oracle):
    for N in (1, 2, 3, 9, 64, 65, 257):
        print(f"case bh small N={N}", flush=True)
        w = wl.r3_compactified(N=max(N, 16), seed=16)
        x, M, h = w.x[:N].copy(), w.M[:N].copy(), w.soft[:N].copy()
        (d, keep), kw = _desc(steps, w, 0.6, "none")
        Fo = oracle.forces_bh(oracle.params(0), 0.6, w.particle_radius, x, M, h)
        F = steps.force_eval(d, x, M, h)
        if N == 1:
            assert np.array_equal(F, np.zeros((1, 3)))
        else:
            assert rel_err(F, Fo).max() <= 1e-12, N


def test_bh_resident_sim_and_repeatability(steps, oracle):
    print("case bh resident sim", flush=True)
    w = wl.r3_compactified(N=20000, seed=17)
    (d, keep), kw = _desc(steps, w, 0.5)
    F1 = steps.force_eval(d, w.x, w.M, w.soft)
    sim = steps.Sim(d, w.N)
    sim.set_state(w.x, w.M, w.soft, w.v)
    for _ in range(2):  # the scratch keeps the previous tree: a second evaluation is the same
        sim.forces()
        assert np.array_equal(sim.get_forces(), F1)
    assert sim.last_kernel_ms()[0] > 0
    with pytest.raises(_abi.StepsError) as ei:
        sim.set_precision(steps.PREC_MIXED)
    assert ei.value.code == _abi.STEPS_ERR_ARG
    sim.close()


def test_bh_refusals(steps):
    print("case bh refusals", flush=True)
    w = wl.r3_compactified(N=100, seed=18)
    bad = [steps.make_desc(steps.TOPO_R3, bh_theta=-0.5),
           steps.make_desc(steps.TOPO_R3, bh_theta=0.5, bh_particle_radius=-1.0),
           steps.make_desc(steps.TOPO_R3, bh_theta=0.5, precision=steps.PREC_MIXED),
           steps.make_desc(steps.TOPO_T3, is_periodic=1, L=10.0, bh_theta=float("nan"))]
    for d, keep in bad:
        with pytest.raises(_abi.StepsError) as ei:
            steps.force_eval(d, w.x, w.M, w.soft)
        assert ei.value.code == _abi.STEPS_ERR_ARG
    srv = steps.Server(streams=1)
    d, keep = steps.make_desc(steps.TOPO_R3, bh_theta=0.5, bh_particle_radius=w.particle_radius)
    with pytest.raises(_abi.StepsError) as ei:
        srv.add_sim(d, w.x, w.M, w.soft)
    assert ei.value.code == _abi.STEPS_ERR_ARG
    srv.close()


# ---- periodic topologies (forces.cc:582-772 T3, 881-1217 S1R2) ------------------

def _periodic_case(steps, oracle, mode, **extra):
    """(workload, device descriptor + keepalive, oracle params) of one periodic
    octree mode; extra goes to the device descriptor (bh_random, radial_bh_table)."""
    if mode.startswith("t3"):
        ip = 1 if mode == "t3_nearest" else 2
        w = wl.t3_lattice(n_side=16, seed=21, is_periodic=ip)
        order = {"t3_nearest": 2, "t3_ewald_o2": 2, "t3_ewald_o4": 4, "t3_ewald_o6": 6}[mode]
        T = steps.t3_table(2, w.L) if ip >= 2 else None
        d = steps.make_desc(steps.TOPO_T3, is_periodic=ip, L=w.L, t3_table=T, interp_order=order, bh_theta=0.5,
                            bh_particle_radius=w.particle_radius, **extra)
        p = oracle.params(1, is_periodic=ip, L=w.L, t3_table=T, interp_order=order)
        return w, d, p
    w = wl.s1r2_cylinder(n_omega=16, n_phi=16, n_z=16, seed=22, is_periodic=1 if mode == "s1r2_nearest" else 2)
    kw = dict(L=w.L, Rsim=w.Rsim, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere)
    bh = dict(bh_theta=0.5, bh_particle_radius=w.particle_radius, **extra)
    if mode == "s1r2_nearest":
        R = steps.radial_table(w.Rsim, 0.5 * w.L, 1000, 1000)
        d = steps.make_desc(steps.TOPO_S1R2, is_periodic=1, radial_table=R, **kw, **bh)
        p = oracle.params(2, is_periodic=1, radial_table=R, **kw)
    elif mode == "s1r2_images":
        R = steps.radial_table(w.Rsim, w.L * (2 + 1 - 0.4), 1000, 1000)
        d = steps.make_desc(steps.TOPO_S1R2, is_periodic=2, s1r2_lookup=0, radial_table=R, **kw, **bh)
        p = oracle.params(2, is_periodic=2, s1r2_lookup=0, cut_inclusive=1, radial_table=R, **kw)
    else:
        order = {"s1r2_ngp": 0, "s1r2_cic": 2, "s1r2_tsc": 4}[mode]
        T, shape = steps.s1r2_table(2, w.L, w.Rsim)
        d = steps.make_desc(steps.TOPO_S1R2, is_periodic=2, s1r2_table=T, s1r2_shape=shape, interp_order=order,
                            **kw, **bh)
        p = oracle.params(2, is_periodic=2, s1r2_table=T, s1r2_shape=shape, interp_order=order, **kw)
    return w, d, p


@pytest.mark.parametrize("mode", ["t3_nearest", "t3_ewald_o2", "t3_ewald_o4", "t3_ewald_o6", "s1r2_nearest",
                                  "s1r2_ngp", "s1r2_cic", "s1r2_tsc", "s1r2_images"])
def test_bh_periodic_matches_oracle(steps, oracle, mode):
    print(f"case bh periodic {mode}", flush=True)
    w, (d, keep), p = _periodic_case(steps, oracle, mode)
    Fo = oracle.forces_bh(p, 0.5, w.particle_radius, w.x, w.M, w.soft)
    F = steps.force_eval(d, w.x, w.M, w.soft)
    e = rel_err(F, Fo)
    report(f"bh {mode} N={w.N}", e)
    assert e.max() <= 1e-12
    lo, hi = 100, 900
    assert np.array_equal(steps.force_eval(d, w.x, w.M, w.soft, id_min=lo, id_max=hi), F[lo:hi + 1])


# ---- randomised octree (RANDOMIZE_BH) and the radial correction -----------------

_U = {1: [0.83, 0.12, 0.47, 0.66, 0.29, 0.91], 2: [0.05, 0.98, 0.51, 0.37, 0.74, 0.18]}


@pytest.mark.parametrize("mode,draw", [("r3", 1), ("r3", 2), ("t3_nearest", 1), ("t3_ewald_o4", 2),
                                       ("s1r2_nearest", 1), ("s1r2_cic", 2), ("s1r2_images", 1)])
def test_bh_randomised_matches_oracle(steps, oracle, mode, draw):
    print(f"case bh randomised {mode} draw {draw}", flush=True)
    u = _U[draw]
    if mode == "r3":
        w = wl.r3_compactified(N=8192, seed=31)
        (d, keep), kw = _desc(steps, w, 0.5)
        d.bh_randomize = 1
        for k in range(6):
            d.bh_random[k] = u[k]
        p = oracle.params(0, **kw)
    else:
        w, (d, keep), p = _periodic_case(steps, oracle, mode, bh_random=u)
    Fo = oracle.forces_bh_ex(p, 0.5, w.particle_radius, w.x, w.M, w.soft, u=u)
    F = steps.force_eval(d, w.x, w.M, w.soft)
    e = rel_err(F, Fo)
    report(f"bh randomised {mode} N={w.N}", e)
    # rounding of the centres of mass and sums, as in the plain octree; rows
    # whose pair sum nearly cancels the external term reach ~1e-12
    assert e.max() <= 5e-12 and np.percentile(e, 99) <= 1e-12
    # and it is not the plain octree
    assert not np.array_equal(F, oracle.forces_bh(p, 0.5, w.particle_radius, w.x, w.M, w.soft))


@pytest.mark.parametrize("mode,randomised", [("r3", False), ("r3", True), ("s1r2_nearest", False),
                                             ("s1r2_cic", True)])
def test_bh_radial_correction_matches_oracle(steps, oracle, mode, randomised):
    print(f"case bh radial correction {mode} randomised={randomised}", flush=True)
    u = _U[1] if randomised else None
    extra = dict(bh_random=u) if randomised else {}
    if mode == "r3":
        w = wl.r3_compactified(N=8192, seed=32)
        Rsim = float(np.sqrt((w.x ** 2).sum(axis=1)).max())
        kw = dict(cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere, Rsim=Rsim)
        tab = (np.linspace(-1.0, 2.0, 24) ** 3) * 1e-3 * w.mass_in_unit_sphere * Rsim
        d = steps.make_desc(steps.TOPO_R3, bh_theta=0.5, bh_particle_radius=w.particle_radius, radial_bh_table=tab,
                            **kw, **extra)
        p = oracle.params(0, **kw)
    else:
        w0 = wl.s1r2_cylinder(n_omega=16, n_phi=16, n_z=16, seed=22, is_periodic=1 if mode == "s1r2_nearest" else 2)
        tab = np.linspace(-1.0, 2.0, 24) * 1e-2 * w0.mass_in_unit_sphere * w0.Rsim
        w, d, p = _periodic_case(steps, oracle, mode, radial_bh_table=tab, **extra)
    d, keep = d
    Fo = oracle.forces_bh_ex(p, 0.5, w.particle_radius, w.x, w.M, w.soft, u=u, radial_bh=tab)
    F = steps.force_eval(d, w.x, w.M, w.soft)
    e = rel_err(F, Fo)
    report(f"bh radial correction {mode} randomised={randomised}", e)
    assert e.max() <= 5e-12 and np.percentile(e, 99) <= 1e-12
    assert not np.array_equal(F, oracle.forces_bh_ex(p, 0.5, w.particle_radius, w.x, w.M, w.soft, u=u))


def test_bh_randomised_resident_sim(steps, oracle):
    print("case bh randomised resident sim", flush=True)
    w = wl.r3_compactified(N=12000, seed=33)
    (d, keep), kw = _desc(steps, w, 0.6)
    p = oracle.params(0, **kw)
    d.bh_randomize = 1
    for k in range(6):
        d.bh_random[k] = _U[1][k]
    sim = steps.Sim(d, w.N)
    sim.set_state(w.x, w.M, w.soft, w.v)
    for draw in (1, 2, 1):  # new draws before each evaluation, as the reference's rand() per call
        sim.set_bh_random(_U[draw])
        sim.forces()
        Fo = oracle.forces_bh_ex(p, 0.6, w.particle_radius, w.x, w.M, w.soft, u=_U[draw])
        assert rel_err(sim.get_forces(), Fo).max() <= 5e-12, draw
    with pytest.raises(_abi.StepsError) as ei:
        sim.set_bh_random([1.5, 0, 0, 0, 0, 0])
    assert ei.value.code == _abi.STEPS_ERR_ARG
    sim.close()
    (d2, keep2), _ = _desc(steps, w, 0.6)
    plain = steps.Sim(d2, w.N)
    with pytest.raises(_abi.StepsError) as ei:
        plain.set_bh_random(_U[1])
    assert ei.value.code == _abi.STEPS_ERR_STATE
    plain.close()


def test_bh_randomised_and_radial_refusals(steps):
    print("case bh randomised/radial refusals", flush=True)
    w = wl.r3_compactified(N=100, seed=18)
    tab = np.ones(8)
    bad = [steps.make_desc(steps.TOPO_R3, bh_random=_U[1]),  # no octree
           steps.make_desc(steps.TOPO_R3, Rsim=10.0, radial_bh_table=tab),  # no octree
           steps.make_desc(steps.TOPO_R3, bh_theta=0.5, bh_random=[0.5, 0.5, 1.5, 0, 0, 0]),
           steps.make_desc(steps.TOPO_T3, is_periodic=1, L=10.0, Rsim=5.0, bh_theta=0.5, radial_bh_table=tab),
           steps.make_desc(steps.TOPO_R3, bh_theta=0.5, radial_bh_table=tab)]  # Rsim = 0
    for d, keep in bad:
        with pytest.raises(_abi.StepsError) as ei:
            steps.force_eval(d, w.x, w.M, w.soft)
        assert ei.value.code == _abi.STEPS_ERR_ARG
    d, keep = steps.make_desc(steps.TOPO_R3, bh_theta=0.5, Rsim=10.0, radial_bh_table=tab[:1])
    with pytest.raises(_abi.StepsError) as ei:
        steps.force_eval(d, w.x, w.M, w.soft)
    assert ei.value.code == _abi.STEPS_ERR_TABLE


@pytest.mark.parametrize("randomised", [False, True])
def test_bh_full_size_sampled_rows(steps, oracle, randomised):
    """The headline workload (N = 2^21, theta 0.5): the device octree over all
    rows against the oracle's pointer octree on three row ranges (the oracle
    builds the whole tree for each range)."""
    print(f"case bh full size N=2^21 randomised={randomised}", flush=True)
    w = wl.r3_compactified(N=2 ** 21, seed=1003)
    (d, keep), kw = _desc(steps, w, 0.5)
    u = _U[2] if randomised else None
    if randomised:
        d.bh_randomize = 1
        for k in range(6):
            d.bh_random[k] = u[k]
    sim = steps.Sim(d, w.N)
    sim.set_state(w.x, w.M, w.soft, w.v)
    sim.forces()
    F = sim.get_forces()
    sim.close()
    p = oracle.params(0, **kw)
    errs = []
    for lo in (0, w.N // 2 - 256, w.N - 512):
        Fo = oracle.forces_bh_ex(p, 0.5, w.particle_radius, w.x, w.M, w.soft, u=u, id_min=lo, id_max=lo + 511)
        errs.append(rel_err(F[lo:lo + 512], Fo))
    e = np.concatenate(errs)
    report(f"bh full size randomised={randomised} (1536 rows)", e)
    assert e.max() <= 5e-12 and np.median(e) <= 1e-14


@pytest.mark.parametrize("mode", ["t3_ewald_o4", "s1r2_cic"])
def test_bh_periodic_full_size_sampled_rows(steps, oracle, mode):
    """The periodic measurement configurations at theta 0.5 (T3 Ewald with
    Catmull-Rom table interpolation on the 64^3 lattice; S1xR2 CIC on the
    N = 524288 cylinder): the device octree over all rows against the
    oracle's pointer octree on three 512-row ranges."""
    print(f"case bh periodic full size {mode}", flush=True)
    if mode == "t3_ewald_o4":
        w = wl.t3_lattice()
        T = steps.t3_table(2, w.L)
        d, keep = steps.make_desc(steps.TOPO_T3, is_periodic=2, L=w.L, t3_table=T, interp_order=4, bh_theta=0.5,
                                  bh_particle_radius=w.particle_radius)
        p = oracle.params(1, is_periodic=2, L=w.L, t3_table=T, interp_order=4)
    else:
        w = wl.s1r2_cylinder()
        T, shape = steps.s1r2_table(2, w.L, w.Rsim)
        kw = dict(L=w.L, Rsim=w.Rsim, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere)
        d, keep = steps.make_desc(steps.TOPO_S1R2, is_periodic=2, s1r2_table=T, s1r2_shape=shape, interp_order=2,
                                  bh_theta=0.5, bh_particle_radius=w.particle_radius, **kw)
        p = oracle.params(2, is_periodic=2, s1r2_table=T, s1r2_shape=shape, interp_order=2, **kw)
    sim = steps.Sim(d, w.N)
    sim.set_state(w.x, w.M, w.soft, w.v)
    sim.forces()
    F = sim.get_forces()
    sim.close()
    got, ref, sums = [], [], []
    for lo in (0, w.N // 2 - 256, w.N - 512):
        got.append(F[lo:lo + 512])
        ref.append(oracle.forces_bh_ex(p, 0.5, w.particle_radius, w.x, w.M, w.soft, id_min=lo, id_max=lo + 511))
        sums.append(oracle.forces_bh_ex(p, 0.5, w.particle_radius, w.x, w.M, w.soft, id_min=lo, id_max=lo + 511,
                                        com_from_children=True))
    got, ref, sums = np.concatenate(got), np.concatenate(ref), np.concatenate(sums)
    e, es, gap = rel_err(got, ref), rel_err(got, sums), rel_err(sums, ref)
    report(f"bh periodic full size {mode} N={w.N} (1536 rows)", e)
    report(f"bh periodic full size {mode}: against centres of mass summed from children", es)
    report(f"bh periodic full size {mode}: the oracle's running averages against those sums", gap)
    # The device sums each node's centre of mass from its children; the
    # reference keeps running averages over thousands of insertions, whose
    # rounding a lattice row's cancelling pulls magnify to ~3e-11 at 64^3.
    # Against the same tree with summed centres of mass the device agrees to
    # rounding, and its distance from the reference's arithmetic is that gap.
    assert es.max() <= 1e-12 and np.median(e) <= 1e-12
    assert np.all(e <= gap + es + 1e-15)
