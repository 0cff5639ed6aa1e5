"""GPU parity of the periodic force variants and of the device-built tables.

T3 (forces_cuda.cu:567-645) and S1R2 (:652-864) forces are compared with the
oracle on the same inputs and the same tables (median relative error <= 1e-12);
the tables the device builds are compared with the oracle's tables.
"""
import numpy as np
import pytest

from conftest import rel_err, report
from steps_amd import workloads as wl

pytestmark = pytest.mark.gpu

L_T3 = 100.0 / 0.6766


@pytest.fixture(scope="module")
def t3_tab(oracle):
    return oracle.t3_table(2, L_T3)


def test_device_t3_table_matches_oracle(steps, t3_tab):
    print("case t3 table is_periodic=2", flush=True)
    T = steps.t3_table(2, L_T3)
    assert T.shape == t3_tab.shape
    scale = np.abs(t3_tab).max()
    d = np.abs(T - t3_tab).max() / scale
    print(f"[T3 table device vs oracle] max |dT|/max|T| = {d:.3e}", flush=True)
    assert d <= 1e-12


def test_device_s1r2_table_matches_oracle(steps, oracle):
    print("case s1r2 table is_periodic=2 Lz=200 Rsim=1000", flush=True)
    T, shape = steps.s1r2_table(2, 200.0, 1000.0)
    To, shape_o = oracle.s1r2_table(2, 200.0, 1000.0)
    assert shape == shape_o == (1440, 128)
    d = np.abs(T - To).max() / np.abs(To).max()
    print(f"[S1R2 table device vs oracle] max |dT|/max|T| = {d:.3e}", flush=True)
    # device and host libm (erfc/exp/sin/cos) differ by ulps; the table is a
    # difference of near-equal forces, which amplifies that to ~1e-12
    assert d <= 1e-10


def test_device_radial_table_matches_oracle(steps, oracle):
    for (R, Lz, size, acc) in [(1000.0, 100.0, 1000, 1000), (500.0, 260.0, 500, 7500)]:
        T = steps.radial_table(R, Lz, size, acc)
        To = oracle.radial_table(R, Lz, size, acc)
        assert np.allclose(T, To, rtol=1e-12, atol=0), np.abs(T / To - 1).max()


@pytest.mark.parametrize("is_periodic", [2, 1])
def test_t3_forces(steps, oracle, t3_tab, is_periodic):
    print(f"case t3 forces n_side=16 is_periodic={is_periodic}", flush=True)
    w = wl.t3_lattice(n_side=16, seed=1001, is_periodic=is_periodic)
    tab = t3_tab if is_periodic >= 2 else None
    d, keep = steps.make_desc(steps.TOPO_T3, is_periodic=is_periodic, L=w.L, cosmology=1, comoving=1,
                              t3_table=tab)
    F = steps.force_eval(d, w.x, w.M, w.soft)
    p = oracle.params(1, is_periodic=is_periodic, L=w.L, cosmology=1, comoving=1, t3_table=tab, interp_order=4)
    Fo = oracle.forces(p, w.x, w.M, w.soft)
    e = rel_err(F, Fo)
    report(f"T3 is_periodic={is_periodic} N={w.N}", e)
    assert np.median(e) <= 1e-12
    assert np.percentile(e, 99) <= 1e-9


def test_t3_row_subset_and_shuffled_input(steps, oracle, t3_tab):
    """A row subset agrees with the oracle and with the all-rows evaluation,
    and the result does not depend on the order of the input particles beyond
    summation order."""
    print("case t3 row subset / shuffled input", flush=True)
    w = wl.t3_lattice(n_side=16, seed=1001)
    d, keep = steps.make_desc(steps.TOPO_T3, is_periodic=2, L=w.L, t3_table=t3_tab)
    F = steps.force_eval(d, w.x, w.M, w.soft)
    Fs = steps.force_eval(d, w.x, w.M, w.soft, id_min=100, id_max=1699)
    p = oracle.params(1, is_periodic=2, L=w.L, t3_table=t3_tab, interp_order=4)
    Fo = oracle.forces(p, w.x, w.M, w.soft)
    report("T3 row subset vs oracle", rel_err(Fs, Fo[100:1700]))
    assert np.median(rel_err(Fs, Fo[100:1700])) <= 1e-12
    e = rel_err(Fs, F[100:1700])
    report("T3 row subset vs all rows", e)
    assert np.median(e) <= 1e-12 and np.percentile(e, 99) <= 1e-10
    perm = np.random.default_rng(7).permutation(w.N)
    Fp = steps.force_eval(d, w.x[perm], w.M[perm], w.soft[perm])
    e = rel_err(Fp, F[perm])
    report("T3 shuffled input vs original order", e)
    assert np.median(e) <= 1e-12 and np.percentile(e, 99) <= 1e-10


@pytest.mark.parametrize("is_periodic", [2, 1])
def test_t3_mixed_precision(steps, oracle, t3_tab, is_periodic):
    """Mixed-precision T3 (tricubic Ewald correction from an FP32 table, or
    the nearest image only) at N = 48^3: every row against the FP64 path,
    sampled rows against the oracle (pair-sum bar p99 <= 1e-5; T3 has no
    external term); deterministic; Ewald: a row subset, which takes the kernel
    without the LDS window, bitwise equal to the full evaluation (window)."""
    print(f"case t3 mixed is_periodic={is_periodic} n_side=48", flush=True)
    w = wl.t3_lattice(n_side=48, seed=1001, is_periodic=is_periodic)
    tab = t3_tab if is_periodic >= 2 else None
    kw = dict(is_periodic=is_periodic, L=w.L, t3_table=tab)
    d64, k64 = steps.make_desc(steps.TOPO_T3, **kw)
    dmx, kmx = steps.make_desc(steps.TOPO_T3, precision=steps.PREC_MIXED, **kw)
    F64 = steps.force_eval(d64, w.x, w.M, w.soft)
    Fmx = steps.force_eval(dmx, w.x, w.M, w.soft)
    assert not np.array_equal(Fmx, F64)                 # the FP32 path really ran
    e = rel_err(Fmx, F64)
    report(f"T3 mixed is_periodic={is_periodic} N={w.N} vs fp64", e)
    assert np.median(e) <= 2e-6 and np.percentile(e, 99) <= 1e-5
    rows = np.linspace(0, w.N - 1, 32).astype(np.int32)
    p = oracle.params(1, is_periodic=is_periodic, L=w.L, t3_table=tab, interp_order=4)
    eo = rel_err(Fmx[rows], oracle.forces_rows(p, w.x, w.M, w.soft, rows))
    report(f"T3 mixed is_periodic={is_periodic} sampled rows vs oracle", eo)
    assert np.median(eo) <= 2e-6
    assert np.array_equal(steps.force_eval(dmx, w.x, w.M, w.soft), Fmx)
    if is_periodic >= 2:
        Fs = steps.force_eval(dmx, w.x, w.M, w.soft, id_min=37, id_max=w.N - 1)
        assert np.array_equal(Fs, Fmx[37:])


def _foldable(T, ng):
    """The library's fold rule: exact antisymmetry in x and in y, except that
    components vanishing by symmetry may hold noise below 2^-52 max|T|, and
    the plane a mirror maps onto itself is exempt from the negated check."""
    T4 = T.reshape(ng, ng, ng, 3)
    eps0 = np.ldexp(np.abs(T4).max(), -52)

    def same(b, a):
        return (b == a) | ((np.abs(a) <= eps0) & (np.abs(b) <= eps0))

    centre = np.zeros(ng, bool)
    if ng % 2:
        centre[ng // 2] = True
    mx, my = T4[::-1], T4[:, ::-1]
    okx = ((same(mx[..., 0], -T4[..., 0]) | centre[:, None, None]) & same(mx[..., 1], T4[..., 1])
           & same(mx[..., 2], T4[..., 2]))
    oky = (same(my[..., 0], T4[..., 0]) & (same(my[..., 1], -T4[..., 1]) | centre[None, :, None])
           & same(my[..., 2], T4[..., 2]))
    return bool(okx.all() and oky.all())


def test_t3_tables_take_the_folded_layout(steps, t3_tab):
    """The oracle's table and the device-built one are antisymmetric in x and
    y to the library's rule, so the pair kernel reads them folded."""
    ng = round((t3_tab.size // 3) ** (1 / 3))
    assert _foldable(t3_tab, ng)
    assert _foldable(steps.t3_table(2, L_T3), ng)


def test_t3_unfolded_layout(steps, oracle, t3_tab):
    """Tables without the symmetry are read unfolded.  (1) One entry that
    vanishes by symmetry raised just above the noise bound: forces agree with
    the folded evaluation of the original table to rounding.  (2) The table
    perturbed by 1e-3 max|T| noise: forces agree with the oracle on that
    same table."""
    print("case t3 unfolded layout n_side=16", flush=True)
    ng = round((t3_tab.size // 3) ** (1 / 3))
    w = wl.t3_lattice(n_side=16, seed=1001)
    d0, keep0 = steps.make_desc(steps.TOPO_T3, is_periodic=2, L=w.L, t3_table=t3_tab)
    F0 = steps.force_eval(d0, w.x, w.M, w.soft)
    T1 = t3_tab.copy().reshape(ng, ng, ng, 3)
    T1[ng // 2 - 10, ng // 2, ng // 2, 1] = 4.0 * np.ldexp(np.abs(t3_tab).max(), -52)
    T1 = T1.reshape(-1)
    assert not _foldable(T1, ng)
    d1, keep1 = steps.make_desc(steps.TOPO_T3, is_periodic=2, L=w.L, t3_table=T1)
    F1 = steps.force_eval(d1, w.x, w.M, w.soft)
    e = rel_err(F1, F0)
    report("T3 unfolded (noise entry) vs folded", e)
    assert np.median(e) <= 1e-14 and e.max() <= 1e-12
    T2 = t3_tab + 1e-3 * np.abs(t3_tab).max() * np.random.default_rng(11).normal(size=t3_tab.size)
    assert not _foldable(T2, ng)
    d2, keep2 = steps.make_desc(steps.TOPO_T3, is_periodic=2, L=w.L, t3_table=T2)
    F2 = steps.force_eval(d2, w.x, w.M, w.soft)
    p = oracle.params(1, is_periodic=2, L=w.L, t3_table=T2, interp_order=4)
    e = rel_err(F2, oracle.forces(p, w.x, w.M, w.soft))
    report("T3 perturbed table (unfolded) vs oracle", e)
    assert np.median(e) <= 1e-12 and np.percentile(e, 99) <= 1e-9


def _antisymmetric_table(ng, seed):
    """A random table with the reference table's symmetry, exactly: Tx odd
    in x and even in y, Ty even in x and odd in y, Tz even in both."""
    R = np.random.default_rng(seed).normal(size=(ng, ng, ng, 3)) * 1e-4
    A, B, C = R[..., 0], R[..., 1], R[..., 2]
    A = A - A[::-1]
    A = A + A[:, ::-1]
    B = B + B[::-1]
    B = B - B[:, ::-1]
    C = C + C[::-1]
    C = C + C[:, ::-1]
    return np.ascontiguousarray(np.stack([A, B, C], axis=-1).reshape(-1))


@pytest.mark.parametrize("ng", [16, 15])
def test_t3_synthetic_tables_even_and_odd_grid(steps, oracle, ng):
    """Folded tables of even and odd Ng (the table window's node range
    differs between the two) and an unfolded one, against the oracle."""
    print(f"case t3 synthetic table ng={ng}", flush=True)
    w = wl.t3_lattice(n_side=16, seed=1001)
    T = _antisymmetric_table(ng, 17 + ng)
    assert _foldable(T, ng)
    Tr = np.random.default_rng(ng).normal(size=ng ** 3 * 3) * 1e-4
    assert not _foldable(Tr, ng)
    for name, tab in (("folded", T), ("unfolded", Tr)):
        d, keep = steps.make_desc(steps.TOPO_T3, is_periodic=2, L=w.L, t3_table=tab)
        F = steps.force_eval(d, w.x, w.M, w.soft)
        p = oracle.params(1, is_periodic=2, L=w.L, t3_table=tab, interp_order=4)
        e = rel_err(F, oracle.forces(p, w.x, w.M, w.soft))
        report(f"T3 ng={ng} {name} table", e)
        assert np.median(e) <= 1e-12 and np.percentile(e, 99) <= 1e-9


def test_t3_clustered_and_outside_box_input(steps, oracle, t3_tab):
    """Inputs the pencil ordering and the table window treat specially: a
    tight cluster plus scattered particles (many pencils of one or two
    particles, so windows change at almost every source), positions shifted
    by whole box lengths (periodic pencil bins), and N below one workgroup."""
    print("case t3 clustered / outside-box / small N", flush=True)
    L = L_T3
    rng = np.random.default_rng(23)
    n = 3000
    x = np.concatenate([L * (0.3 + 0.02 * rng.normal(size=(2400, 3))), rng.uniform(0, L, size=(n - 2400, 3))])
    x = np.mod(x, L)
    M = np.full(n, 1.0)
    h = np.full(n, 0.01)
    for label, xx in (("clustered", x), ("shifted by whole boxes", x + L * rng.integers(-2, 3, size=(n, 3)))):
        d, keep = steps.make_desc(steps.TOPO_T3, is_periodic=2, L=L, t3_table=t3_tab)
        F = steps.force_eval(d, xx, M, h)
        p = oracle.params(1, is_periodic=2, L=L, t3_table=t3_tab, interp_order=4)
        e = rel_err(F, oracle.forces(p, xx, M, h))
        report(f"T3 {label}", e)
        assert np.median(e) <= 1e-12 and np.percentile(e, 99) <= 1e-9
    xs = x[:100]
    d, keep = steps.make_desc(steps.TOPO_T3, is_periodic=2, L=L, t3_table=t3_tab)
    F = steps.force_eval(d, xs, M[:100], h[:100])
    p = oracle.params(1, is_periodic=2, L=L, t3_table=t3_tab, interp_order=4)
    e = rel_err(F, oracle.forces(p, xs, M[:100], h[:100]))
    report("T3 N=100", e)
    assert np.median(e) <= 1e-12


@pytest.mark.parametrize("mode", ["t3", "s1r2", "s1r2_mixed"])
def test_nearest_image_far_tiles(steps, oracle, mode):
    """Nearest-image modes (IS_PERIODIC = 1) at N >= 65536, where rows and
    sources are Morton-ordered and tiles that no lane can soften and that no
    lane sees across a wrap take the R3 loop with per-lane shifts: sampled rows
    against the oracle, and a row subset (input order, so other waves and
    other far-tile decisions) bitwise equal to the full evaluation."""
    print(f"case nearest image far tiles {mode}", flush=True)
    if mode == "t3":
        w = wl.t3_lattice(n_side=48, seed=1001, is_periodic=1)
        kw = dict(is_periodic=1, L=w.L, cosmology=1, comoving=1)
        d, keep = steps.make_desc(steps.TOPO_T3, **kw)
        p = oracle.params(1, interp_order=4, **kw)
    else:
        w = wl.s1r2_cylinder(n_omega=64, n_phi=32, n_z=32, seed=1002)
        radial = oracle.radial_table(w.Rsim, 0.5 * w.L, 1000, 1000)    # main.cc:1302
        kw = dict(is_periodic=1, L=w.L, Rsim=w.Rsim, cosmology=1, comoving=1,
                  mass_in_unit_sphere=w.mass_in_unit_sphere, radial_table=radial)
        prec = steps.PREC_MIXED if mode == "s1r2_mixed" else steps.PREC_FP64
        d, keep = steps.make_desc(steps.TOPO_S1R2, precision=prec, **kw)
        p = oracle.params(2, cut_inclusive=0, **kw)
    assert w.N >= 65536
    F = steps.force_eval(d, w.x, w.M, w.soft)
    rows = np.linspace(0, w.N - 1, 64).astype(np.int32)
    e = rel_err(F[rows], oracle.forces_rows(p, w.x, w.M, w.soft, rows))
    report(f"{mode} nearest N={w.N} sampled", e)
    if mode == "s1r2_mixed":
        # FP32 pairs against the FP64 evaluation of every row (bar of the
        # mixed path with the comoving term: p99 3e-5)
        d64, keep64 = steps.make_desc(steps.TOPO_S1R2, **kw)
        ea = rel_err(F, steps.force_eval(d64, w.x, w.M, w.soft))
        report(f"{mode} nearest N={w.N} all rows vs fp64", ea)
        assert np.median(e) <= 2e-6 and np.percentile(ea, 99) <= 3e-5 and np.median(ea) <= 2e-6
    else:
        assert np.median(e) <= 1e-12 and e.max() <= 1e-10
    Fs = steps.force_eval(d, w.x, w.M, w.soft, id_min=37, id_max=w.N - 1)
    assert np.array_equal(Fs, F[37:])


def test_t3_full_size_sampled_rows(steps, oracle, t3_tab):
    """The 64^3 T3 configuration at full size, checked on 96 sampled rows."""
    print("case t3 forces n_side=64 (N=262144) sampled rows", flush=True)
    w = wl.t3_lattice(n_side=64, seed=1001)
    d, keep = steps.make_desc(steps.TOPO_T3, is_periodic=2, L=w.L, t3_table=t3_tab)
    sim = steps.Sim(d, w.N)
    sim.set_state(w.x, w.M, w.soft)
    sim.forces()
    F = sim.get_forces()
    ms = sim.last_kernel_ms()
    print(f"[T3 N=262144] pair kernel {ms[0]:.1f} ms -> {w.N ** 2 / ms[0] / 1e6:.1f} G pair/s", flush=True)
    sim.close()
    rows = np.linspace(0, w.N - 1, 96).astype(np.int32)
    p = oracle.params(1, is_periodic=2, L=w.L, t3_table=t3_tab, interp_order=4)
    Fo = oracle.forces_rows(p, w.x, w.M, w.soft, rows)
    e = rel_err(F[rows], Fo)
    report("T3 N=262144 sampled", e)
    assert np.median(e) <= 1e-12
    # no external term and an antisymmetric correction: sum_i m_i F_i = 0
    mF = w.M[:, None] * F
    net = np.abs(mF.sum(axis=0)).max() / np.abs(mF).sum(axis=0).max()
    print(f"[T3 N=262144] |sum m F| / sum |m F| = {net:.3e}", flush=True)
    assert net <= 1e-12


@pytest.fixture(scope="module")
def cyl(oracle):
    w = wl.s1r2_cylinder(n_omega=32, n_phi=8, n_z=16, seed=1002)
    T, shape = oracle.s1r2_table(2, w.L, w.Rsim)
    return w, T, shape


@pytest.mark.parametrize("order", [0, 2, 4])
def test_s1r2_lookup_forces(steps, oracle, cyl, order):
    print(f"case s1r2 lookup order={order}", flush=True)
    w, T, shape = cyl
    kw = dict(is_periodic=2, L=w.L, Rsim=w.Rsim, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere,
              s1r2_lookup=1, interp_order=order, s1r2_table=T, s1r2_shape=shape)
    d, keep = steps.make_desc(steps.TOPO_S1R2, **kw)
    F = steps.force_eval(d, w.x, w.M, w.soft)
    Fo = oracle.forces(oracle.params(2, **kw), w.x, w.M, w.soft)
    e = rel_err(F, Fo)
    report(f"S1R2 lookup order={order} N={w.N}", e)
    assert np.median(e) <= 1e-12
    assert np.percentile(e, 99) <= 1e-9


@pytest.mark.parametrize("order", [2, 4])
def test_s1r2_table_window_matches_global_reads(steps, oracle, cyl, order):
    """The CIC / TSC kernels stage the rho rows of each source tile in LDS
    when the workgroup's pairs fit a window of rows, else read the global
    table.  Input order (windows fit), a row subset (other workgroups, other
    windows) and shuffled rows and sources (windows mostly do not fit) give the
    same values: bitwise for the subset, to summation order for the shuffle;
    and the oracle agrees on sampled rows."""
    print(f"case s1r2 table window order={order}", flush=True)
    _, T, shape = cyl
    w = wl.s1r2_cylinder(n_omega=64, n_phi=16, n_z=32, seed=1002)
    kw = dict(is_periodic=2, L=w.L, Rsim=w.Rsim, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere,
              s1r2_lookup=1, interp_order=order, s1r2_table=T, s1r2_shape=shape)
    d, keep = steps.make_desc(steps.TOPO_S1R2, **kw)
    F = steps.force_eval(d, w.x, w.M, w.soft)
    # a subset that keeps the row-block count keeps the source splits, hence
    # the summation order
    Fs = steps.force_eval(d, w.x, w.M, w.soft, id_min=37, id_max=w.N - 1)
    report(f"S1R2 window order={order} row subset vs all rows", rel_err(Fs, F[37:]))
    assert np.array_equal(Fs, F[37:])
    rows = np.linspace(0, w.N - 1, 48).astype(np.int32)
    e = rel_err(F[rows], oracle.forces_rows(oracle.params(2, **kw), w.x, w.M, w.soft, rows))
    report(f"S1R2 window order={order} N={w.N} sampled rows", e)
    assert np.median(e) <= 1e-12 and e.max() <= 1e-9
    perm = np.random.default_rng(11).permutation(w.N)
    Fp = steps.force_eval(d, w.x[perm], w.M[perm], w.soft[perm])
    ep = rel_err(Fp, F[perm])
    report(f"S1R2 window order={order} shuffled vs input order", ep)
    assert np.median(ep) <= 1e-12 and np.percentile(ep, 99) <= 1e-10


@pytest.mark.parametrize("mode", ["nearest", "images"])
def test_s1r2_radial_corrected_forces(steps, oracle, cyl, mode):
    print(f"case s1r2 {mode}", flush=True)
    w, _, _ = cyl
    if mode == "nearest":
        is_p, lookup = 1, 1
        radial = oracle.radial_table(w.Rsim, 0.5 * w.L, 1000, 1000)    # main.cc:1302
    else:
        is_p, lookup = 2, 0
        radial = oracle.radial_table(w.Rsim, w.L * (is_p + 1 - 0.4), 1000, 1000)  # main.cc:1282
    kw = dict(is_periodic=is_p, L=w.L, Rsim=w.Rsim, cosmology=1, comoving=1,
              mass_in_unit_sphere=w.mass_in_unit_sphere, s1r2_lookup=lookup, radial_table=radial)
    d, keep = steps.make_desc(steps.TOPO_S1R2, **kw)
    F = steps.force_eval(d, w.x, w.M, w.soft)
    Fo = oracle.forces(oracle.params(2, cut_inclusive=0, **kw), w.x, w.M, w.soft)
    e = rel_err(F, Fo)
    report(f"S1R2 {mode} N={w.N}", e)
    assert np.median(e) <= 1e-12


def test_s1r2_full_size_sampled_rows(steps, oracle):
    """The 524288-particle S1R2 configuration (CIC and TSC), 64 sampled rows."""
    w = wl.s1r2_cylinder(seed=1002)
    T, shape = steps.s1r2_table(2, w.L, w.Rsim)
    rows = np.linspace(0, w.N - 1, 64).astype(np.int32)
    for order in (2, 4):
        print(f"case s1r2 N={w.N} order={order} sampled rows", flush=True)
        kw = dict(is_periodic=2, L=w.L, Rsim=w.Rsim, cosmology=1, comoving=1,
                  mass_in_unit_sphere=w.mass_in_unit_sphere, s1r2_lookup=1, interp_order=order, s1r2_table=T,
                  s1r2_shape=shape)
        d, keep = steps.make_desc(steps.TOPO_S1R2, **kw)
        sim = steps.Sim(d, w.N)
        sim.set_state(w.x, w.M, w.soft)
        sim.forces()
        F = sim.get_forces()
        ms = sim.last_kernel_ms()
        print(f"[S1R2 N={w.N} order={order}] pair kernel {ms[0]:.1f} ms -> {w.N ** 2 / ms[0] / 1e6:.1f} G pair/s",
              flush=True)
        sim.close()
        Fo = oracle.forces_rows(oracle.params(2, **kw), w.x, w.M, w.soft, rows)
        e = rel_err(F[rows], Fo)
        report(f"S1R2 N={w.N} order={order} sampled", e)
        assert np.median(e) <= 1e-12


def test_s1r2_ngp_has_no_mixed_path(steps, oracle, cyl):
    w, T, shape = cyl
    d, keep = steps.make_desc(steps.TOPO_S1R2, is_periodic=2, L=w.L, Rsim=w.Rsim, s1r2_lookup=1, interp_order=0,
                              s1r2_table=T, s1r2_shape=shape, precision=1)
    with pytest.raises(steps.StepsError):
        steps.force_eval(d, w.x, w.M, w.soft)


@pytest.mark.parametrize("order", [2, 4, "nearest"])
def test_s1r2_mixed_precision_error(steps, oracle, cyl, order):
    """FP32 pair arithmetic (nearest z-image, FP32-interpolated Ewald
    correction) with FP64 accumulation across tiles, against the FP64 path and
    the oracle: p99 relative error <= 1e-5 for the pair sum alone (cosmology
    off).  With the comoving term the total force of a particle near the axis
    is a small difference of the pair sum and the background term
    2 pi rho_c Omega_m (x, y, 0), so the same absolute error is a larger
    relative one: p99 <= 3e-5 there."""
    print(f"case s1r2 mixed order={order}", flush=True)
    w, T, shape = cyl
    for cosmology in (0, 1):
        if order == "nearest":
            radial = oracle.radial_table(w.Rsim, 0.5 * w.L, 1000, 1000)
            kw = dict(is_periodic=1, L=w.L, Rsim=w.Rsim, cosmology=cosmology, comoving=cosmology,
                      mass_in_unit_sphere=w.mass_in_unit_sphere, radial_table=radial)
        else:
            kw = dict(is_periodic=2, L=w.L, Rsim=w.Rsim, cosmology=cosmology, comoving=cosmology,
                      mass_in_unit_sphere=w.mass_in_unit_sphere, s1r2_lookup=1, interp_order=order,
                      s1r2_table=T, s1r2_shape=shape)
        d64, k1 = steps.make_desc(steps.TOPO_S1R2, precision=0, **kw)
        dmx, k2 = steps.make_desc(steps.TOPO_S1R2, precision=1, **kw)
        F64 = steps.force_eval(d64, w.x, w.M, w.soft)
        Fmx = steps.force_eval(dmx, w.x, w.M, w.soft)
        Fo = oracle.forces(oracle.params(2, **kw), w.x, w.M, w.soft)
        e = rel_err(Fmx, F64)
        eo = rel_err(Fmx, Fo)
        report(f"S1R2 mixed order={order} cosmology={cosmology} vs fp64", e)
        report(f"S1R2 mixed order={order} cosmology={cosmology} vs oracle", eo)
        bound = 1e-5 if cosmology == 0 else 3e-5
        assert np.percentile(e, 99) <= bound
        assert np.percentile(eo, 99) <= bound
        assert np.median(e) <= 2e-6
        assert not np.array_equal(Fmx, F64)     # the FP32 path really ran


@pytest.mark.parametrize("N", [2, 65, 1000])
@pytest.mark.parametrize("mode", ["t3_ewald", "t3_nearest", "s1r2_cic", "s1r2_tsc"])
def test_periodic_ragged_sizes(steps, oracle, t3_tab, cyl, mode, N):
    """Sizes that fill neither a 64-source tile nor a 256-row block (the
    clamped lanes and partial tiles of the window kernels): FP64 against the
    oracle on every row, mixed against FP64."""
    print(f"case periodic ragged {mode} N={N}", flush=True)
    rng = np.random.default_rng(N)
    M = rng.uniform(0.5, 2.0, N)
    if mode.startswith("t3"):
        L = L_T3
        x = rng.uniform(0.0, L, (N, 3))
        h = np.full(N, 0.02 * L / 64)
        ip = 2 if mode == "t3_ewald" else 1
        tab = t3_tab if ip == 2 else None
        kw = dict(is_periodic=ip, L=L, t3_table=tab)
        d, keep = steps.make_desc(steps.TOPO_T3, **kw)
        dm, keepm = steps.make_desc(steps.TOPO_T3, precision=steps.PREC_MIXED, **kw)
        p = oracle.params(1, interp_order=4, **kw)
    else:
        w, T, shape = cyl
        r = w.Rsim * np.sqrt(rng.uniform(0.0, 1.0, N))
        phi = rng.uniform(0.0, 2 * np.pi, N)
        x = np.stack([r * np.cos(phi), r * np.sin(phi), rng.uniform(-w.L / 2, w.L / 2, N)], axis=1)
        h = np.full(N, 0.5)
        kw = dict(is_periodic=2, L=w.L, Rsim=w.Rsim, s1r2_lookup=1, interp_order=2 if mode == "s1r2_cic" else 4,
                  s1r2_table=T, s1r2_shape=shape)
        d, keep = steps.make_desc(steps.TOPO_S1R2, **kw)
        dm, keepm = steps.make_desc(steps.TOPO_S1R2, precision=steps.PREC_MIXED, **kw)
        p = oracle.params(2, **kw)
    F = steps.force_eval(d, x, M, h)
    e = rel_err(F, oracle.forces(p, x, M, h))
    report(f"{mode} N={N} vs oracle", e)
    assert np.median(e) <= 1e-12 and e.max() <= 1e-9
    Fm = steps.force_eval(dm, x, M, h)
    em = rel_err(Fm, F)
    report(f"{mode} N={N} mixed vs fp64", em)
    assert np.isfinite(Fm).all()
    if N >= 65:  # with two particles one nearly cancelling force would decide the statistic
        assert np.median(em) <= 1e-5 and np.percentile(em, 90) <= 1e-4


def test_table_edited_in_place_is_uploaded_again(steps, t3_tab):
    """The one-shot entry point keeps the device copy of a table between calls
    and re-uploads it when the caller's table changes.  A single entry edited
    in place (same pointer, same size) at a node a pair interpolates from, and
    not on any stride a sampled fingerprint would read, must reach the device:
    the forces then equal those of a fresh copy of the edited table bitwise."""
    print("case t3 table edited in place", flush=True)
    L, ng = L_T3, 63
    x = np.array([[10.0, 20.0, 30.0], [27.3, 25.1, 109.8]])
    M = np.array([1.0, 2.0])
    h = np.array([0.01, 0.01])
    tab = t3_tab.copy()
    d, keep = steps.make_desc(steps.TOPO_T3, is_periodic=2, L=L, t3_table=tab)
    live = keep[0]
    assert np.shares_memory(live, tab)
    F0 = steps.force_eval(d, x, M, h)
    dd = x[1] - x[0]
    dd -= L * np.round(dd / L)                       # nearest image, as the kernel takes it
    u = (dd + 0.5 * L) * ng / L - 0.5                # grid coordinate of the displacement
    ix, iy, iz = (int(round(c)) for c in u)          # a node inside the tricubic stencil
    idx = ((ix * ng + iy) * ng + iz) * 3 + 0
    n = live.size
    assert idx % max(1, n // 4096) != 0 and idx != n - 1
    live[idx] += 50.0 * np.abs(t3_tab).max()
    F1 = steps.force_eval(d, x, M, h)
    assert not np.array_equal(F1, F0)
    d2, keep2 = steps.make_desc(steps.TOPO_T3, is_periodic=2, L=L, t3_table=live.copy())
    F2 = steps.force_eval(d2, x, M, h)
    assert np.array_equal(F1, F2)
    live[idx] = t3_tab[idx]                           # and back
    assert np.array_equal(steps.force_eval(d, x, M, h), F0)


@pytest.mark.parametrize("is_periodic", [3, 4])
def test_device_t3_table_finer_grids(steps, oracle, is_periodic):
    """IS_PERIODIC 3 and 4 (main.cc:430-447): 127^3 and 255^3 tables with
    wider cuts, built on the device; sampled nodes against the oracle's Ewald
    correction evaluated directly at the node (ewald_space.cc:210-286)."""
    print(f"case t3 table is_periodic={is_periodic}", flush=True)
    L = L_T3
    T = steps.t3_table(is_periodic, L)
    alpha, rel, rec, ng = oracle.t3_params(is_periodic, L)
    assert ng == (127 if is_periodic == 3 else 255) and T.size == ng ** 3 * 3
    T = T.reshape(ng, ng, ng, 3)
    scale = np.abs(T).max()
    rng = np.random.default_rng(is_periodic)
    idx = rng.integers(0, ng, size=(120, 3))
    idx[:4] = [[0, 0, 0], [ng - 1, ng // 2, 3], [ng // 2, ng // 2, ng // 2], [1, ng - 2, ng // 3]]
    worst = 0.0
    for i, j, k in idx:
        r = (np.array([i, j, k]) + 0.5) * (L / ng) - 0.5 * L
        D = oracle.t3_correction(r, L, alpha, rel, rec)
        worst = max(worst, np.abs(T[i, j, k] - D).max() / scale)
    print(f"[T3 table is_periodic={is_periodic}] sampled max |dT|/max|T| = {worst:.3e}", flush=True)
    assert worst <= 1e-11


@pytest.mark.parametrize("is_periodic", [3, 4])
def test_device_s1r2_table_finer_grids(steps, oracle, is_periodic):
    """IS_PERIODIC 3 and 4 (main.cc:574-602): Nz = 256 and 512 with wider
    cuts; sampled nodes against the oracle's singly periodic pair force minus
    the nearest-image Newton force (ewald_space.cc:754-798)."""
    print(f"case s1r2 table is_periodic={is_periodic}", flush=True)
    Lz, Rsim = 200.0, 1000.0
    T, (nrho, nz) = steps.s1r2_table(is_periodic, Lz, Rsim)
    alpha, rel, rec, nz_o, nrho_o = oracle.s1r2_params(is_periodic, Lz, Rsim)
    assert (nrho, nz) == (nrho_o, nz_o) and nz == (256 if is_periodic == 3 else 512)
    T = T.reshape(nrho, nz, 2)
    scale = np.abs(T).max()
    drho, dzg = 2.25 * Rsim / (nrho - 1), Lz / nz
    rng = np.random.default_rng(10 + is_periodic)
    worst = 0.0
    for ir, iz in list(zip(rng.integers(0, nrho, 150), rng.integers(0, nz, 150))) + [(0, 0), (nrho - 1, nz - 1)]:
        rho, z = ir * drho, (iz + 0.5) * dzg - 0.5 * Lz
        Fp = oracle.s1r2_force_pair((rho, 0.0, z), Lz, alpha, rel, rec)
        r2 = rho * rho + z * z                       # |z| < Lz/2: z is its own nearest image
        newton = np.array([-rho, -z]) * r2 ** -1.5 if r2 > 0 else np.zeros(2)
        worst = max(worst, np.abs(T[ir, iz] - (np.array([Fp[0], Fp[2]]) - newton)).max() / scale)
    print(f"[S1R2 table is_periodic={is_periodic}] sampled max |dT|/max|T| = {worst:.3e}", flush=True)
    assert worst <= 1e-10


def test_forces_with_finer_tables(steps, oracle):
    """The pair kernels on the IS_PERIODIC = 3 tables (T3 127^3: the global-
    table kernel, the LDS window needs Ng <= 119; S1R2 Nz = 256, CIC and TSC)
    against the oracle on the same tables, sampled rows."""
    print("case forces with is_periodic=3 tables", flush=True)
    w = wl.t3_lattice(n_side=24, seed=1001, is_periodic=3)
    T = steps.t3_table(3, w.L)
    d, keep = steps.make_desc(steps.TOPO_T3, is_periodic=3, L=w.L, t3_table=T)
    F = steps.force_eval(d, w.x, w.M, w.soft)
    rows = np.linspace(0, w.N - 1, 48).astype(np.int32)
    p = oracle.params(1, is_periodic=3, L=w.L, t3_table=T, interp_order=4)
    e = rel_err(F[rows], oracle.forces_rows(p, w.x, w.M, w.soft, rows))
    report("T3 is_periodic=3 (127^3) sampled rows", e)
    assert np.median(e) <= 1e-12 and e.max() <= 1e-10
    c = wl.s1r2_cylinder(n_omega=16, n_phi=8, n_z=16, seed=6, is_periodic=3)
    Ts, shape = steps.s1r2_table(3, c.L, c.Rsim)
    kw = dict(is_periodic=3, L=c.L, Rsim=c.Rsim, cosmology=1, comoving=1, mass_in_unit_sphere=c.mass_in_unit_sphere)
    rows = np.linspace(0, c.N - 1, 64).astype(np.int32)
    for order in (2, 4):
        d, keep = steps.make_desc(steps.TOPO_S1R2, s1r2_lookup=1, interp_order=order, s1r2_table=Ts,
                                  s1r2_shape=shape, **kw)
        F = steps.force_eval(d, c.x, c.M, c.soft)
        p = oracle.params(2, s1r2_lookup=1, interp_order=order, s1r2_table=Ts, s1r2_shape=shape, **kw)
        e = rel_err(F[rows], oracle.forces_rows(p, c.x, c.M, c.soft, rows))
        report(f"S1R2 is_periodic=3 (Nz=256) order={order} sampled rows", e)
        assert np.median(e) <= 1e-12 and e.max() <= 1e-10


@pytest.mark.parametrize("mode", ["s1r2_cic", "s1r2_tsc", "s1r2_nearest", "t3_nearest"])
def test_mixed_full_size(steps, oracle, mode):
    """The mixed path at full size (S1R2 N = 524288, T3 nearest image on the
    64^3 lattice), all rows against FP64 for the pair sum alone and with the
    comoving term, p99 <= 1e-5 (measured 6e-7 .. 3.8e-6); and 32 sampled rows
    against the oracle, median <= 1e-6 and max <= 1e-4 (too few rows for a
    p99: the sample's worst rows are the tail's).  Pairs within 1e-6 L of L/2
    on a wrapped axis are redone in FP64, so no pair takes another image than
    the reference's; the max error over all rows stays below 1e-3."""
    print(f"case mixed full size {mode}", flush=True)
    if mode == "t3_nearest":
        w = wl.t3_lattice(is_periodic=1)
        topo, base = steps.TOPO_T3, dict(is_periodic=1, L=w.L)
        cases = [(0, base), (1, dict(base, cosmology=1, comoving=1))]
        otop = 1
    else:
        w = wl.s1r2_cylinder(seed=1002, is_periodic=1 if mode == "s1r2_nearest" else 2)
        topo, otop = steps.TOPO_S1R2, 2
        if mode == "s1r2_nearest":
            radial = steps.radial_table(w.Rsim, 0.5 * w.L, 1000, 1000)
            base = dict(is_periodic=1, L=w.L, Rsim=w.Rsim, radial_table=radial)
        else:
            T, shape = steps.s1r2_table(2, w.L, w.Rsim)
            base = dict(is_periodic=2, L=w.L, Rsim=w.Rsim, s1r2_lookup=1, s1r2_table=T, s1r2_shape=shape,
                        interp_order=2 if mode == "s1r2_cic" else 4)
        cases = [(0, base), (1, dict(base, cosmology=1, comoving=1, mass_in_unit_sphere=w.mass_in_unit_sphere))]
    rows = np.linspace(0, w.N - 1, 32).astype(np.int32)
    for cosmology, kw in cases:
        d64, k1 = steps.make_desc(topo, precision=steps.PREC_FP64, **kw)
        dmx, k2 = steps.make_desc(topo, precision=steps.PREC_MIXED, **kw)
        F64 = steps.force_eval(d64, w.x, w.M, w.soft)
        Fmx = steps.force_eval(dmx, w.x, w.M, w.soft)
        e = rel_err(Fmx, F64)
        okw = dict(kw, interp_order=kw.get("interp_order", 4))
        eo = rel_err(Fmx[rows], oracle.forces_rows(oracle.params(otop, **okw), w.x, w.M, w.soft, rows))
        report(f"{mode} mixed N={w.N} cosmology={cosmology} vs fp64", e)
        report(f"{mode} mixed N={w.N} cosmology={cosmology} vs oracle (32 rows)", eo)
        assert np.percentile(e, 99) <= 1e-5 and e.max() <= 1e-3
        assert np.median(eo) <= 1e-6 and eo.max() <= 1e-4
