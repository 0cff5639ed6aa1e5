"""Seeded synthetic workloads for the direct-summation force path.

Random numbers come from SplitMix64 (counter-based, vectorised in numpy), so
every machine produces bit-identical inputs for a given seed.  Units are the
reference's code units: Mpc, 1e11 Msun, G = 1, velocities in UNIT_V
(StePS/src/global_variables.h:16-18).

Configurations (sizes are parameters so tests can use small N):
  r3_compactified   open R^3, stereographic-style shells, m ~ (1+r^2/L^2)^3
  t3_lattice        periodic T^3, 64^3 grid + k^-1 Gaussian displacement field
  s1r2_cylinder     S^1 x R^2, (omega, phi, z) lattice, m ~ (1+R^2/L^2)^2
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

PI = 3.14159265358979323846264338327950288419716939937510
UNIT_T = 47.14829951063323   # Gy
UNIT_V = 20.738652969925447  # km/s

# LCDM parameters of StePS/examples/LCDM_cylindrical_Lz100_D1000.param:3-9
OMEGA_M = 0.3111
OMEGA_L = 0.6889
H0_KMS = 67.66
H0 = H0_KMS / UNIT_V                       # read_paramfile.cc:403
RHO_CRIT = 3.0 * H0 * H0 / (8.0 * PI)      # main.cc:1262

_GOLD = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, stream: int, n: int) -> np.ndarray:
    """n outputs of SplitMix64 started at state seed*2^20 + stream*2^40."""
    base = np.uint64((seed * (1 << 20) + stream * (1 << 40)) & 0xFFFFFFFFFFFFFFFF)
    with np.errstate(over="ignore"):
        z = base + _GOLD * np.arange(1, n + 1, dtype=np.uint64)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """U[0,1) with 53 random bits."""
    return (splitmix64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, n: int) -> np.ndarray:
    """Standard normals by Box-Muller from two uniform streams."""
    u1 = uniform(seed, stream, n)
    u2 = uniform(seed, stream + 1000, n)
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * PI * u2)


@dataclass
class Workload:
    name: str
    topology: int                 # 0 R3, 1 T3, 2 S1R2
    x: np.ndarray                 # [N,3]
    v: np.ndarray                 # [N,3] code units
    M: np.ndarray                 # [N]
    soft: np.ndarray              # [N] softening lengths h_i
    particle_radius: float        # PARTICLE_RADII
    rho_part: float
    L: float = 0.0                # T3 box / S1R2 z period
    Rsim: float = 0.0
    is_periodic: int = 0
    cosmology: int = 1
    comoving: int = 1
    mass_in_unit_sphere: float = 0.0
    a_start: float = 1.0 / 32.0
    extra: dict = field(default_factory=dict)

    @property
    def N(self) -> int:
        return int(self.M.size)


def softening(M: np.ndarray, particle_radius: float):
    """utils.cc:59-82: h_i = cbrt(M_i * 3/(4 pi rho_part)), rho_part from M_min."""
    mmin = float(M.min())
    rho_part = mmin / (4.0 * PI * particle_radius ** 3 / 3.0)
    cb = 3.0 / rho_part / (4.0 * PI)
    return np.cbrt(M * cb), rho_part


def r3_compactified(N: int = 16384, seed: int = 1000, L: float = 100.0, total_mass: float = 1.0e4,
                    omega_max: float = 2.0 * math.atan(10.0), soft_coeff: float = 0.05,
                    sigma_v_kms: float = 50.0, cosmology: int = 1, comoving: int = 1,
                    a_start: float = 1.0 / 32.0) -> Workload:
    """Open R^3: omega_i on a uniform grid in (0, omega_max), r = L tan(omega/2),
    isotropic directions, m ~ (1 + r^2/L^2)^3 summing to total_mass,
    h_i = soft_coeff * m_i^(1/3), Gaussian velocities of sigma_v per component."""
    i = np.arange(N, dtype=np.float64)
    omega = omega_max * (i + 0.5) / N
    r = L * np.tan(0.5 * omega)
    mu = 2.0 * uniform(seed, 0, N) - 1.0
    phi = 2.0 * PI * uniform(seed, 1, N)
    s = np.sqrt(np.maximum(0.0, 1.0 - mu * mu))
    x = np.stack([r * s * np.cos(phi), r * s * np.sin(phi), r * mu], axis=1)
    M = (1.0 + (r / L) ** 2) ** 3
    M *= total_mass / M.sum()
    R0 = soft_coeff * float(M.min()) ** (1.0 / 3.0)
    h, rho_part = softening(M, R0)
    v = np.stack([normal(seed, 2 + k, N) for k in range(3)], axis=1) * (sigma_v_kms / UNIT_V)
    if cosmology == 1 and comoving == 1:
        v /= math.sqrt(a_start)   # main.cc:1155-1163
    misph = 4.0 * PI * RHO_CRIT * OMEGA_M / 3.0 if cosmology == 1 else 0.0  # main.cc:1313
    return Workload("r3_compactified", 0, x, v, M, h, R0, rho_part, cosmology=cosmology,
                    comoving=comoving, mass_in_unit_sphere=misph, a_start=a_start)


def _gaussian_field(n: int, seed: int, stream: int, box: float) -> np.ndarray:
    """Real Gaussian field on an n^3 grid with amplitude ~ k^-1 (P ~ k^-2) for
    0 < |k| <= k_Nyquist, unit rms."""
    w = normal(seed, stream, n ** 3).reshape(n, n, n)
    wk = np.fft.rfftn(w)
    k1 = np.fft.fftfreq(n, d=box / n) * 2.0 * PI
    kz = np.fft.rfftfreq(n, d=box / n) * 2.0 * PI
    kk = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + kz[None, None, :] ** 2)
    knyq = PI * n / box
    amp = np.zeros_like(kk)
    ok = (kk > 0) & (kk <= knyq)
    amp[ok] = 1.0 / kk[ok]
    f = np.fft.irfftn(wk * amp, s=(n, n, n), axes=(0, 1, 2))
    return f / np.sqrt(np.mean(f * f))


def t3_lattice(n_side: int = 64, seed: int = 1001, L: float = 100.0 / 0.6766,
               sigma_cells: float = 0.5, soft_cells: float = 0.02, is_periodic: int = 2) -> Workload:
    """Periodic T^3: n_side^3 equal masses on a grid displaced by a k^-1
    Gaussian field (rms sigma_cells grid spacings per component), wrapped into
    [0, L); M = Omega_m rho_c L^3 / N; PARTICLE_RADII = soft_cells * spacing."""
    N = n_side ** 3
    d = L / n_side
    g = (np.arange(n_side) + 0.5) * d
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    for k in range(3):
        x[:, k] += sigma_cells * d * _gaussian_field(n_side, seed, 10 + k, L).ravel()
    x = np.mod(x, L)
    x[x >= L] -= L
    M = np.full(N, OMEGA_M * RHO_CRIT * L ** 3 / N)
    R0 = soft_cells * d
    h, rho_part = softening(M, R0)
    v = np.zeros_like(x)
    return Workload("t3_lattice", 1, x, v, M, h, R0, rho_part, L=L, is_periodic=is_periodic,
                    cosmology=1, comoving=1)


def s1r2_cylinder(n_omega: int = 128, n_phi: int = 64, n_z: int = 64, seed: int = 1002,
                  Lz: float = 200.0, Lc: float = 100.0, rsim_over_lc: float = 10.0,
                  jitter: float = 0.1, soft_coeff: float = 0.05, is_periodic: int = 2) -> Workload:
    """S^1 x R^2: shells R_a = Lc tan(omega_a/2) with omega up to
    2 atan(Rsim/Lc), golden-angle phi offsets per shell, z planes; isotropic
    Gaussian jitter of rms jitter*local spacing; z wrapped;
    m ~ (1 + R^2/Lc^2)^2 normalised to Omega_m rho_c pi Rsim^2 Lz.

    Rsim is set exactly (Lc * rsim_over_lc): the reference's radial-table
    trapezoid (utils.cc:174-205) evaluates sqrt(R^2 - Y^2) at Y = acc*(R/acc),
    which is NaN whenever that product rounds above R."""
    Rsim = Lc * rsim_over_lc
    omega_max = 2.0 * math.atan(rsim_over_lc)
    a = np.arange(n_omega, dtype=np.float64)
    R_a = Lc * np.tan(0.5 * omega_max * (a + 0.5) / n_omega)
    R_lo = Lc * np.tan(0.5 * omega_max * a / n_omega)
    R_hi = Lc * np.tan(0.5 * omega_max * (a + 1.0) / n_omega)
    spacing = (PI * (R_hi ** 2 - R_lo ** 2) * Lz / (n_phi * n_z)) ** (1.0 / 3.0)
    A, B, Cz = np.meshgrid(np.arange(n_omega), np.arange(n_phi), np.arange(n_z), indexing="ij")
    A, B, Cz = A.ravel(), B.ravel(), Cz.ravel()
    N = A.size
    R = R_a[A]
    phi = 2.0 * PI * (B + 0.5) / n_phi + A * 2.39996
    z = Lz * (Cz + 0.5) / n_z
    x = np.stack([R * np.cos(phi), R * np.sin(phi), z], axis=1)
    s = spacing[A]
    for k in range(3):
        x[:, k] += jitter * s * normal(seed, 20 + k, N)
    x[:, 2] = np.mod(x[:, 2], Lz)
    x[x[:, 2] >= Lz, 2] -= Lz
    M = (1.0 + (R / Lc) ** 2) ** 2
    M *= OMEGA_M * RHO_CRIT * PI * Rsim ** 2 * Lz / M.sum()
    R0 = soft_coeff * float(M.min()) ** (1.0 / 3.0)
    h, rho_part = softening(M, R0)
    v = np.zeros_like(x)
    misph = 2.0 * PI * RHO_CRIT * OMEGA_M  # main.cc:1288
    return Workload("s1r2_cylinder", 2, x, v, M, h, R0, rho_part, L=Lz, Rsim=Rsim,
                    is_periodic=is_periodic, cosmology=1, comoving=1, mass_in_unit_sphere=misph)


def r3_sized(N: int, seed: int) -> Workload:
    return r3_compactified(N=N, seed=seed)


def s1r2_sized(N: int, seed: int) -> Workload:
    """S1R2 generator at about N particles: n_phi = n_z = ceil(sqrt(N/128)) rounded
    up to a multiple of 8, n_omega = N // (n_phi*n_z)."""
    nq = int(math.ceil(math.sqrt(N / 128.0)))
    nq = max(8, 8 * int(math.ceil(nq / 8.0)))
    n_omega = max(1, N // (nq * nq))
    return s1r2_cylinder(n_omega=n_omega, n_phi=nq, n_z=nq, seed=seed)
