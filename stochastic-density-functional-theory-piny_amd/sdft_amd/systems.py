"""Synthetic stochastic-DFT systems: cells, plane-wave sets, grid maps, potentials.

This module builds the *inputs* the filter path consumes, in the storage
conventions of PINY's sDFT code:

* half-sphere plane-wave set with G=0 stored last; Gamma symmetry
  c(-G) = c(G)* (stodft/filter.c:318-706 operates on `cre_up/cim_up`
  arrays of `ncoef` entries per state, entry `ncoef` being G=0);
* FFTW row-major grid (a, b, c) with c fastest; the pack/unpack maps are
  `mapFFTWFilter`/`mapConFFTWFilter` (mathlib/fft_create_package.c:237-248,
  setfft_indx at :981-1016): 0-based linear index of +G and of -G;
* `ak2Kinetic` = |G|^2 (energy/cp/cp_energy_eext.c:2259-2304, no cap);
* real-space Kohn-Sham potential on the same grid (FFTW order, as the
  reference transposes it at energy/cp/cp_energy_ee_rho.c:1944-1970);
* Kleinman-Bylander projectors tabulated on the plane-wave set
  (the quantities the G-space KB path builds per call in
  stodft/reduced-energy/cp-energy-eext-stodft.c:701-1168);
* Rademacher noise orbitals (stodft/gen-noise.c:162-374 semantics:
  real-space values +-sqrt(N), transformed to the coefficient set with
  unit covariance).

Physics of the synthetic systems (analytic, no pseudopotential files):
Si diamond cells; local pseudopotential -Z erf(r/rc)/r (Gaussian-screened
Coulomb, Z=4, rc=1.0 bohr) plus the Hartree and Slater-exchange potential of
a fixed superposition of atomic Gaussian densities; one s and one p KB
projector per atom with radial form r^l exp(-r^2/(2 rc^2)), rc=0.8 bohr,
normalised to unit norm, KB energies `h_s`, `h_p` (Ha).

Everything here is host-side set-up done once per system; the hot path
(H.psi inside the Chebyshev filter) runs in the HIP library.
"""
from __future__ import annotations

import dataclasses
import math
from typing import Optional

import numpy as np

# ---------------------------------------------------------------------------
# Cells and atoms
# ---------------------------------------------------------------------------

_DIAMOND_BASIS = np.array(
    [[0.0, 0.0, 0.0], [0.0, 0.5, 0.5], [0.5, 0.0, 0.5], [0.5, 0.5, 0.0],
     [0.25, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.25, 0.75],
     [0.75, 0.75, 0.25]])


def diamond_supercell(a0: float, rep: int) -> tuple[float, np.ndarray]:
    """Cubic diamond supercell rep x rep x rep of conventional cubic cells of
    side a0 (bohr).  Returns (L, cartesian positions (natm, 3))."""
    cells = np.array([[i, j, k] for i in range(rep) for j in range(rep)
                      for k in range(rep)], dtype=np.float64)
    frac = (cells[:, None, :] + _DIAMOND_BASIS[None, :, :]).reshape(-1, 3) / rep
    L = a0 * rep
    return L, frac * L


def _splitmix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser on uint64 arrays (public-domain mixing function)."""
    x = (x + np.uint64(0x9E3779B97F4A7C15))
    x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return x ^ (x >> np.uint64(31))


def rademacher_signs(seed: int, orbital: int, npts: int) -> np.ndarray:
    """+-1 signs for grid points 0..npts-1 of one orbital.

    Counter-based: sign = top bit of splitmix64(seed<<44 | orbital<<23 | r).
    Grids up to 2^23 points and 2^21 orbitals per seed."""
    if npts > (1 << 23) or orbital >= (1 << 21) or seed >= (1 << 20):
        raise ValueError("noise key out of range")
    with np.errstate(over="ignore"):
        key = (np.uint64(seed) << np.uint64(44)) | (np.uint64(orbital) << np.uint64(23)) \
            | np.arange(npts, dtype=np.uint64)
        h = _splitmix64(key)
    return np.where((h >> np.uint64(63)) == 1, -1.0, 1.0)


# ---------------------------------------------------------------------------
# Plane-wave set and FFTW maps
# ---------------------------------------------------------------------------

def half_sphere(L: float, g2max: float) -> np.ndarray:
    """Integer triplets (ka, kb, kc) of the half sphere |G|^2 <= g2max,
    G = 2 pi k / L, half space: ka>0, or ka=0 & kb>0, or ka=kb=0 & kc>0.
    G=0 is appended last (PINY stores it as coefficient `ncoef`)."""
    b = 2.0 * math.pi / L
    K = int(math.sqrt(g2max) / b) + 1
    r = np.arange(-K, K + 1)
    ka, kb, kc = np.meshgrid(r, r, r, indexing="ij")
    ka, kb, kc = ka.ravel(), kb.ravel(), kc.ravel()
    g2 = b * b * (ka * ka + kb * kb + kc * kc)
    half = (ka > 0) | ((ka == 0) & (kb > 0)) | ((ka == 0) & (kb == 0) & (kc > 0))
    sel = (g2 <= g2max) & half
    k = np.stack([ka[sel], kb[sel], kc[sel]], axis=1)
    # order: ka outermost, then kb, then kc (a layout choice; PINY's kastr order
    # is not observable through the filter's results)
    order = np.lexsort((k[:, 2], k[:, 1], k[:, 0]))
    k = k[order]
    return np.concatenate([k, np.zeros((1, 3), dtype=k.dtype)], axis=0)


def fftw_maps(k: np.ndarray, n: tuple[int, int, int]) -> tuple[np.ndarray, np.ndarray]:
    """mapFFTWFilter / mapConFFTWFilter: 0-based linear grid index of +G and -G
    in the (a, b, c) c-fastest layout (fft_create_package.c:237-248,981-1016).
    The G=0 entry maps to 0 in both."""
    n1, n2, n3 = n
    ia = np.mod(k[:, 0], n1)
    ib = np.mod(k[:, 1], n2)
    ic = np.mod(k[:, 2], n3)
    m = (ia * n2 + ib) * n3 + ic
    ja = np.mod(-k[:, 0], n1)
    jb = np.mod(-k[:, 1], n2)
    jc = np.mod(-k[:, 2], n3)
    mc = (ja * n2 + jb) * n3 + jc
    return m.astype(np.int32), mc.astype(np.int32)


# ---------------------------------------------------------------------------
# System container
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class System:
    name: str
    L: float                      # cubic cell side (bohr)
    n: tuple                      # FFT grid (n1, n2, n3)
    pos: np.ndarray               # (natm, 3) cartesian (bohr)
    k: np.ndarray                 # (ncoef, 3) integer G triplets, G=0 last
    map: np.ndarray               # (ncoef,) int32
    map_conj: np.ndarray          # (ncoef,) int32
    ak2: np.ndarray               # (ncoef,) |G|^2
    gvec: np.ndarray              # (ncoef, 3) cartesian G
    vks: np.ndarray               # (n1*n2*n3,) real, FFTW order
    proj: np.ndarray              # (nproj, ncoef) complex Q_p(G)
    proj_energy: np.ndarray       # (nproj,) E_p = h_l / V (Ha)
    proj_atom: np.ndarray         # (nproj,) atom index of projector p
    proj_l: np.ndarray            # (nproj,) angular momentum
    proj_m: np.ndarray            # (nproj,) real-Ylm component (0; x,y,z -> 0,1,2)
    z_val: float = 4.0
    rc_loc: float = 1.0
    rc_nl: float = 0.8
    h_kb: tuple = (2.0, 1.0)
    sigma_rho: float = 1.2

    @property
    def ncoef(self) -> int:
        return int(self.k.shape[0])

    @property
    def ngrid(self) -> int:
        return int(self.n[0] * self.n[1] * self.n[2])

    @property
    def volume(self) -> float:
        return self.L ** 3

    @property
    def natm(self) -> int:
        return int(self.pos.shape[0])

    @property
    def nproj(self) -> int:
        return int(self.proj.shape[0])

    @property
    def nelec(self) -> float:
        return self.z_val * self.natm


def _grid_freqs(n: int) -> np.ndarray:
    f = np.arange(n)
    return np.where(f <= n // 2, f, f - n)


def _structure_factor(kfull: np.ndarray, L: float, pos: np.ndarray) -> np.ndarray:
    """S(G) = sum_a exp(-i G.R_a) on an arbitrary list of integer triplets."""
    frac = pos / L
    out = np.zeros(kfull.shape[0], dtype=np.complex128)
    chunk = 64
    for i in range(0, frac.shape[0], chunk):
        ph = 2.0 * math.pi * (kfull @ frac[i:i + chunk].T)
        out += np.exp(-1j * ph).sum(axis=1)
    return out


def _structure_factor_grid(n: tuple, L: float, pos: np.ndarray) -> np.ndarray:
    """S(G) on the full FFT grid (FFTW order) from per-axis phase tables:
    exp(-i G.R) = e_a(ka) e_b(kb) e_c(kc); equal to _structure_factor up to
    rounding, without one complex exponential per grid point and atom."""
    frac = pos / L
    e = [np.exp(-2j * math.pi * np.outer(_grid_freqs(m), frac[:, d])) for d, m in enumerate(n)]
    return np.einsum("ia,ja,ka->ijk", e[0], e[1], e[2], optimize=True).ravel()


def _structure_factor_sep(k: np.ndarray, L: float, pos: np.ndarray) -> np.ndarray:
    """S(G) at the integer triplets k from per-axis phase tables over the
    triplets' bounding box (one contraction over atoms, then a lookup); equal
    to _structure_factor up to rounding."""
    ki = np.asarray(k).astype(np.int64)
    lo, hi = ki.min(axis=0), ki.max(axis=0)
    frac = pos / L
    e = [np.exp(-2j * math.pi * np.outer(np.arange(lo[d], hi[d] + 1), frac[:, d]))
         for d in range(3)]
    cube = np.einsum("ia,ja,ka->ijk", e[0], e[1], e[2], optimize=True)
    j = ki - lo
    return cube[j[:, 0], j[:, 1], j[:, 2]]


def ks_potential(L: float, n: tuple, pos: np.ndarray, z_val: float, rc_loc: float,
                 sigma_rho: float, separable: bool = False) -> np.ndarray:
    """v_KS(r) on the FFT grid, FFTW order, real.

    v_loc(G) = -4 pi Z/(V G^2) exp(-G^2 rc^2/4) S(G)      (erf-screened Coulomb)
    rho(G)   = (Z/V) exp(-G^2 sigma^2/4) S(G)              (atomic Gaussians)
    v_H(G)   = 4 pi rho(G)/G^2 ; G=0 terms dropped (neutral cell)
    v_x(r)   = -(3 rho(r)/pi)^(1/3)                         (Slater exchange)
    """
    n1, n2, n3 = n
    V = L ** 3
    f1, f2, f3 = _grid_freqs(n1), _grid_freqs(n2), _grid_freqs(n3)
    ka, kb, kc = np.meshgrid(f1, f2, f3, indexing="ij")
    kfull = np.stack([ka.ravel(), kb.ravel(), kc.ravel()], axis=1).astype(np.float64)
    g2 = (2.0 * math.pi / L) ** 2 * (kfull ** 2).sum(axis=1)
    S = _structure_factor_grid(n, L, pos) if separable else _structure_factor(kfull, L, pos)
    with np.errstate(divide="ignore", invalid="ignore"):
        vloc = -4.0 * math.pi * z_val / (V * g2) * np.exp(-g2 * rc_loc ** 2 / 4.0) * S
        rhog = (z_val / V) * np.exp(-g2 * sigma_rho ** 2 / 4.0) * S
        vh = 4.0 * math.pi * rhog / g2
    vloc[0] = 0.0
    vh[0] = 0.0
    N = n1 * n2 * n3
    vg = (vloc + vh).reshape(n1, n2, n3)
    v_r = np.fft.ifftn(vg).real * N
    rho_r = np.fft.ifftn(rhog.reshape(n1, n2, n3)).real * N
    rho_r = np.maximum(rho_r, 0.0)
    vx = -np.cbrt(3.0 * rho_r / math.pi)
    return np.ascontiguousarray((v_r + vx).ravel())


def _real_ylm(l: int, m: int, g: np.ndarray) -> np.ndarray:
    """Real spherical harmonics for l<=1 at directions g (rows).  At g=0 the
    l=1 harmonics are defined as 0 (their projector transform vanishes)."""
    if l == 0:
        return np.full(g.shape[0], 1.0 / math.sqrt(4.0 * math.pi))
    gn = np.linalg.norm(g, axis=1)
    with np.errstate(invalid="ignore", divide="ignore"):
        u = np.where(gn > 0, g[:, m] / np.where(gn > 0, gn, 1.0), 0.0)
    return math.sqrt(3.0 / (4.0 * math.pi)) * u


def kb_table(L: float, pos: np.ndarray, k: np.ndarray, rc_nl: float,
             h_kb: tuple) -> tuple:
    """Projector table Q_p(G) = 4 pi i^l N_l F_l(|G|) Y_p(G^) e^{i G.R_a} on the
    half sphere (G=0 last), with F_l(g) = sqrt(pi/2) rc^(2l+3) g^l e^{-g^2 rc^2/2}
    the Bessel transform of r^l e^{-r^2/(2 rc^2)} and N_l its unit-norm factor.
    Then <beta_p|psi> = V^(-1/2) sum_{all G} c(G) Q_p(G) and the KB operator
    is sum_p |beta_p> h_l <beta_p|, so E_p = h_l / V."""
    V = L ** 3
    b = 2.0 * math.pi / L
    g = k.astype(np.float64) * b
    gn = np.linalg.norm(g, axis=1)
    frac = pos / L
    rows, energy, atom, ls, ms = [], [], [], [], []
    for l, nm in ((0, 1), (1, 3)):
        norm = 1.0 / math.sqrt(math.gamma(l + 1.5) * rc_nl ** (2 * l + 3) / 2.0)
        F = math.sqrt(math.pi / 2.0) * rc_nl ** (2 * l + 3) * gn ** l * np.exp(-gn ** 2 * rc_nl ** 2 / 2.0)
        radial = 4.0 * math.pi * (1j ** l) * norm * F
        for m in range(nm):
            y = _real_ylm(l, m, g)
            base = radial * y
            for a in range(pos.shape[0]):
                ph = np.exp(2j * math.pi * (k @ frac[a]))
                rows.append(base * ph)
                energy.append(h_kb[l] / V)
                atom.append(a)
                ls.append(l)
                ms.append(m)
    # order projectors atom-major: (atom, l, m)
    order = np.lexsort((np.array(ms), np.array(ls), np.array(atom)))
    proj = np.array(rows)[order]
    return (np.ascontiguousarray(proj), np.array(energy)[order], np.array(atom)[order],
            np.array(ls)[order], np.array(ms)[order])


@dataclasses.dataclass
class RealSpaceProjectors:
    """The real-space KB tables of PSEUDO_REAL (typedefs_cp.h: numGridNlppMap,
    gridNlppMap, vnlPhiAtomGridRe/Im, vpsNormList, volElem) as one CSR list
    over projectors: projector p has values[offsets[p]:offsets[p+1]] at the
    FFTW linear grid indices index[...] and KB energy energy[p]."""
    offsets: np.ndarray    # (nproj + 1,) int64
    index: np.ndarray      # (nent,) int32
    values: np.ndarray     # (nent,) beta_p(r)
    energy: np.ndarray     # (nproj,) h_p
    vol_elem: float        # V / N
    atom: np.ndarray       # (nproj,) atom of projector p


def real_space_projectors(sysm, rcut: float = 4.0) -> RealSpaceProjectors:
    """The system's KB projectors sampled on the FFT grid: for every atom the
    grid points within rcut (minimum image, rcut < L/2) and, in the order of
    kb_table (atom, l, m), beta(r) = N_l r^l e^{-r^2/(2 rc^2)} Y_lm(r^) with
    real Y_lm (so the reference's m > 0 factor 2 is inside the harmonics) and
    energy h_l."""
    n = sysm.n[0]
    L = sysm.L
    h = L / n
    if not (2.0 * rcut < L):
        raise ValueError("rcut must be below half the cell")
    rc = sysm.rc_nl
    offs, idx, vals, en, atm = [0], [], [], [], []
    for a, x in enumerate(sysm.pos):
        lo = np.floor((x - rcut) / h).astype(int)
        hi = np.ceil((x + rcut) / h).astype(int)
        ii = [np.arange(lo[d], hi[d] + 1) for d in range(3)]
        I, J, K = np.meshgrid(ii[0], ii[1], ii[2], indexing="ij")
        d = np.stack([I.ravel() * h - x[0], J.ravel() * h - x[1], K.ravel() * h - x[2]], axis=1)
        r2 = (d * d).sum(axis=1)
        sel = r2 <= rcut * rcut
        d, r2 = d[sel], r2[sel]
        lin = ((np.mod(I.ravel()[sel], n) * n + np.mod(J.ravel()[sel], n)) * n
               + np.mod(K.ravel()[sel], n)).astype(np.int32)
        g = np.exp(-r2 / (2.0 * rc * rc))
        for l, nm in ((0, 1), (1, 3)):
            norm = 1.0 / math.sqrt(math.gamma(l + 1.5) * rc ** (2 * l + 3) / 2.0)
            for m in range(nm):
                if l == 0:
                    v = norm * g / math.sqrt(4.0 * math.pi)
                else:
                    v = norm * math.sqrt(3.0 / (4.0 * math.pi)) * d[:, m] * g
                idx.append(lin)
                vals.append(v)
                offs.append(offs[-1] + lin.size)
                en.append(sysm.h_kb[l])
                atm.append(a)
    return RealSpaceProjectors(np.array(offs, dtype=np.int64), np.concatenate(idx),
                               np.concatenate(vals), np.array(en), sysm.L ** 3 / float(n ** 3),
                               np.array(atm))


def real_space_projector_gradients(sysm, rs: RealSpaceProjectors) -> np.ndarray:
    """The gradient tables of real_space_projectors' values (PINY's
    vnlPhiDx/Dy/DzAtomGrid convention: d beta_p / d r at the grid point,
    r - R_a the minimum-image offset from the atom; -d beta_p / d R_a), as
    an array (3, nent) aligned with rs.values: s: -N0 g x_j / rc^2 Y_00,
    p_m: N1 sqrt(3/(4 pi)) (delta_mj - x_m x_j / rc^2) g, g = exp(-r^2 / 2 rc^2)."""
    n = sysm.n[0]
    h = sysm.L / n
    rc = sysm.rc_nl
    out = np.zeros((3, int(rs.offsets[-1])))
    nproj = len(rs.energy)
    p = 0
    while p < nproj:
        a = int(rs.atom[p])
        sl = slice(rs.offsets[p], rs.offsets[p + 1])
        lin = rs.index[sl].astype(np.int64)
        ijk = np.stack([lin // (n * n), (lin // n) % n, lin % n], axis=1).astype(np.float64)
        d = ijk * h - sysm.pos[a]
        d -= sysm.L * np.round(d / sysm.L)
        r2 = (d * d).sum(axis=1)
        g = np.exp(-r2 / (2.0 * rc * rc))
        k = 0
        for l, nm in ((0, 1), (1, 3)):
            norm = 1.0 / math.sqrt(math.gamma(l + 1.5) * rc ** (2 * l + 3) / 2.0)
            for m in range(nm):
                if p + k >= nproj or int(rs.atom[p + k]) != a:
                    break
                slk = slice(rs.offsets[p + k], rs.offsets[p + k + 1])
                for j in range(3):
                    if l == 0:
                        v = -norm / math.sqrt(4.0 * math.pi) * g * d[:, j] / (rc * rc)
                    else:
                        v = norm * math.sqrt(3.0 / (4.0 * math.pi)) * g * (
                            (1.0 if m == j else 0.0) - d[:, m] * d[:, j] / (rc * rc))
                    out[j, slk] = v
                k += 1
        p += k
    return out


def build_system(name: str, a0: float, rep: int, n: int, g2max: float = 12.0,
                 pos: Optional[np.ndarray] = None, with_nl: bool = True,
                 h_kb: tuple = (2.0, 1.0), cell: Optional[float] = None,
                 separable_sf: bool = False) -> System:
    L, p = diamond_supercell(a0, rep)
    if cell is not None:
        L = float(cell)
    if pos is not None:
        p = np.asarray(pos, dtype=np.float64)
    k = half_sphere(L, g2max)
    grid = (n, n, n)
    kmax = np.abs(k).max(axis=0)
    if np.any(2 * kmax >= np.array(grid)):
        raise ValueError("grid too small for the plane-wave sphere")
    m, mc = fftw_maps(k, grid)
    gvec = k.astype(np.float64) * (2.0 * math.pi / L)
    ak2 = (gvec ** 2).sum(axis=1)
    vks = ks_potential(L, grid, p, 4.0, 1.0, 1.2, separable=separable_sf)
    if with_nl:
        proj, pe, pa, pl, pm = kb_table(L, p, k, 0.8, h_kb)
    else:
        proj = np.zeros((0, k.shape[0]), dtype=np.complex128)
        pe = np.zeros(0)
        pa = pl = pm = np.zeros(0, dtype=np.int64)
    return System(name=name, L=L, n=grid, pos=p, k=k, map=m, map_conj=mc, ak2=ak2,
                  gvec=gvec, vks=vks, proj=proj, proj_energy=pe, proj_atom=pa,
                  proj_l=pl, proj_m=pm, h_kb=h_kb)


@dataclasses.dataclass
class DensitySphere:
    """The density plane-wave half sphere |G|^2 <= 4 g2max (PINY's large
    density grid package): the G-set of rho(G), v_H(G) and v_ext(G)."""
    n: tuple
    k: np.ndarray
    map: np.ndarray
    map_conj: np.ndarray
    ak2: np.ndarray
    gvec: np.ndarray

    @property
    def ncoef(self) -> int:
        return int(self.k.shape[0])


def density_sphere(sysm: System, g2max: float = 12.0) -> DensitySphere:
    """Half sphere of the density cutoff (4x the wave-function cutoff in
    energy) on the system's grid, G = 0 last, with its FFTW-order maps."""
    k = half_sphere(sysm.L, 4.0 * g2max)
    if np.any(2 * np.abs(k).max(axis=0) >= np.array(sysm.n)):
        raise ValueError("grid too small for the density sphere")
    m, mc = fftw_maps(k, sysm.n)
    gvec = k.astype(np.float64) * (2.0 * math.pi / sysm.L)
    return DensitySphere(n=tuple(sysm.n), k=k, map=m, map_conj=mc, ak2=(gvec ** 2).sum(axis=1),
                         gvec=gvec)


def vloc_form_factor(sysm: System, dens: DensitySphere) -> np.ndarray:
    """Per-atom local form factor on the density sphere (one atom type):
    f(G) = -4 pi Z / (V G^2) exp(-G^2 rc^2 / 4), zero at G = 0, so that
    v_ext(G) = f(G) sum_a exp(-i G.R_a)."""
    g2 = dens.ak2
    out = np.zeros((1, dens.ncoef))
    nz = g2 > 0
    out[0, nz] = -4.0 * math.pi * sysm.z_val / (sysm.volume * g2[nz]) * np.exp(
        -g2[nz] * sysm.rc_loc ** 2 / 4.0)
    return out


def vloc_density(sysm: System, dens: DensitySphere, separable: bool = False) -> np.ndarray:
    """v_ext(G) on the density sphere: the erf-screened Coulomb local
    potential of the atoms, -4 pi Z / (V G^2) exp(-G^2 rc^2 / 4) S(G), zero
    at G = 0 (the same local potential ks_potential puts into v_KS);
    separable = True takes S(G) from per-axis tables (large cells)."""
    V = sysm.volume
    S = (_structure_factor_sep(dens.k, sysm.L, sysm.pos) if separable
         else _structure_factor(dens.k.astype(np.float64), sysm.L, sysm.pos))
    g2 = dens.ak2
    out = np.zeros(dens.ncoef, dtype=np.complex128)
    nz = g2 > 0
    out[nz] = (-4.0 * math.pi * sysm.z_val / (V * g2[nz]) * np.exp(-g2[nz] * sysm.rc_loc ** 2 / 4.0)
               * S[nz])
    return out


def noise_orbitals(sysm: System, nstate: int, seed: int, first: int = 0) -> np.ndarray:
    """Stochastic orbitals chi_s(G), complex (nstate, ncoef), G=0 last.

    Real-space Rademacher values +-sqrt(N) (stodft/gen-noise.c:186-195 uses
    ranValue = sqrt(nfft2)) transformed with c(G) = (1/N) sum_r chi(r) e^{-iG.r},
    which gives E[c c^dagger] = identity on the coefficient set; G=0 is real."""
    n1, n2, n3 = sysm.n
    N = sysm.ngrid
    out = np.empty((nstate, sysm.ncoef), dtype=np.complex128)
    root = math.sqrt(N)
    for s in range(nstate):
        chi = rademacher_signs(seed, first + s, N) * root
        F = np.fft.fftn(chi.reshape(n1, n2, n3)).ravel() / N
        c = F[sysm.map]
        c[-1] = F[0].real
        out[s] = c
    return out


# ---------------------------------------------------------------------------
# The benchmark configurations (cell, grid, orbitals, degree, temperature)
# ---------------------------------------------------------------------------

@dataclasses.dataclass
class Workload:
    name: str
    a0: float
    rep: int
    n: int
    nstate: int
    block: int
    degree: int
    kt: float
    seed: int


WORKLOADS = {
    "si8": Workload("si8-32^3-B64-L256", 10.26, 1, 32, 64, 64, 256, 0.01, 0),
    "si64": Workload("si64-64^3-B256-L1024", 10.26, 2, 64, 256, 256, 1024, 0.01, 1),
    "si512": Workload("si512-128^3-B512x128-L2048", 10.26, 4, 128, 512, 128, 2048, 0.01, 2),
    "si1728": Workload("si1728-192^3-B1024x128-L3000", 10.26, 6, 192, 1024, 128, 3000, 0.005, 4),
}


def displaced_positions(pos: np.ndarray, sigma: float, seed: int) -> np.ndarray:
    """Atoms displaced by independent N(0, sigma^2) offsets per coordinate
    (numpy PCG64 stream of `seed`), as the Si1728 force workload specifies."""
    rng = np.random.default_rng(seed)
    return np.asarray(pos, dtype=np.float64) + rng.normal(0.0, sigma, size=np.shape(pos))


# The large workloads apply their KB projectors in real space (the reference's
# nlpp_real_on path): s + 3 p Gaussian channels per atom sampled within this
# cutoff (8217 grid points per atom at the 0.3206-bohr spacing of both cells).
WORKLOAD_RS_CUTOFF = 4.0


def workload_system(key: str) -> tuple:
    """(System, RealSpaceProjectors) of the Si512 / Si1728 workloads: the
    4x4x4 / 6x6x6 diamond supercell (a = 41.04 / 61.56 bohr), 12 Ry, 128^3 /
    192^3 grid, Si1728's atoms displaced by N(0, 0.05 bohr) from seed 4; the
    potential from the per-axis structure-factor tables, and the projectors in
    real space (the System carries no G-space table)."""
    wl = WORKLOADS[key]
    L, p = diamond_supercell(wl.a0, wl.rep)
    if key == "si1728":
        p = displaced_positions(p, 0.05, 4)
    sysm = build_system(wl.name, wl.a0, wl.rep, wl.n, pos=p, with_nl=False, separable_sf=True)
    return sysm, real_space_projectors(sysm, WORKLOAD_RS_CUTOFF)


# ---------------------------------------------------------------------------
# Fragment batch: many small systems cut from the Si1728 cell
# ---------------------------------------------------------------------------

FRAGMENT_GRIDS = (24, 32, 36, 48)


@dataclasses.dataclass
class FragmentSpec:
    index: int
    natm: int
    cell: float
    n: int
    nstate: int
    pos: np.ndarray


def fragment_specs(nfrag: int = 343, seed: int = 3, natm_range: tuple = (8, 64),
                   g2max: float = 12.0, a0: float = 10.26, rep: int = 6,
                   grids: tuple = FRAGMENT_GRIDS) -> list:
    """Fragments of the Si1728 (6x6x6 diamond, a = 61.56 bohr) cell: fragment
    i takes natm ~ U[8, 64] atoms (seeded) nearest, in the max-norm, to a
    random centre -- a cube-shaped cluster -- in a cubic box of the bulk
    density (side a0 (natm/8)^(1/3)), positions wrapped into the box.  Its
    grid is the smallest of `grids` holding the density sphere (n >= 4K + 1,
    K = max |k| of the wavefunction sphere) and its orbital count grows
    linearly from 32 (8 atoms) to 128 (64 atoms), kept even."""
    rng = np.random.default_rng(seed)
    Lbig, pbig = diamond_supercell(a0, rep)
    out = []
    for i in range(nfrag):
        natm = int(rng.integers(natm_range[0], natm_range[1] + 1))
        centre = rng.random(3) * Lbig
        d = pbig - centre
        d -= Lbig * np.round(d / Lbig)
        idx = np.argsort(np.abs(d).max(axis=1), kind="stable")[:natm]
        side = a0 * (natm / 8.0) ** (1.0 / 3.0)
        pos = np.mod(d[idx] + 0.5 * side, side)
        K = int(math.floor(math.sqrt(g2max) * side / (2.0 * math.pi)))
        fit = [g for g in grids if g >= 4 * K + 1]
        n = min(fit) if fit else max(grids)
        frac = (natm - natm_range[0]) / float(natm_range[1] - natm_range[0])
        nstate = 32 + 2 * int(round(frac * 48))
        out.append(FragmentSpec(i, natm, side, n, nstate, pos))
    return out


def build_fragment(spec: FragmentSpec, g2max: float = 12.0, with_nl: bool = True) -> System:
    return build_system("frag%d-%da-%d^3" % (spec.index, spec.natm, spec.n), 10.26, 1, spec.n,
                        g2max=g2max, pos=spec.pos, with_nl=with_nl, cell=spec.cell,
                        separable_sf=True)
