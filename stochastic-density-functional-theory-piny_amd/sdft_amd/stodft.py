"""The reference's filter-path operations, driven over the C ABI.

Names follow stodft/ in the reference: genEnergyMax / genEnergyMin
(calc-spectral-range.c), calcChebyCoeff (coeff.c), calcChemPotCheby
(calc-chempot-chebyshev.c), filterChebyPolyHerm (filter.c),
calcRhoStoHybridCheby (density.c), calcEnergyChemPot (calc-energy.c) and the
orchestration of genStoOrbitalCheby (gen-stodft-wf.c:356-694).  All the
arithmetic runs in libsdft_hip; this module only sequences calls.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from . import hip as _hip


def _real_space_nl_bounds(rs, ngrid) -> tuple[float, float]:
    """Bounds of the real-space KB operator V = vol_elem B E B^T (B: grid x
    projectors, the sampled beta_p): its eigenvalues are those of
    vol_elem E^(1/2) B^T B E^(1/2) (|E| and signs for negative energies), and
    restricting V to the plane-wave set keeps them inside; Gershgorin's
    row sums of that projector-space matrix bound them."""
    import scipy.sparse as sp
    nproj = len(rs.energy)
    rows = np.repeat(np.arange(nproj), np.diff(rs.offsets))
    B = sp.csr_matrix((rs.values[:rs.offsets[-1]], (rows, rs.index[:rs.offsets[-1]])),
                      shape=(nproj, ngrid))
    gram = (B @ B.T).tocsr()
    s = np.sqrt(np.abs(rs.energy) * rs.vol_elem)
    m = sp.diags(s) @ gram @ sp.diags(s)
    absrow = np.asarray(abs(m).sum(axis=1)).ravel()
    diag = m.diagonal()
    off = absrow - np.abs(diag)
    sign = np.sign(rs.energy)
    hi = max(0.0, float(np.max(np.where(sign > 0, diag + off, off))))
    lo = min(0.0, float(np.min(np.where(sign < 0, -(diag + off), -off)))) if np.any(sign < 0) \
        else 0.0
    return lo, hi


def spectral_bounds(sysm, rs=None) -> tuple[float, float]:
    """Rigorous bounds of the spectrum of H = T + v + V_NL on the coefficient
    set: T in [0, max|G|^2/2], v in [min v, max v], and V_NL = B E B^T lies
    in [lambda_min, lambda_max] of |E|^(1/2) Gram |E|^(1/2) (signed by E),
    Gram_pq = Q_p(0)Q_q(0) + 2 Re sum_half conj(Q_p) Q_q; with real-space
    projectors rs (systems.RealSpaceProjectors) instead, the Gershgorin
    bounds of their grid-space operator."""
    lo_nl = hi_nl = 0.0
    if rs is not None:
        lo_nl, hi_nl = _real_space_nl_bounds(rs, sysm.ngrid)
    elif sysm.nproj > 0:
        q = sysm.proj
        gram = np.outer(q[:, -1].real, q[:, -1].real) + 2.0 * np.real(np.conj(q[:, :-1]) @ q[:, :-1].T)
        e = sysm.proj_energy
        s = np.sqrt(np.abs(e))
        m = (s[:, None] * gram * s[None, :]) * np.sign(e)[None, :]
        # eigenvalues of E Gram (similar to the symmetric form for E >= 0)
        ev = np.linalg.eigvals(m).real
        lo_nl, hi_nl = min(0.0, float(ev.min())), max(0.0, float(ev.max()))
    vmin, vmax = float(sysm.vks.min()), float(sysm.vks.max())
    return vmin + lo_nl, 0.5 * float(sysm.ak2.max()) + vmax + hi_nl


def _hpsi(ctx, c, prec):
    f = ctx.hpsi_raw(c, prec)
    h = -0.25 * f
    h[:, -1] = -0.5 * f[:, -1]
    return h


def _gdot(a, b):
    return 2.0 * (np.sum(a.real * b.real) + np.sum(a.imag * b.imag)) - a[..., -1].real @ b[..., -1].real


def gen_energy_max(ctx, niter=100000, seed=7, pad=0.1, conv=1e-5, prec=_hip.PREC_FP64):
    """genEnergyMax (calc-spectral-range.c:33-333) on the device
    (sdft_hip_spectral_bound): power iteration on H until the Rayleigh
    quotient changes by <= conv, padded by +0.1 Ha."""
    from . import systems
    v = systems.noise_orbitals(ctx.sys, 1, seed)[0]
    e, _, _ = ctx.spectral_bound(v, 0, conv=conv, max_iter=niter, prec=prec)
    return e + pad


def gen_energy_min(ctx, emax, niter=100000, seed=8, pad=0.1, conv=1e-5, prec=_hip.PREC_FP64):
    """genEnergyMin (calc-spectral-range.c:339-566) on the device: power
    iteration on (E_max - H); the converged <phi|H|phi> padded by -0.1 Ha."""
    from . import systems
    v = systems.noise_orbitals(ctx.sys, 1, seed)[0]
    e, _, _ = ctx.spectral_bound(v, 1, emax=emax, conv=conv, max_iter=niter, prec=prec)
    return e - pad


@dataclasses.dataclass
class FilterResult:
    mu: float
    ne: float
    coeffs: np.ndarray
    moments: np.ndarray
    filtered: np.ndarray
    emin: float
    emax: float


def gen_sto_orbital_cheby(ctx, chi, L, beta, ne_true, mu_init, gap, emin=None, emax=None,
                          occ=2.0, prec=_hip.PREC_FP64):
    """genStoOrbitalCheby (gen-stodft-wf.c:356-694) with chem_pot_opt = cheby:
    spectral range -> moments -> mu -> sqrt(f) coefficients -> filter."""
    if emin is None or emax is None:
        emin, emax = spectral_bounds(ctx.sys)
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    nsto = chi.shape[0]
    mom = ctx.cheby_moments(chi, L, scale, emean, prec)
    mu, ne = ctx.chem_pot(L, emin, emax, beta, mom, nsto, ne_true, mu_init, gap, occ)
    coeffs = ctx.cheby_coeffs(L, emin, emax, mu, beta, 2)
    acc, _ = ctx.cheby_filter(chi, coeffs, scale, emean, prec)
    return FilterResult(mu, ne, coeffs, mom, acc, emin, emax)
