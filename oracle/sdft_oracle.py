"""CPU FP64 oracle for PINY's stochastic-DFT Chebyshev filter path.

TEST INFRASTRUCTURE ONLY.  Only `tests/`, `__graft_entry__.smoke()` and the
`cpu_baseline` leg of `bench.py` may import this module, and only as the
checker (or as the timed CPU baseline).  The product path never calls it.

Each function restates one piece of the reference algorithm (cited by
file:line under the reference tree) in plain numpy, FP64 throughout, with
orbitals held as complex arrays (nstate, ncoef) whose last entry is G=0
(real).  The plane-wave set, grid maps, potentials and projector tables are
inputs (built by the package's system builder); the oracle restates what the
filter path does with them.

Parity status of each function (how it is pinned; see tests/):
  pack_pair / unpack_pair_sub / hpsi_raw   : brute-force O(N^2) DFT on tiny
      grids, plane-wave eigenvector closed form (H c = (|G|^2/2 + v0) c),
      Hermiticity of H
  nl_apply_gspace                           : Hermiticity, gradient of the NL
      energy by finite differences, projector table vs real-space quadrature
  nl_apply_real                             : agreement with the G-space form
      built from the analytic Bessel transform (two independent forms),
      Hermiticity
  gen_energy_bound                          : converges to the extreme
      eigenvalues of the dense H (eigvalsh)
  newton_filter / newton_coeffs / leja_order : scalar Newton polynomial on
      exact eigenvectors for every mu column; agreement with the Chebyshev
      filter of the same smooth function to the interpolation error
  norm_h_cheby / filter_cheby               : scalar recurrence on exact
      eigenvectors, dense-H spectral oracle (sqrt f(H) chi by eigh)
  cheby_coeffs / round_fft                  : direct cosine sums, fit-error
      criterion of testChebyCoeff, roundFFT closed cases
  cheby_moments / num_elec_cheby / chem_pot_cheby : doubling identities vs
      direct <chi|T_k|chi>, eigenvalue trace on dense H, monotone N(mu)
  density_real                              : sum(rho)/N = occ <chi|chi>/Nsto
      identity and the brute-force real-space sum
  energy_ke / energy_nl / nl_forces         : closed forms and FD derivatives
  nl_forces_real                            : central differences of energy_nl_real
      with the projectors resampled at displaced atoms
  energy_nl_real                            : equals occ/Nsto sum_s <psi_s|V_NL|psi_s>
      with V_NL psi from nl_apply_real through hpsi_raw (Gamma dot of the
      raw-force difference)
  local_forces                              : central differences of E_ext
      (ks_potential's external energy) w.r.t. the atom positions
  xc_lda / ks_potential                     : published correlation energies
      (Ceperley-Alder points), exchange closed form, PZ branch continuity,
      v_xc = d(rho eps_xc)/d rho by finite differences, Hartree energy and
      potential of a plane-wave density in closed form
"""
from __future__ import annotations

import math

import numpy as np

try:  # optional multi-threaded FFT for the timed CPU baseline
    import scipy.fft as _sfft
except Exception:  # pragma: no cover
    _sfft = None

_FFT_WORKERS = 1


def set_fft_workers(n: int) -> None:
    """Threads used by the oracle's 3D FFTs (scipy.fft); 1 = numpy."""
    global _FFT_WORKERS
    _FFT_WORKERS = max(1, int(n))


def _ifftn_unnorm(z):
    """FFTW_BACKWARD (e^{+iGr}, unnormalised) as planned at
    mathlib/fft_create_package.c:1685-1687."""
    if _sfft is not None and _FFT_WORKERS > 1:
        return _sfft.ifftn(z, norm="forward", workers=_FFT_WORKERS)
    return np.fft.ifftn(z) * z.size


def _fftn(z):
    """FFTW_FORWARD (e^{-iGr}, unnormalised)."""
    if _sfft is not None and _FFT_WORKERS > 1:
        return _sfft.fftn(z, workers=_FFT_WORKERS)
    return np.fft.fftn(z)


# ---------------------------------------------------------------------------
# Pack / FFT / unpack  (mathlib/fft_package.c)
# ---------------------------------------------------------------------------

def pack_pair(c1, c2, mp, mpc, shape):
    """dble_pack_coef_fftw3d_filter, fft_package.c:3195-3254: zero the grid,
    Z[map(G)] = c1(G) + i c2(G), Z[mapC(G)] = c1(G)* + i c2(G)*,
    Z[0] = c1re(0) + i c2re(0)."""
    N = int(np.prod(shape))
    z = np.zeros(N, dtype=np.complex128)
    a, b = c1[:-1], c2[:-1]
    z[mp[:-1]] = (a.real - b.imag) + 1j * (a.imag + b.real)
    z[mpc[:-1]] = (a.real + b.imag) + 1j * (-a.imag + b.real)
    z[0] = c1[-1].real + 1j * c2[-1].real
    return z.reshape(shape)


def to_r(z):
    """para_fft_gen3d_fwd_to_r_fftw3d_filter, fft_package.c:3333-3401."""
    return _ifftn_unnorm(z)


def to_g(z):
    """para_fft_gen3d_bck_to_g_fftw3d_filter, fft_package.c:3407-3504."""
    return _fftn(z)


def unpack_pair_sub(f1, f2, F, mp, mpc):
    """dble_upack_coef_sum_fftw3d_filter, fft_package.c:3513-3599 (FFTW
    branch :3573-3595): with F the forward-FFT output scaled by 1/N,
      f1re -= 2(F(G)+F(-G))re   f1im -= 2(F(G)-F(-G))im
      f2re -= 2(F(G)+F(-G))im   f2im -= 2(-F(G)+F(-G))re
    and at G=0 f1re -= 2 F(0)re, f2re -= 2 F(0)im."""
    Fl = F.ravel() / F.size
    t = Fl[mp[:-1]]
    tc = Fl[mpc[:-1]]
    f1[:-1] -= 2.0 * ((t.real + tc.real) + 1j * (t.imag - tc.imag))
    f2[:-1] -= 2.0 * ((t.imag + tc.imag) + 1j * (-t.real + tc.real))
    f1[-1] -= 2.0 * Fl[0].real
    f2[-1] -= 2.0 * Fl[0].imag


# ---------------------------------------------------------------------------
# Nonlocal Kleinman-Bylander (G-space)
# ---------------------------------------------------------------------------

def nl_projections(C, proj):
    """Z[s, p] = sum_{all G} c_s(G) Q_p(G) = c(0)Q(0) + 2 sum_half Re(c Q).

    The G-space KB matrix of controlEwdNonlocFilter / getNlmatFilter
    (stodft/reduced-energy/cp-energy-eext-stodft.c:865-946,1105-1168) with the
    structure factor, radial transform and Ylm folded into the table Q."""
    if proj.shape[0] == 0:
        return np.zeros((C.shape[0], 0))
    Zh = 2.0 * (C[:, :-1].real @ proj[:, :-1].real.T - C[:, :-1].imag @ proj[:, :-1].imag.T)
    Z0 = np.outer(C[:, -1].real, proj[:, -1].real)
    return Zh + Z0


def nl_apply_gspace(C, proj, energy):
    """Raw NL force f = -d<psi|V_NL|psi>/dc* with the reference's scaling
    (-4 H c for G != 0, -2 H c at G = 0; energy-wrapper-scf.c:1072-1076,
    getnl_fcoef energy/cp/cp_energy_eext_nonloc.c:1669+):
      f(G) = -4 sum_p E_p Z_p conj(Q_p(G))  (G != 0),  f(0) = -2 sum_p E_p Z_p Q_p(0)."""
    f = np.zeros_like(C)
    if proj.shape[0] == 0:
        return f
    Z = nl_projections(C, proj) * energy[None, :]
    f[:, :-1] = -4.0 * (Z @ np.conj(proj[:, :-1]))
    f[:, -1] = -2.0 * (Z @ proj[:, -1].real)
    return f


def nl_energy_per_state(C, proj, energy):
    """<psi_s|V_NL|psi_s> = sum_p E_p Z_p^2 (sumNlPot, cp-energy-eext-stodft.c:650-696)."""
    if proj.shape[0] == 0:
        return np.zeros(C.shape[0])
    Z = nl_projections(C, proj)
    return (Z * Z * energy[None, :]).sum(axis=1)


def nl_apply_real(z, rs):
    """nlppKBRealEnergyFilter (energy/cp/cp_energy_eext_nonloc_real_filter.c:
    110-450) on one grid array z (the reference runs it on the real and the
    imaginary part of the packed pair separately, :36-105; the operator is
    real, so one pass over complex z is the same): per projector p the
    neighbourhood values are gathered (wfNbhd, :229-241), dot = volElem *
    sum beta_p psi * vpsNorm_p (:318-331), forceTemp += dot * beta_p (daxpy),
    and the per-atom results are scattered onto the grid (:430-441)."""
    out = np.zeros_like(z)
    for p in range(len(rs.energy)):
        sl = slice(rs.offsets[p], rs.offsets[p + 1])
        ix = rs.index[sl]
        b = rs.values[sl]
        dot = rs.vol_elem * (b @ z[ix]) * rs.energy[p]
        out[ix] += dot * b
    return out


# ---------------------------------------------------------------------------
# H.psi (raw force)
# ---------------------------------------------------------------------------

def hpsi_raw(C, sysm, with_nl=True, nl_real=None):
    """calcCoefForceWrapSCF (stodft/reduced-energy/energy-wrapper-scf.c:1066-1197)
    -> calcNonLocalPseudoScf (G-space KB) + coefForceCalcHybridSCF
    (cp-energy-ee-rho-stodft.c:38-355): per orbital pair pack -> G->r FFT ->
    v.psi (cp_vpsi_threads, cp_energy_ee_rho.c:2817-2845) -> r->G FFT ->
    unpack-subtract; an odd last orbital is paired with zero (identical math
    to the single pack); kinetic f -= 2|G|^2 c (G != 0) and f_im(0) = 0.
    With nl_real (systems.RealSpaceProjectors) the KB term is the real-space
    one, added to v.psi on the grid (controlPotentialEnergyFilter,
    cp_energy_eext_nonloc_real_filter.c:66-100), instead of the G-space one."""
    nstate, ncoef = C.shape
    shape = tuple(sysm.n)
    v = sysm.vks.reshape(shape)
    f = np.zeros_like(C)
    if with_nl and nl_real is None and sysm.nproj > 0:
        f += nl_apply_gspace(C, sysm.proj, sysm.proj_energy)
    zero = np.zeros(ncoef, dtype=np.complex128)
    for s in range(0, nstate, 2):
        c1 = C[s]
        c2 = C[s + 1] if s + 1 < nstate else zero
        z = to_r(pack_pair(c1, c2, sysm.map, sysm.map_conj, shape))
        if with_nl and nl_real is not None:
            z = z * v + nl_apply_real(z.ravel(), nl_real).reshape(shape)
        else:
            z *= v
        F = to_g(z)
        f1 = f[s]
        f2 = f[s + 1] if s + 1 < nstate else zero.copy()
        unpack_pair_sub(f1, f2, F, sysm.map, sysm.map_conj)
    f[:, :-1] -= 2.0 * sysm.ak2[None, :-1] * C[:, :-1]
    f[:, -1] = f[:, -1].real
    return f


def hpsi(C, sysm, with_nl=True):
    """Plain H c from the raw force: H c = -f/4 (G != 0), -f/2 (G = 0)."""
    f = hpsi_raw(C, sysm, with_nl)
    h = -0.25 * f
    h[:, -1] = -0.5 * f[:, -1]
    return h


def gamma_dot(a, b):
    """Gamma-point inner product 2 sum_{G!=0} Re(a* b) + a0 b0 (coded in the
    reference as 2 ddot over all entries minus a0 b0,
    calc-chempot-chebyshev.c:725-729)."""
    return 2.0 * (np.sum(a.real * b.real, axis=-1) + np.sum(a.imag * b.imag, axis=-1)) \
        - a[..., -1].real * b[..., -1].real


# ---------------------------------------------------------------------------
# Spectral range
# ---------------------------------------------------------------------------

def gen_energy_bound(v0, sysm, which, emax=0.0, conv=1e-5, max_iter=100000, with_nl=True):
    """genEnergyMax (which 0) / genEnergyMin (which 1), stodft/calc-spectral-
    range.c:33-333 / :339-566: normalise phi; energyOld = 1e7, energy = 1e5;
    while |energyOld - energy| > conv: H phi from the raw force (x -1/4,
    G = 0 x -1/2), energy = <phi|H phi> (Gamma dot), phi = H phi (or
    emax phi - H phi), normalised.  Returns (energy, phi, iterations); the
    reference's bounds are energy + 0.1 / energy - 0.1."""
    phi = np.array(v0, dtype=np.complex128).reshape(1, -1)
    phi /= math.sqrt(gamma_dot(phi, phi).sum())
    e_old, e = 1.0e7, 1.0e5
    it = 0
    while abs(e_old - e) > conv and it < max_iter:
        f = hpsi_raw(phi, sysm, with_nl)
        h = -0.25 * f
        h[:, -1] = -0.5 * f[:, -1]
        e_old = e
        e = float(gamma_dot(h, phi).sum())
        w = h if which == 0 else emax * phi - h
        phi = w / math.sqrt(gamma_dot(w, w).sum())
        it += 1
    return e, phi[0], it


# ---------------------------------------------------------------------------
# Chebyshev recurrence and filter
# ---------------------------------------------------------------------------

def norm_h_cheby(c, f, t_prev, k, scale, emean):
    """normHCheby recurrence part, stodft/normh.c:259-338 (in place on f):
    k == 1: f = s1 f + p c (G != 0), f_re(0) = s2 f_re(0) + p c_re(0)
    k > 1 : f = 2 s1 f + 2 p c - t_prev, likewise at G=0 with 2 s2
    with s1 = -scale/4, s2 = -scale/2, p = -scale*emean; f_im(0) is left as
    the raw force has it (0)."""
    s1 = -scale * 0.25
    s2 = -scale * 0.5
    p = -scale * emean
    if k == 1:
        f[:, :-1] = s1 * f[:, :-1] + p * c[:, :-1]
        f[:, -1] = (s2 * f[:, -1].real + p * c[:, -1].real) + 1j * f[:, -1].imag
    else:
        f[:, :-1] = 2.0 * s1 * f[:, :-1] + 2.0 * p * c[:, :-1] - t_prev[:, :-1]
        f[:, -1] = (2.0 * s2 * f[:, -1].real + 2.0 * p * c[:, -1].real - t_prev[:, -1].real) \
            + 1j * f[:, -1].imag
    return f


def filter_cheby(chi, coeffs, sysm, scale, emean, with_nl=True, moments=False, nl_real=None):
    """filterChebyPolyHerm, stodft/filter.c:318-706 (numChemPot = 1):
    acc = c_0 chi; for k = 1..L-1: T_k by normHCheby, acc += c_k T_k.
    Returns (acc, T_{L-1}) -- on return the reference leaves cre_up = T_{L-1}.
    With moments=True also returns the optional moment accumulation of
    filter.c:458-510,572-594 (sum_s <chi_s|T_k chi_s>, k < L)."""
    L = len(coeffs)
    acc = coeffs[0] * chi
    c = chi.copy()
    t_prev = chi.copy()
    mom = np.zeros(L)
    if moments:
        mom[0] = gamma_dot(chi, chi).sum()
    for k in range(1, L):
        f = hpsi_raw(c, sysm, with_nl, nl_real)
        norm_h_cheby(c, f, t_prev, k, scale, emean)
        if k > 1:
            t_prev = c
        c = f
        if moments:
            mom[k] = gamma_dot(chi, c).sum()
        acc = acc + coeffs[k] * c
    if moments:
        return acc, c, mom
    return acc, c


# ---------------------------------------------------------------------------
# Chebyshev coefficients (stodft/coeff.c)
# ---------------------------------------------------------------------------

def newton_filter(chi, coeffs, samp_point, sysm, scale, emean, with_nl=True, nl_real=None):
    """filterNewtonPolyHerm (stodft/filter.c:33-312) with normHNewtonHerm
    (stodft/normh.c:35-176): P_0 = chi; P_n = (scale (H - emean) - z_{n-1})
    P_{n-1}, computed from the raw force as f * (-scale/4) (G != 0, real part
    at G = 0 * -scale/2) + (-scale emean - z) P (normh.c:101-136); out[mu] =
    sum_n coeffs[n, mu] P_n.  Returns (out (nmu, nstate, ncoef), P_{L-1})."""
    co = np.asarray(coeffs, dtype=np.float64)
    if co.ndim == 1:
        co = co[:, None]
    L, nmu = co.shape
    P = chi.copy()
    out = np.stack([co[0, m] * chi for m in range(nmu)])
    for n in range(1, L):
        f = hpsi_raw(P, sysm, with_nl, nl_real)
        pref = -scale * emean - samp_point[n - 1]
        nxt = -0.25 * scale * f + pref * P
        nxt[:, -1] = (-0.5 * scale * f[:, -1].real + pref * P[:, -1].real) \
            + 1j * (f[:, -1].imag + pref * P[:, -1].imag)
        P = nxt
        for m in range(nmu):
            out[m] += co[n, m] * P
    return out, P


def newton_coeffs(fun, nodes):
    """Newton divided differences of fun at nodes (the interpolation form
    the Newton filter evaluates: p(x) = sum_n a_n prod_{j<n} (x - z_j))."""
    z = np.asarray(nodes, dtype=np.float64)
    a = np.array([fun(x) for x in z], dtype=np.float64)
    for k in range(1, len(z)):
        a[k:] = (a[k:] - a[k - 1:-1]) / (z[k:] - z[:-k])
    return a


def leja_order(nodes):
    """Leja ordering (start at the largest |z|, then maximise the product of
    distances), the stable order for Newton interpolation on an interval."""
    z = list(np.asarray(nodes, dtype=np.float64))
    out = [max(z, key=abs)]
    z.remove(out[0])
    while z:
        nxt = max(z, key=lambda x: np.prod([abs(x - y) for y in out]))
        out.append(nxt)
        z.remove(nxt)
    return np.array(out)


def newton_samp_points(L, energy_min, newton_scale, smin=-2.0, smax=2.0):
    """genSampNewtonHermit, stodft/coeff.c:423-533: 32 L equally spaced
    candidates on [Smin, Smax] (Smin = -2, Smax = 2, init.c:240-241,442-443),
    the first sample point is Smin, then each next point is the candidate
    that maximises sum_j log((x - z_j)^2) over the points already chosen (a
    discrete Leja sequence; a candidate closer than 1e-10 to a chosen point
    scores -1e30, ties go to the lowest index as the reference's per-thread
    strict-greater scan and thread-order merge do).  newton_scale is
    newtonInfo->scale = (Smax - Smin)/energyDiff (gen-stodft-wf.c:216-220);
    returns (sampPoint, sampPointUnscale) with sampPointUnscale =
    (sampPoint - Smin)/newton_scale + energyMin."""
    ncand = 32 * L
    delta = (smax - smin) / float(ncand - 1)
    cand = smin + np.arange(ncand) * delta
    samp = np.empty(L)
    samp[0] = cand[0]

    def score(z):
        d = cand - z
        d = d * d
        out = np.full(ncand, -1.0e30)
        ok = d >= 1.0e-20
        out[ok] = np.log(d[ok])
        return out

    obj = score(samp[0])
    for i in range(1, L):
        k = int(np.argmax(obj))
        if not obj[k] > -100000.0:  # the reference's scan starts from -1e5
            raise ValueError("no admissible sample candidate left")
        samp[i] = cand[k]
        obj += score(samp[i])
    unscale = (samp - smin) * (1.0 / newton_scale) + energy_min
    return samp, unscale


def newton_hermite_coeffs(samp, unscale, chem_pots, beta):
    """genCoeffNewtonHermit, stodft/coeff.c:199-420, without energy windows
    (energyWindowOn = 0) and with SQRTFERMI: for each chemical potential the
    Newton divided differences of sqrt(f_FD) at the sample points, by the
    reference's forward substitution (a_0 = g(x_0), a_1 = (g_1 - a_0)/(z_1 -
    z_0), a_i = (g_i - a_0 - sum_{j=1}^{i-1} a_j prod_{m<j}(z_i - z_m)) /
    prod_{m<i}(z_i - z_m), accumulated in the reference's order).  Returns
    expanCoeff as (L, nmu): row iPoly holds expanCoeff[iPoly*numChemPot + imu]."""
    z = [float(v) for v in samp]
    L = len(z)
    mus = np.atleast_1d(np.asarray(chem_pots, dtype=np.float64))
    out = np.zeros((L, mus.size))
    for imu, mu in enumerate(mus):
        g = [math.sqrt(float(fermi(x, mu, beta))) for x in unscale]
        a = [0.0] * L
        a[0] = g[0]
        if L > 1:
            a[1] = (g[1] - a[0]) / (z[1] - z[0])
        for i in range(2, L):
            zi = z[i]
            sm = g[i] - a[0]
            prod = 1.0
            for j in range(1, i):
                prod *= zi - z[j - 1]
                sm -= a[j] * prod
            prod *= zi - z[i - 1]
            a[i] = sm / prod
        out[:, imu] = a
    return out


def newton_fit_error(coeffs, samp, emin, emax, newton_scale, chem_pots, beta, npoint=1000):
    """calcFitErrorNewton, stodft/coeff.c:644-760 (energyWindowOn = 0): the
    largest |sum_n a_n prod_{j<n}(s - z_j) - sqrt(f_FD(x))| over npoint
    midpoints x_i = emin + (i + 1/2)(emax - emin)/npoint, s = (x - emean)
    newton_scale, over all chemical potentials."""
    co = np.asarray(coeffs, dtype=np.float64)
    if co.ndim == 1:
        co = co[:, None]
    mus = np.atleast_1d(np.asarray(chem_pots, dtype=np.float64))
    emean = 0.5 * (emin + emax)
    x = emin + (np.arange(npoint) + 0.5) * (emax - emin) / npoint
    s = (x - emean) * newton_scale
    err = -1.0e30
    for imu, mu in enumerate(mus):
        val = np.full(npoint, co[0, imu])
        prod = np.ones(npoint)
        for n in range(1, co.shape[0]):
            prod = prod * (s - samp[n - 1])
            val = val + co[n, imu] * prod
        ref = np.sqrt(fermi(x, mu, beta))
        err = max(err, float(np.max(np.abs(val - ref))))
    return err


def round_fft(n_in: int) -> int:
    """roundFFT, coeff.c:1556-1600: nearest 2^n 3^m to n_in by the
    reference's search (2^nmax first, then 2^i 3^(m+1) with m = floor of the
    log ratio, strict improvement only)."""
    ln2, ln3 = math.log(2.0), math.log(3.0)
    lnn = math.log(n_in)
    nmax = int(lnn / ln2) + 1
    best = abs(2.0 ** nmax - n_in)
    n, m = nmax, 0
    for i in range(nmax):
        x = max((lnn - i * ln2) / ln3, 0.0)
        mnow = int(x) + 1
        diff = abs(2.0 ** i * 3.0 ** mnow - n_in)
        if diff < best:
            m, n, best = mnow, i, diff
    return (2 ** n) * (3 ** m)


def fermi(x, mu, beta):
    """fermiExpReal, mathlib/mathlib.c:781-785."""
    with np.errstate(over="ignore"):
        return 1.0 / (1.0 + np.exp(beta * (x - mu)))


def cheby_coeffs(L, emin, emax, mu, beta, fun_flag=2, grid=None):
    """calcChebyCoeff, coeff.c:1093-1223 for funFlag 1 (f) and 2 (sqrt f):
    x_j = (dE/2) cos(2 pi j/M) + E_mean, M = roundFFT(32 L);
    c_k = 2 Re(DFT_k)/M (k >= 1), c_0 = Re(DFT_0)/M."""
    M = grid if grid is not None else round_fft(32 * L)
    dE = (emax - emin) * 0.5
    em = 0.5 * (emax + emin)
    j = np.arange(M)
    x = dE * np.cos(2.0 * math.pi * j / M) + em
    fv = fermi(x, mu, beta)
    if fun_flag == 2:
        fv = np.sqrt(fv)
    X = np.fft.fft(fv)
    c = 2.0 * X[:L].real / M
    c[0] = X[0].real / M
    return c


def cheby_eval(coeffs, y):
    """sum_k c_k T_k(y) by the three-term recurrence (testChebyCoeff,
    coeff.c:1319-1460)."""
    t0 = np.ones_like(y)
    out = coeffs[0] * t0
    if len(coeffs) == 1:
        return out
    t1 = y.copy()
    out = out + coeffs[1] * t1
    for k in range(2, len(coeffs)):
        t2 = 2.0 * y * t1 - t0
        out = out + coeffs[k] * t2
        t0, t1 = t1, t2
    return out


# ---------------------------------------------------------------------------
# Moments and chemical potential (stodft/calc-chempot-chebyshev.c)
# ---------------------------------------------------------------------------

def cheby_moments(chi, sysm, L, scale, emean, with_nl=True, nl_real=None):
    """calcChebyMoments, calc-chempot-chebyshev.c:596-913 (summed over states,
    not yet normalised): mu_0 = <chi|chi>, mu_1 = <chi|T_1>, then for
    k = 2..L/2 one normHCheby per step and
      mu_k = <chi|T_k>, mu_{2k-1} = 2<T_k|T_{k-1}> - mu_1,
      mu_{2k} = 2<T_k|T_k> - mu_0
    (the direct value of mu_k, written last, wins for k < L/2+1).
    Returns mu_0..mu_L."""
    nmom = L // 2 + 1 if L % 2 == 0 else (L + 1) // 2
    mom = np.zeros(L + 1)
    mom[0] = gamma_dot(chi, chi).sum()
    c = chi.copy()
    f = hpsi_raw(c, sysm, with_nl, nl_real)
    norm_h_cheby(c, f, chi, 1, scale, emean)
    mom[1] = gamma_dot(chi, f).sum()
    t_prev = chi.copy()
    c = f
    for k in range(2, nmom):
        f = hpsi_raw(c, sysm, with_nl, nl_real)
        norm_h_cheby(c, f, t_prev, k, scale, emean)
        mom[k] = gamma_dot(chi, f).sum()
        mom[2 * k - 1] = 2.0 * gamma_dot(f, c).sum() - mom[1]
        mom[2 * k] = 2.0 * gamma_dot(f, f).sum() - mom[0]
        t_prev = c
        c = f
    return mom


def num_elec_cheby(mu, moments_norm, L, emin, emax, beta, occ=2.0):
    """calcNumElecCheby, calc-chempot-chebyshev.c:918-972:
    N(mu) = occ sum_{k<L} c_k(mu; f) mu_k."""
    c = cheby_coeffs(L, emin, emax, mu, beta, fun_flag=1)
    return occ * float(np.dot(c, moments_norm[:L]))


def chem_pot_cheby(moments, nsto, L, emin, emax, beta, ne_true, mu_init, gap,
                   occ=2.0, tol_rel=1e-11, tol_abs=None, max_iter=200):
    """calcChemPotCheby, calc-chempot-chebyshev.c:30-315: normalise moments by
    the stochastic orbital count; bracket [mu0 -+ gap/2], widen by gap until
    it brackets N_e; one linear-interpolation step; then bisection until
    |N(mu) - N_e| <= tol (1e-11 N_e in the reference)."""
    mnorm = moments / float(nsto)
    tol = tol_abs if tol_abs is not None else tol_rel * ne_true

    def N(mu):
        return num_elec_cheby(mu, mnorm, L, emin, emax, beta, occ)

    lo, hi = mu_init - 0.5 * gap, mu_init + 0.5 * gap
    nlo, nhi = N(lo), N(hi)
    it = 0
    while nhi < ne_true and it < max_iter:
        lo, hi = hi, hi + gap
        nlo, nhi = N(lo), N(hi)
        it += 1
    while nlo > ne_true and it < max_iter:
        hi, lo = lo, lo - gap
        nlo, nhi = N(lo), N(hi)
        it += 1
    mu = (ne_true - nlo) * (hi - lo) / (nhi - nlo) + lo
    nn = N(mu)
    it = 0
    while abs(nn - ne_true) > tol and it < max_iter:
        if nn > ne_true:
            hi, nhi = mu, nn
        if nn < ne_true:
            lo, nlo = mu, nn
        mu = 0.5 * (lo + hi)
        nn = N(mu)
        it += 1
    return mu, nn


# ---------------------------------------------------------------------------
# Density and energies
# ---------------------------------------------------------------------------

def density_real(C, sysm, occ_over_nsto):
    """rhoCalcRealStoHybrid (stodft/density.c:43-274) + the occ/N_sto scale of
    calcRhoStoHybridCheby (:1478): rho(r) = occ/N_sto sum_s psi_s(r)^2 with
    psi from the pair-packed G->r FFT (sum_rho, fft_package.c:2502-2516).
    Returned in the FFTW (a, b, c; c fastest) grid order."""
    nstate, ncoef = C.shape
    shape = tuple(sysm.n)
    rho = np.zeros(shape)
    zero = np.zeros(ncoef, dtype=np.complex128)
    for s in range(0, nstate, 2):
        c1 = C[s]
        c2 = C[s + 1] if s + 1 < nstate else zero
        z = to_r(pack_pair(c1, c2, sysm.map, sysm.map_conj, shape))
        rho += z.real ** 2 + z.imag ** 2
    return rho.ravel() * occ_over_nsto


def energy_ke(C, ak2, occ, nsto):
    """calc-energy.c:133-141,190: E_kin = occ/2 sum_s sum_{G!=0} 2|G|^2 |c|^2 / N_sto."""
    e = 2.0 * (ak2[None, :-1] * (C[:, :-1].real ** 2 + C[:, :-1].imag ** 2)).sum()
    return occ * 0.5 * e / nsto


def energy_nl(C, proj, energy, occ, nsto):
    """calc-energy.c:160-191: E_NL = occ sum_s <psi_s|V_NL|psi_s> / N_sto."""
    return occ * nl_energy_per_state(C, proj, energy).sum() / nsto


def energy_nl_real(C, sysm, rs, occ, nsto):
    """Real-space KB energy, nlppKBRealEnergy (energy/cp/cp_energy_eext_nonloc_real.c:
    137-372): per orbital the grid values psi(r) (the unnormalised G->r
    transform), per projector channel dot = volElem sum_r beta_p(r) psi(r)
    (:255-265, :284-286) and E += dot^2 vpsNorm_p / V (:270, :302; the caller's
    energy_p carries the factor 2 of the m > 0 rows), scaled by occ / N_sto
    (stodft/calc-energy.c:160-191).  Each packed pair's real and imaginary
    parts are its two orbitals."""
    nstate = C.shape[0]
    shape = tuple(sysm.n)
    zero = np.zeros(C.shape[1], dtype=np.complex128)
    e = 0.0
    for s in range(0, nstate, 2):
        c2 = C[s + 1] if s + 1 < nstate else zero
        z = to_r(pack_pair(C[s], c2, sysm.map, sysm.map_conj, shape)).ravel()
        for p in range(len(rs.energy)):
            sl = slice(rs.offsets[p], rs.offsets[p + 1])
            d = rs.vol_elem * (rs.values[sl] @ z[rs.index[sl]])
            e += (d.real ** 2 + d.imag ** 2) * rs.energy[p] / sysm.volume
    return occ * e / nsto


def nl_forces_real(C, sysm, rs, grad, occ, nsto):
    """Real-space KB Hellmann-Feynman forces, nlppKBRealEnergyForce
    (energy/cp/cp_energy_eext_nonloc_real_force.c:36-240): per orbital the
    grid values psi(r), per projector channel dot = volElem sum_r beta_p psi
    and dotDev_j = volElem sum_r (d beta_p / d r_j) psi (:163-209, the gradient
    tables vnlPhiD{x,y,z}AtomGrid of :483-494), and the force on the channel's
    atom (d beta / d R_a = -d beta / d r) F_j += 2 vpsNorm_p / V dot dotDev_j
    (the reference's 4 vpsNorm of the m > 0 rows is the caller's doubled
    energy_p), scaled by occ / N_sto.  grad: (3, nent) aligned with
    rs.values.  Returns (natm, 3)."""
    nstate = C.shape[0]
    shape = tuple(sysm.n)
    zero = np.zeros(C.shape[1], dtype=np.complex128)
    F = np.zeros((sysm.natm, 3))
    for s in range(0, nstate, 2):
        c2 = C[s + 1] if s + 1 < nstate else zero
        z = to_r(pack_pair(C[s], c2, sysm.map, sysm.map_conj, shape)).ravel()
        for p in range(len(rs.energy)):
            sl = slice(rs.offsets[p], rs.offsets[p + 1])
            zz = z[rs.index[sl]]
            dot = rs.vol_elem * (rs.values[sl] @ zz)
            for j in range(3):
                dev = rs.vol_elem * (grad[j, sl] @ zz)
                F[rs.atom[p], j] += 2.0 * rs.energy[p] / sysm.volume * (
                    dot.real * dev.real + dot.imag * dev.imag)
    return F * occ / nsto


def nl_forces(C, sysm, occ, nsto):
    """Hellmann-Feynman NL force on each atom (calc-nuclei-force.c:302-489,
    G-space branch): F_a = -dE_NL/dR_a with Q_p(G) = q_p(G) e^{iG.R_a}, so
    dZ_p/dR_a = -2 sum_half G Im(c(G) Q_p(G)); E_NL = occ/Nsto sum E_p Z^2."""
    proj, E = sysm.proj, sysm.proj_energy
    Z = nl_projections(C, proj)                     # (s, p)
    cq = C[:, None, :-1] * proj[None, :, :-1]       # (s, p, g)
    dZ = -2.0 * np.einsum("spg,gx->spx", cq.imag, sysm.gvec[:-1])
    contrib = 2.0 * (E[None, :, None] * Z[:, :, None] * dZ).sum(axis=0)  # (p, 3)
    F = np.zeros((sysm.natm, 3))
    np.add.at(F, sysm.proj_atom, -contrib)
    return F * occ / nsto


# ---------------------------------------------------------------------------
# Dense references for tiny systems
# ---------------------------------------------------------------------------

def gamma_real_basis_dim(ncoef):
    return 2 * ncoef - 1


def to_real_vec(C):
    """Orthonormal real coordinates of Gamma coefficient vectors:
    x = [sqrt2 Re c(G!=0), sqrt2 Im c(G!=0), c(0)], so x.y = gamma_dot."""
    r2 = math.sqrt(2.0)
    return np.concatenate([r2 * C[..., :-1].real, r2 * C[..., :-1].imag, C[..., -1:].real], axis=-1)


def from_real_vec(x, ncoef):
    r2 = math.sqrt(2.0)
    m = ncoef - 1
    c = (x[..., :m] + 1j * x[..., m:2 * m]) / r2
    return np.concatenate([c, x[..., 2 * m:2 * m + 1] + 0j], axis=-1)


def dense_hamiltonian(sysm, with_nl=True):
    """H in the orthonormal real Gamma basis, built column by column from
    hpsi (the FAST_FILTER idea of filter.c:1134-1526: exact spectral form)."""
    nc = sysm.ncoef
    dim = gamma_real_basis_dim(nc)
    E = np.eye(dim)
    C = from_real_vec(E, nc)
    H = to_real_vec(hpsi(C, sysm, with_nl))
    return H.T


# ---------------------------------------------------------------------------
# Kohn-Sham potential from the density (cpGetVksStodft, LDA, periodic)
# ---------------------------------------------------------------------------

XC_PZ_LDA, XC_PW_LDA, XC_PADE_LDA = 0, 1, 2


def xc_lda(rho, kind):
    """LDA exchange-correlation of the physical density rho (e/bohr^3) on a
    grid: returns (v_xc, e_density, mu_density) with v_xc the potential,
    e_density = rho * eps_xc and mu_density the v_xc * rho term that
    excpot_* sum into muxc.  Restates excpot_pz_lda
    (energy/cp/xc_functionals.c:268-362: Perdew-Zunger, rho <= 0 clamped to
    1e-10), excpot_pw_lda (:368-480: Perdew-Wang 1992, same clamp) and
    excpot_pade_lda (:32-101: Goedecker-Teter-Hutter Pade fit, no clamp, no
    muxc contribution), without the LYP switch (lyp_fact = 1)."""
    rho = np.asarray(rho, dtype=np.float64)
    if kind == XC_PADE_LDA:
        a = (0.4581652932831429, 2.217058676663745, 0.7405551735357053, 0.01968227878617998)
        b = (1.0, 4.504130959426697, 1.110667363742916, 0.02359291751427506)
        rs = np.cbrt(3.0 / (4.0 * math.pi * rho))
        num = ((a[3] * rs + a[2]) * rs + a[1]) * rs + a[0]
        den = (((b[3] * rs + b[2]) * rs + b[1]) * rs + b[0]) * rs
        drs = -rs / (3.0 * rho)
        num_p = (3.0 * a[3] * rs + 2.0 * a[2]) * rs + a[1]
        den_p = ((4.0 * b[3] * rs + 3.0 * b[2]) * rs + 2.0 * b[1]) * rs + b[0]
        dfact = ((den * num_p - den_p * num) / (den * den)) * drs
        v = -(dfact * rho + num / den)
        return v, -(num / den) * rho, np.zeros_like(rho)
    r = np.where(rho <= 0.0, 1.0e-10, rho)
    rs = np.cbrt(3.0 / (4.0 * math.pi * r))
    sq = np.sqrt(rs)
    vx = -np.cbrt(3.0 * r / math.pi)
    if kind == XC_PZ_LDA:
        g, b1, b2 = -0.14230, 1.05290, 0.33340
        A, B, C, D = 0.03110, -0.0480, 0.00200, -0.01160
        hi = rs >= 1.0
        cf1 = 1.0 + b1 * sq + b2 * rs
        cf2 = 1.0 + (7.0 / 6.0) * b1 * sq + (4.0 / 3.0) * b2 * rs
        ec_hi = g / cf1
        mu_hi = ec_hi * (cf2 / cf1)
        lr = np.log(rs)
        ec_lo = A * lr + B + C * rs * lr + D * rs
        mu_lo = A * lr + (B - A / 3.0) + (2.0 / 3.0) * C * rs * lr + (1.0 / 3.0) * (2.0 * D - C) * rs
        ec = np.where(hi, ec_hi, ec_lo)
        mu = np.where(hi, mu_hi, mu_lo)
    elif kind == XC_PW_LDA:
        a, al1 = 0.03109070, 0.213700
        be = (7.59570, 3.58760, 1.63820, 0.492940)
        cf1 = 2.0 * a * (be[0] * sq + be[1] * rs + be[2] * rs ** 1.5 + be[3] * rs * rs)
        cf2 = 2.0 * a * (0.5 * be[0] / sq + be[1] + 1.5 * be[2] * sq + 2.0 * be[3] * rs)
        lg = np.log(1.0 + 1.0 / cf1)
        opa = 1.0 + al1 * rs
        ec = -2.0 * a * opa * lg
        mu = -2.0 * a * (1.0 + 2.0 * al1 * rs / 3.0) * lg + (-2.0 * a) * rs * opa / (3.0 * cf1 * (1.0 + cf1)) * cf2
    else:
        raise ValueError("unknown LDA kind")
    v = vx + mu
    return v, (0.75 * vx + ec) * r, v * r


def ks_potential(rho_raw, vext, dens, volume, kind):
    """v_KS(r) (FFTW order) and the energies from the stochastic density.

    rho_raw: the density as the filter path produces it (sum rho / N = N_e,
    i.e. the physical density times the cell volume), FFTW order.  dens: the
    density plane-wave half sphere (k, map, map_conj, ak2; G = 0 last);
    vext: v_ext(G) on it (plane-wave amplitudes).  Restates
    calcRhoStoRecipFullg (stodft/density.c:699-870: pack rho, r->G FFT,
    unpack the sphere, divide rho(r) by the volume) and cpGetVksStodft
    (stodft/reduced-energy/cp-energy-ee-rho-stodft.c:644-900, iperd = 3, no
    dual grid): Q(G) = (1/N) sum_r rho_raw e^{-iG.r}, v(G) = vext(G) +
    4 pi Q(G) / (|G|^2 V) for G != 0 and vext(0) at G = 0; E_H = sum over
    the half sphere of 4 pi |Q|^2 / (|G|^2 V); E_ext = 2 sum_half Re(vext
    conj(Q)) + vext(0) Q(0); v(r) = sum_G v(G) e^{iG.r} over the full sphere;
    then the LDA potential of rho_raw / V is added (excpot_*).  Returns
    (v_ks, dict(eh, eext, exc, muxc, Q))."""
    n = tuple(int(x) for x in dens.n)
    N = n[0] * n[1] * n[2]
    rho_raw = np.asarray(rho_raw, dtype=np.float64).reshape(n)
    F = np.fft.fftn(rho_raw).ravel() / N
    Q = F[dens.map]
    nc = Q.shape[0]
    ak2 = np.asarray(dens.ak2, dtype=np.float64)
    gh = np.zeros(nc)
    gh[:-1] = 4.0 * math.pi / (ak2[:-1] * volume)
    v = np.asarray(vext, dtype=np.complex128) + gh * Q
    eh = float(np.sum(gh[:-1] * np.abs(Q[:-1]) ** 2))
    eext = float(2.0 * np.sum((vext[:-1] * np.conj(Q[:-1])).real) + (vext[-1] * np.conj(Q[-1])).real)
    full = np.zeros(N, dtype=np.complex128)
    full[dens.map[:-1]] = v[:-1]
    full[dens.map_conj[:-1]] = np.conj(v[:-1])
    full[dens.map[-1]] = v[-1].real
    v_r = (np.fft.ifftn(full.reshape(n)) * N).real.ravel()
    rho_phys = rho_raw.ravel() / volume
    vxc, ed, md = xc_lda(rho_phys, kind)
    vscale = volume / N
    out = dict(eh=eh, eext=eext, exc=float(ed.sum() * vscale), muxc=float(md.sum() * vscale), Q=Q)
    return v_r + vxc, out


def local_forces(Q, dens, pos, vloc_type, atom_type):
    """Local pseudopotential Hellmann-Feynman forces from rho(G): the local
    part of calcEnergyForce (stodft/calc-nuclei-force.c:254 ->
    calcLocExtPostScf, reduced-energy/energy-wrapper-post-scf.c:140-220 ->
    control_ewd_loc, without the ion-ion Ewald term).  With v_ext(G) =
    sum_a f_type(a)(G) e^{-iG.R_a} and E_ext = sum_G Re(v_ext(G) conj Q(G)),
    F_a = -dE_ext/dR_a = 2 sum_{half sphere, G != 0} G f(G) (cos(G.R_a) Im Q
    + sin(G.R_a) Re Q).  Q: rho(G) on the density sphere (ks_potential's
    out['Q']), vloc_type: (ntype, ncoef) real form factors, pos (natm, 3)."""
    G = np.asarray(dens.gvec)[:-1]
    q = np.asarray(Q)[:-1]
    F = np.zeros((pos.shape[0], 3))
    for a in range(pos.shape[0]):
        f = np.asarray(vloc_type)[atom_type[a]][:-1]
        ph = G @ pos[a]
        t = f * (np.cos(ph) * q.imag + np.sin(ph) * q.real)
        F[a] = 2.0 * (G * t[:, None]).sum(axis=0)
    return F
