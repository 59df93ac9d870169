"""GPU parity at the large grids of the Si512 (128^3) and Si1728 (192^3)
workloads.

The plane-wave spheres here have the same largest index K as those
workloads (K = 22 -> 45 kb rows at 128^3, K = 33 -> 67 rows at 192^3), so the
kernels run with the same shapes and LDS layouts -- at 192^3 in FP64 the
plane, density and noise kernels take their split forms -- while the cell
is the 64-atom one (a = 20.52 bohr, higher cutoff) so that the oracle and the
system set-up stay fast.  Checks: H.psi against the oracle (FP64 1e-12,
FP32 2e-5 relative), the density against the oracle (1e-12), device noise
against the host stream (1e-12), a short filter (1e-10) and Hermiticity of
the device H (size-independent property).
"""
import math
import sys

import numpy as np
import pytest

from oracle import sdft_oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope="module")
def hip():
    from sdft_amd import hip as H
    H.load()
    return H


def _system(n, K):
    from sdft_amd import systems
    g2 = ((K + 0.5) * 2.0 * math.pi / 20.52) ** 2
    s = systems.build_system("si64-%d^3-K%d" % (n, K), 10.26, 2, n, g2max=g2, separable_sf=True)
    assert int(np.abs(s.k).max()) == K
    return s


@pytest.fixture(scope="module", params=[(128, 22), (192, 33)], ids=["128", "192"])
def large(request, hip):
    n, K = request.param
    s = _system(n, K)
    return s, hip.Context(s)


def test_hpsi_large_grids(hip, large):
    from sdft_amd import systems
    s, ctx = large
    say("case: hpsi %s ncoef=%d" % (s.name, s.ncoef))
    C = systems.noise_orbitals(s, 3, seed=31)
    ref = O.hpsi_raw(C, s)
    assert rel(ctx.hpsi_raw(C), ref) < 1e-12
    assert rel(ctx.hpsi_raw(C, hip.PREC_FP32), ref) < 2e-5


def test_hermitian_large_grids(hip, large):
    from sdft_amd import systems
    s, ctx = large
    say("case: hermiticity %s" % s.name)
    a = systems.noise_orbitals(s, 1, seed=41)
    b = systems.noise_orbitals(s, 1, seed=42)
    for prec, tol in ((hip.PREC_FP64, 1e-12), (hip.PREC_FP32, 1e-5)):
        fa = ctx.hpsi_raw(a, prec)
        fb = ctx.hpsi_raw(b, prec)
        # <a|H b> = <H a|b> with H c = -f/4 (G != 0), -f/2 (G = 0)
        ha, hb = -0.25 * fa, -0.25 * fb
        ha[:, -1] *= 2.0
        hb[:, -1] *= 2.0
        lhs = O.gamma_dot(a, hb).sum()
        rhs = O.gamma_dot(ha, b).sum()
        assert abs(lhs - rhs) < tol * abs(O.gamma_dot(ha, ha).sum()) ** 0.5 * \
            abs(O.gamma_dot(b, b).sum()) ** 0.5


def test_density_and_noise_large_grids(hip, large):
    from sdft_amd import systems
    s, ctx = large
    say("case: device noise + density %s" % s.name)
    got = ctx.noise(3, seed=5, first=2)
    ref = systems.noise_orbitals(s, 3, seed=5, first=2)
    assert rel(got, ref) < 1e-12
    rho_g = ctx.density(ref, 2.0 / 3)
    rho_r = O.density_real(ref, s, 2.0 / 3)
    assert rel(rho_g, rho_r) < 1e-12


def test_filter_large_grids(hip, large):
    from sdft_amd import stodft, systems
    s, ctx = large
    say("case: short filter %s" % s.name)
    emin, emax = stodft.spectral_bounds(s)
    L = 4
    coeffs = O.cheby_coeffs(L, emin, emax, 0.3, 100.0, 2)
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    chi = systems.noise_orbitals(s, 2, seed=51)
    acc_g, tl_g = ctx.cheby_filter(chi, coeffs, scale, emean)
    acc_r, tl_r = O.filter_cheby(chi, coeffs, s, scale, emean)
    assert rel(acc_g, acc_r) < 1e-10
    assert rel(tl_g, tl_r) < 1e-10


def test_filter_many_steps_large_grids(hip, large):
    """A longer filter at the Si512 / Si1728 sphere shapes (L = 64 at 128^3,
    32 at 192^3; one orbital pair) against the oracle: FP64 filtered orbitals
    and density within 1e-10 relative, FP32 orbitals within 1e-4, so the
    recurrence's error growth over many steps stays bounded on the split
    (192^3 FP64) and unsplit forms alike."""
    from sdft_amd import stodft, systems
    s, ctx = large
    L = 64 if s.n[0] == 128 else 32
    say("case: filter L=%d %s (oracle ~%d s)" % (L, s.name, 20 if L == 64 else 30))
    emin, emax = stodft.spectral_bounds(s)
    coeffs = O.cheby_coeffs(L, emin, emax, 0.3, 100.0, 2)
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    chi = systems.noise_orbitals(s, 2, seed=57)
    acc_r, _ = O.filter_cheby(chi, coeffs, s, scale, emean)
    acc_g, _ = ctx.cheby_filter(chi, coeffs, scale, emean)
    assert rel(acc_g, acc_r) < 1e-10, rel(acc_g, acc_r)
    rho_g = ctx.density(acc_g, 1.0)
    rho_r = O.density_real(acc_r, s, 1.0)
    assert rel(rho_g, rho_r) < 1e-10, rel(rho_g, rho_r)
    acc32, _ = ctx.cheby_filter(chi, coeffs, scale, emean, prec=hip.PREC_FP32)
    say("  %s L=%d: FP64 orbitals %.3g density %.3g, FP32 orbitals %.3g" % (
        s.name, L, rel(acc_g, acc_r), rel(rho_g, rho_r), rel(acc32, acc_r)))
    assert rel(acc32, acc_r) < 1e-4, rel(acc32, acc_r)


def test_real_space_kb_large_grids(hip, large):
    """The per-atom real-space KB at 128^3 and 192^3 (at 192^3 in FP64 the
    plane stage runs its split form on both sides of the real-space grid):
    the projectors of 8 atoms with a 2-bohr cutoff (about 27k grid points per
    atom at 192^3), H.psi against the oracle's literal gather / dot /
    scatter (FP64 1e-12, FP32 2e-5 relative) and the NL energy against
    nlppKBRealEnergy restated (1e-12)."""
    import dataclasses

    from sdft_amd import systems
    s, _ = large
    say("case: per-atom real-space KB %s (8 atoms)" % s.name)
    rs = systems.real_space_projectors(s, 2.0)
    rs = dataclasses.replace(rs, offsets=rs.offsets[:33], energy=rs.energy[:32],
                             atom=rs.atom[:32])
    ctx = hip.Context(s, with_nl=False)
    ctx.set_projectors_real(rs)
    C = systems.noise_orbitals(s, 3, seed=61)
    ref = O.hpsi_raw(C, s, nl_real=rs)
    e64 = rel(ctx.hpsi_raw(C), ref)
    e32 = rel(ctx.hpsi_raw(C, hip.PREC_FP32), ref)
    _, enl = ctx.energies(C, 2.0, 3.0)
    enl_ref = O.energy_nl_real(C, s, rs, 2.0, 3.0)
    say("  %s: hpsi fp64 %.3g fp32 %.3g, enl %.12g / %.12g" % (s.name, e64, e32, enl, enl_ref))
    assert e64 < 1e-12 and e32 < 2e-5, (e64, e32)
    assert abs(enl - enl_ref) < 1e-12 * abs(enl_ref)
    ctx.close()
