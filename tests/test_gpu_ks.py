"""GPU parity of the Kohn-Sham potential from the density
(sdft_hip_ks_init / sdft_hip_ks_potential) against the oracle's restatement
of calcRhoStoRecipFullg + cpGetVksStodft: v_KS at 1e-10 relative, the four
energies at 1e-10 relative, both density layouts, the three LDA forms, the
density kernel's split form at 128^3; and the filter path using the new v_KS.
"""
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


def _density(sysm, nstate=6, seed=41):
    """A stochastic density of noise orbitals (the filter path's own
    density kernel output convention: sum rho / N = occ <chi|chi> / nsto)."""
    from sdft_amd import systems
    chi = systems.noise_orbitals(sysm, nstate, seed=seed)
    return O.density_real(chi, sysm, 2.0 / nstate)


@pytest.mark.parametrize("name,rep,n", [("si8", 1, 32), ("si64", 2, 64)])
@pytest.mark.parametrize("xc", [0, 1, 2])
def test_ks_potential_matches_oracle(hip, name, rep, n, xc):
    from sdft_amd import systems
    say("case: ks_potential %s xc=%d" % (name, xc))
    s = systems.build_system(name, 10.26, rep, n)
    dens = systems.density_sphere(s)
    vext = systems.vloc_density(s, dens)
    rho = _density(s)
    ref, en = O.ks_potential(rho, vext, dens, s.volume, xc)
    ctx = hip.Context(s, with_nl=False)
    ctx.ks_init(dens)
    v, eg = ctx.ks_potential(rho, vext, s.volume, xc=xc)
    assert rel(v, ref) < 1e-10, rel(v, ref)
    for k in ("eh", "eext", "exc"):
        assert abs(eg[k] - en[k]) <= 1e-10 * abs(en[k]), (k, eg[k], en[k])
    assert abs(eg["muxc"] - en["muxc"]) <= 1e-10 * max(abs(en["muxc"]), 1e-300)
    # generic (a fastest) layout in and out
    gen = rho.reshape(s.n).transpose(2, 1, 0).ravel()
    vg, _ = ctx.ks_potential(gen, vext, s.volume, xc=xc, layout=hip.RHO_GENERIC)
    assert rel(vg, ref.reshape(s.n).transpose(2, 1, 0).ravel()) < 1e-10
    ctx.close()


def test_ks_potential_feeds_the_filter(hip):
    """After sdft_hip_ks_potential the context's H uses the new v_KS: H.psi
    against the oracle with sysm.vks replaced."""
    import dataclasses
    from sdft_amd import systems
    say("case: ks_potential -> hpsi si8")
    s = systems.build_system("si8", 10.26, 1, 32)
    dens = systems.density_sphere(s)
    vext = systems.vloc_density(s, dens)
    rho = _density(s)
    ref, _ = O.ks_potential(rho, vext, dens, s.volume, 0)
    ctx = hip.Context(s)
    ctx.ks_init(dens)
    ctx.ks_potential(rho, vext, s.volume, xc=0)
    chi = systems.noise_orbitals(s, 4, seed=5)
    s2 = dataclasses.replace(s, vks=ref)
    assert rel(ctx.hpsi_raw(chi), O.hpsi_raw(chi, s2)) < 1e-11
    ctx.close()


def test_ks_potential_128_split_density_kernel(hip):
    """128^3 (density sphere of 91 kb rows: the split form of the density
    kernel and the chunked field transform)."""
    import math
    from sdft_amd import systems
    say("case: ks_potential 128^3, density sphere K = 45")
    g2 = (22.5 * 2.0 * math.pi / 20.52) ** 2
    s = systems.build_system("si64-128^3-K22", 10.26, 2, 128, g2max=g2, with_nl=False,
                             separable_sf=True)
    dens = systems.density_sphere(s, g2)
    assert int(np.abs(dens.k).max()) == 45
    vext = systems.vloc_density(s, dens)
    rho = np.abs(np.fft.ifftn(np.fft.fftn(s.vks.reshape(s.n)))).ravel() * 1e-2 + 1e-3
    ref, en = O.ks_potential(rho, vext, dens, s.volume, 1)
    ctx = hip.Context(s, with_nl=False)
    ctx.ks_init(dens)
    v, eg = ctx.ks_potential(rho, vext, s.volume, xc=1)
    assert rel(v, ref) < 1e-10, rel(v, ref)
    assert abs(eg["eh"] - en["eh"]) <= 1e-10 * abs(en["eh"])
    ctx.close()


def test_ks_rejects_bad_calls(hip):
    from sdft_amd import systems
    say("case: ks argument checks")
    s = systems.build_system("si8", 10.26, 1, 32)
    ctx = hip.Context(s, with_nl=False)
    rho = np.ones(s.ngrid)
    p = rho.ctypes.data_as(hip._dp)
    with pytest.raises(hip.SdftError):  # no ks_init: rejected by the library
        hip._check(ctx.lib.sdft_hip_ks_potential(ctx.h, p, 0, p, p, 1.0, 0, None, None))
    dens = systems.density_sphere(s)
    ctx.ks_init(dens)
    with pytest.raises(hip.SdftError):
        ctx.ks_potential(rho, np.zeros(dens.ncoef, complex), s.volume, xc=7)
    with pytest.raises(hip.SdftError):
        ctx.ks_potential(rho, np.zeros(dens.ncoef, complex), -1.0)
    ctx.close()


@pytest.mark.parametrize("name,rep,n", [("si8", 1, 32), ("si64", 2, 64)])
def test_local_forces_match_oracle(hip, name, rep, n):
    """Local pseudopotential forces from the device density vs the oracle
    (itself pinned by finite differences of E_ext), 1e-10 relative to the
    largest component."""
    from sdft_amd import systems
    say("case: local forces %s" % name)
    s = systems.build_system(name, 10.26, rep, n)
    dens = systems.density_sphere(s)
    vext = systems.vloc_density(s, dens)
    rho = _density(s)
    _, en = O.ks_potential(rho, vext, dens, s.volume, 0)
    ff = systems.vloc_form_factor(s, dens)
    types = np.zeros(s.natm, dtype=np.int32)
    ref = O.local_forces(en["Q"], dens, s.pos, ff, types)
    ctx = hip.Context(s, with_nl=False)
    ctx.ks_init(dens)
    with pytest.raises(hip.SdftError):  # no density on the device yet
        ctx.local_forces(s.pos, types, ff, dens.gvec)
    ctx.ks_potential(rho, vext, s.volume, xc=0)
    F = ctx.local_forces(s.pos, types, ff, dens.gvec)
    assert rel(F, ref) < 1e-10, rel(F, ref)
    # two atom types (the same form factor scaled) exercise the type table
    types2 = (np.arange(s.natm) % 2).astype(np.int32)
    ff2 = np.concatenate([ff, 0.5 * ff])
    ref2 = O.local_forces(en["Q"], dens, s.pos, ff2, types2)
    assert rel(ctx.local_forces(s.pos, types2, ff2, dens.gvec), ref2) < 1e-10
    ctx.close()
