"""GPU parity of the HIP path against the CPU FP64 oracle (same seeded inputs).

Every call goes through the C ABI (libsdft_hip.so via ctypes).  Tolerances:
FP64 form <= 1e-12 relative on H.psi and <= 1e-10 relative on the filtered
orbitals / density (the plain form's bar: 1e-10 relative in density); FP32
form checked against FP64 with the bound written in each test.
"""
import sys

import numpy as np
import pytest

from oracle import sdft_oracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    from sdft_amd import hip as H
    H.load()
    return H


@pytest.fixture(scope="module")
def si8_ctx(hip, si8_system):
    return hip.Context(si8_system)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def bounds(sysm):
    from sdft_amd import stodft
    return stodft.spectral_bounds(sysm)


@pytest.mark.parametrize("n", [24, 32, 36, 48, 64, 96, 128, 192])
def test_lds_fft_matches_numpy(hip, n):
    say(f"case: selftest_fft n={n}")
    rng = np.random.default_rng(n)
    x = rng.standard_normal((37, n)) + 1j * rng.standard_normal((37, n))
    for prec, tol in ((hip.PREC_FP64, 1e-14), (hip.PREC_FP32, 3e-6)):
        yb = hip.selftest_fft(x, +1, prec)
        yf = hip.selftest_fft(x, -1, prec)
        ref_b = np.fft.ifft(x, axis=1) * n
        ref_f = np.fft.fft(x, axis=1)
        assert rel(yb, ref_b) < tol * np.log2(n), (n, prec, rel(yb, ref_b))
        assert rel(yf, ref_f) < tol * np.log2(n), (n, prec, rel(yf, ref_f))


@pytest.mark.parametrize("nstate", [1, 2, 5, 8])
def test_hpsi_fp64_si8(hip, si8_ctx, si8_system, nstate):
    from sdft_amd import systems
    say(f"case: hpsi fp64 si8 nstate={nstate}")
    chi = systems.noise_orbitals(si8_system, nstate, seed=0)
    f_gpu = si8_ctx.hpsi_raw(chi, hip.PREC_FP64)
    f_ref = O.hpsi_raw(chi, si8_system)
    assert rel(f_gpu, f_ref) < 1e-12
    assert np.all(f_gpu[:, -1].imag == 0.0)


def test_hpsi_fp32_si8(hip, si8_ctx, si8_system):
    from sdft_amd import systems
    say("case: hpsi fp32 si8")
    chi = systems.noise_orbitals(si8_system, 6, seed=3)
    f32 = si8_ctx.hpsi_raw(chi, hip.PREC_FP32)
    f_ref = O.hpsi_raw(chi, si8_system)
    # FP32 FFT / v.psi / KB GEMM: relative error bounded by ~ a few 1e-6
    assert rel(f32, f_ref) < 2e-5


def test_hpsi_state_counts_grow_and_shrink(hip, si8_system):
    """One context applied to blocks of 2, 8, 40, 3 and 33 states in turn:
    the KB projection buffers (split partials, Z = A X, the padded Y rows)
    follow the state count, both where they are reallocated (8 -> 40) and
    where the earlier allocation is reused (2 -> 8, 40 -> 3 -> 33)."""
    from sdft_amd import systems
    ctx = hip.Context(si8_system)
    for i, nstate in enumerate((2, 8, 40, 3, 33)):
        say(f"case: hpsi si8 state-count sequence nstate={nstate}")
        chi = systems.noise_orbitals(si8_system, nstate, seed=40 + i)
        ref = O.hpsi_raw(chi, si8_system)
        assert rel(ctx.hpsi_raw(chi, hip.PREC_FP64), ref) < 1e-12, nstate
        assert rel(ctx.hpsi_raw(chi, hip.PREC_FP32), ref) < 2e-5, nstate


def test_hpsi_local_only(hip, si8_system):
    from sdft_amd import systems
    say("case: hpsi local-only si8")
    ctx = hip.Context(si8_system, with_nl=False)
    chi = systems.noise_orbitals(si8_system, 4, seed=5)
    f_gpu = ctx.hpsi_raw(chi)
    f_ref = O.hpsi_raw(chi, si8_system, with_nl=False)
    assert rel(f_gpu, f_ref) < 1e-12
    ctx.close()


def test_cheby_coeffs(hip, si8_ctx):
    say("case: cheby coeffs")
    for L, fun in ((256, 2), (300, 1), (1024, 2)):
        c_gpu = si8_ctx.cheby_coeffs(L, -0.9, 6.7, 0.21, 100.0, fun)
        c_ref = O.cheby_coeffs(L, -0.9, 6.7, 0.21, 100.0, fun)
        assert np.max(np.abs(c_gpu - c_ref)) < 1e-13


def test_filter_fp64_si8(hip, si8_ctx, si8_system):
    from sdft_amd import systems
    say("case: filter fp64 si8 L=32 B=8")
    emin, emax = bounds(si8_system)
    L = 32
    coeffs = O.cheby_coeffs(L, emin, emax, 0.2, 100.0, 2)
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    chi = systems.noise_orbitals(si8_system, 8, seed=0)
    acc_g, tl_g, mom_g = si8_ctx.cheby_filter(chi, coeffs, scale, emean, moments=True)
    acc_r, tl_r, mom_r = O.filter_cheby(chi, coeffs, si8_system, scale, emean, moments=True)
    assert rel(acc_g, acc_r) < 1e-10
    assert rel(tl_g, tl_r) < 1e-10
    assert np.max(np.abs(mom_g - mom_r)) / np.max(np.abs(mom_r)) < 1e-10


def test_filter_fp32_si8(hip, si8_ctx, si8_system):
    from sdft_amd import systems
    say("case: filter fp32 si8 L=64 B=8")
    emin, emax = bounds(si8_system)
    L = 64
    coeffs = O.cheby_coeffs(L, emin, emax, 0.2, 100.0, 2)
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    chi = systems.noise_orbitals(si8_system, 8, seed=2)
    acc_g, _ = si8_ctx.cheby_filter(chi, coeffs, scale, emean, prec=hip.PREC_FP32)
    acc_r, _ = O.filter_cheby(chi, coeffs, si8_system, scale, emean)
    assert rel(acc_g, acc_r) < 1e-4


def test_moments_and_chem_pot(hip, si8_ctx, si8_system):
    from sdft_amd import systems
    say("case: moments + mu si8 L=40 B=6")
    emin, emax = bounds(si8_system)
    L = 40
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    chi = systems.noise_orbitals(si8_system, 6, seed=1)
    m_g = si8_ctx.cheby_moments(chi, L, scale, emean)
    m_r = O.cheby_moments(chi, si8_system, L, scale, emean)
    assert np.max(np.abs(m_g - m_r)) / np.max(np.abs(m_r)) < 1e-10
    beta = 1.0 / 0.01
    ne = 32.0
    mu_r, n_r = O.chem_pot_cheby(m_r, 6, L, emin, emax, beta, ne, 0.2, 0.4)
    mu_g, n_g = si8_ctx.chem_pot(L, emin, emax, beta, m_r, 6, ne, 0.2, 0.4)
    assert abs(n_g - ne) <= 1e-11 * ne * 1.0001
    assert abs(mu_g - mu_r) < 1e-9


def test_density_and_energies(hip, si8_ctx, si8_system):
    from sdft_amd import systems
    say("case: density + energies si8 B=7")
    C = systems.noise_orbitals(si8_system, 7, seed=4) * 0.3
    rho_g = si8_ctx.density(C, 2.0 / 7)
    rho_r = O.density_real(C, si8_system, 2.0 / 7)
    assert rel(rho_g, rho_r) < 1e-12
    eke, enl = si8_ctx.energies(C, 2.0, 7)
    assert abs(eke - O.energy_ke(C, si8_system.ak2, 2.0, 7)) < 1e-12 * abs(eke)
    enl_r = O.energy_nl(C, si8_system.proj, si8_system.proj_energy, 2.0, 7)
    assert abs(enl - enl_r) < 1e-12 * abs(enl_r)


def test_against_golden_fixture(hip, si8_ctx, si8_system):
    """The HIP path against the committed oracle fixture (Si8, L=32, B=4)."""
    import os
    say("case: golden fixture si8 L=32 B=4")
    g = dict(np.load(os.path.join(os.path.dirname(__file__), "golden", "si8_L32_B4.npz")))
    chi = g["chi"]
    emin, emax = g["emin_emax"]
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    assert rel(si8_ctx.hpsi_raw(chi), g["f"]) < 1e-12
    acc, tl = si8_ctx.cheby_filter(chi, g["coeffs"], scale, emean)
    assert rel(acc, g["acc"]) < 1e-10 and rel(tl, g["tl"]) < 1e-10
    _, t4 = si8_ctx.cheby_filter(chi, np.ones(5), scale, emean)
    assert rel(t4, g["T4"]) < 1e-11
    m = si8_ctx.cheby_moments(chi, 32, scale, emean)
    assert np.max(np.abs(m - g["moments"])) < 1e-10 * np.max(np.abs(g["moments"]))
    mu, nmu = si8_ctx.chem_pot(32, emin, emax, 100.0, g["moments"], 4, 32.0, 0.3, 0.4)
    assert abs(mu - g["mu"][0]) < 1e-9
    coeffs = si8_ctx.cheby_coeffs(32, emin, emax, g["mu"][0], 100.0, 2)
    assert np.max(np.abs(coeffs - g["coeffs"])) < 1e-13
    rho = si8_ctx.density(g["acc"], 2.0 / 4)
    assert abs(rho.sum() - g["rho_sum"][0]) < 1e-11 * g["rho_sum"][0]
    assert rel(rho[::16], g["rho_sub"]) < 1e-11
    eke, enl = si8_ctx.energies(g["acc"], 2.0, 4)
    assert np.allclose([eke, enl], g["energies"], rtol=1e-12)


def test_nl_forces_fp64(hip, si8_ctx, si8_system):
    """HF nonlocal forces on the GPU vs the oracle (itself pinned by finite
    differences of E_NL in tests/test_oracle.py)."""
    from sdft_amd import systems
    say("case: NL forces si8 B=6")
    C = systems.noise_orbitals(si8_system, 6, seed=8) * 0.2
    F = si8_ctx.nl_forces(C, 2.0, 6)
    F_ref = O.nl_forces(C, si8_system, 2.0, 6)
    assert np.max(np.abs(F - F_ref)) < 1e-10 * np.max(np.abs(F_ref))
    # translation invariance is not expected (local + NL); the sum over
    # projectors of one atom is what the oracle pins.  Also an odd state count:
    F5 = si8_ctx.nl_forces(C[:5], 2.0, 5)
    assert np.max(np.abs(F5 - O.nl_forces(C[:5], si8_system, 2.0, 5))) < 1e-10 * np.max(np.abs(F_ref))


@pytest.mark.parametrize("name,n,g2,nst,seed,first", [("si8", 32, 12.0, 5, 0, 0),
                                                      ("tiny", 24, 3.0, 4, 7, 11)])
def test_device_noise_matches_host_stream(hip, name, n, g2, nst, seed, first):
    """genNoiseOrbitalReal on the device reproduces the host builder's
    counter-based Rademacher noise transformed by numpy's FFT."""
    from sdft_amd import systems
    say(f"case: device noise {name} nstate={nst} seed={seed}")
    s = systems.build_system(name, 10.26, 1, n, g2max=g2)
    ctx = hip.Context(s)
    got = ctx.noise(nst, seed, first)
    ref = systems.noise_orbitals(s, nst, seed, first)
    assert rel(got, ref) < 1e-12
    # the resident block filters its own noise
    blk = hip.Block(ctx, nst, hip.PREC_FP64)
    blk.noise(seed, first)
    assert rel(blk.get(1), ref) < 1e-12
    blk.close()
    ctx.close()


@pytest.mark.parametrize("nproj", [32, 31])
def test_hpsi_real_space_kb(hip, si8_system, nproj):
    """sdft_hip_set_projectors_real against the oracle's literal real-space
    KB (nlppKBRealEnergyFilter restated), in both forms: the per-atom form on
    the real-space grid (nl_form 0, the default) and the device-built G-space
    table (nl_form 1).  H.psi FP64 1e-12, FP32 2e-5 relative; a short filter
    1e-10; the NL energy against nlppKBRealEnergy restated, 1e-12.  nproj = 31
    leaves the last atom with three channels (a partial projector group)."""
    import dataclasses

    from sdft_amd import systems
    rs = systems.real_space_projectors(si8_system, 4.0)
    rs = dataclasses.replace(rs, offsets=rs.offsets[:nproj + 1], energy=rs.energy[:nproj],
                             atom=rs.atom[:nproj])
    ctx = hip.Context(si8_system, with_nl=False)
    ctx.set_projectors_real(rs)
    C = systems.noise_orbitals(si8_system, 5, seed=21)
    ref = O.hpsi_raw(C, si8_system, nl_real=rs)
    emin, emax = bounds(si8_system)
    L = 12
    coeffs = O.cheby_coeffs(L, emin - 1.0, emax + 1.0, 0.2, 100.0, 2)
    scale, emean = 2.0 / (emax - emin + 2.0), 0.5 * (emin + emax)
    acc, _ = O.filter_cheby(C, coeffs, si8_system, scale, emean, nl_real=rs)
    enl_ref = O.energy_nl_real(C, si8_system, rs, 2.0, 5.0)
    for form in (0, 1):
        say("case: hpsi real-space KB si8 nproj=%d nl_form=%d" % (nproj, form))
        ctx.set_nl_form(form)
        e64 = rel(ctx.hpsi_raw(C), ref)
        e32 = rel(ctx.hpsi_raw(C, hip.PREC_FP32), ref)
        acc_g, _ = ctx.cheby_filter(C, coeffs, scale, emean)
        ef = rel(acc_g, acc)
        _, enl = ctx.energies(C, 2.0, 5.0)
        say("  hpsi fp64 %.3g fp32 %.3g filter %.3g enl %.12g / %.12g" % (e64, e32, ef, enl,
                                                                         enl_ref))
        assert e64 < 1e-12 and e32 < 2e-5 and ef < 1e-10, (form, e64, e32, ef)
        assert abs(enl - enl_ref) < 1e-12 * abs(enl_ref), (form, enl, enl_ref)
    ctx.close()


def test_hpsi_real_space_kb_state_counts_grow_and_shrink(hip, si8_system):
    """One context with the per-atom real-space KB applied to blocks of 2, 9,
    40, 3 and 33 states in turn: the pair grids, per-pair projections and
    energy partials it keeps follow the pair count, both where they are
    reallocated (9 -> 40) and where the earlier allocation is reused (2 -> 9
    after a first call, 40 -> 3 -> 33).  H.psi FP64 1e-12, FP32 2e-5; NL
    energy 1e-12 relative, each against the oracle's real-space KB."""
    from sdft_amd import systems
    rs = systems.real_space_projectors(si8_system, 4.0)
    ctx = hip.Context(si8_system, with_nl=False)
    ctx.set_projectors_real(rs)
    for i, nstate in enumerate((2, 9, 40, 3, 33)):
        say(f"case: hpsi real-space KB si8 state-count sequence nstate={nstate}")
        chi = systems.noise_orbitals(si8_system, nstate, seed=60 + i)
        ref = O.hpsi_raw(chi, si8_system, nl_real=rs)
        assert rel(ctx.hpsi_raw(chi, hip.PREC_FP64), ref) < 1e-12, nstate
        assert rel(ctx.hpsi_raw(chi, hip.PREC_FP32), ref) < 2e-5, nstate
        _, enl = ctx.energies(chi, 2.0, float(nstate))
        enl_ref = O.energy_nl_real(chi, si8_system, rs, 2.0, float(nstate))
        assert abs(enl - enl_ref) < 1e-12 * abs(enl_ref), (nstate, enl, enl_ref)
    ctx.close()


def test_density_generic_layout_and_accumulate(hip, si8_ctx, si8_system):
    """SDFT_RHO_GENERIC: the density in PINY's generic-FFT order (a fastest),
    which the reference converts to and from the FFTW order by a pure
    transpose (energy/cp/cp_energy_ee_rho.c:1944-1956: generic index
    (c n2 + b) n1 + a <-> FFTW index (a n2 + b) n3 + c); accumulate = 1 adds
    to the caller's array (the reference sums rho over state groups)."""
    from sdft_amd import systems
    say("case: density generic layout + accumulate")
    C = systems.noise_orbitals(si8_system, 5, seed=61)
    n = si8_system.n[0]
    rho_f = si8_ctx.density(C, 0.4, layout=hip.RHO_FFTW)
    rho_g = si8_ctx.density(C, 0.4, layout=hip.RHO_GENERIC)
    assert np.array_equal(rho_g, rho_f.reshape(n, n, n).transpose(2, 1, 0).ravel())
    rho_ref = O.density_real(C, si8_system, 0.4)
    assert rel(rho_f, rho_ref) < 1e-12
    acc = rho_f.copy()
    si8_ctx.density(C, 0.4, layout=hip.RHO_FFTW, accumulate_into=acc)
    assert rel(acc, 2.0 * rho_f) < 1e-15


def test_device_density_through_piny_files(hip, si8_ctx, si8_system, tmp_path):
    """The device density (generic order) written as PINY's final-density
    file and density checkpoint (host C) reads back to 16 digits: the
    checkpoint holds rho * volume, as stodft/checkpointIO.c:100-106 writes
    it; the filtered orbitals go through a frag-MO file the same way."""
    from sdft_amd import pinyfiles, systems
    say("case: device density -> PINY files")
    C = systems.noise_orbitals(si8_system, 4, seed=71)
    rho = si8_ctx.density(C, 0.5, layout=hip.RHO_GENERIC)
    assert rel(rho, O.density_real(C, si8_system, 0.5).reshape(
        (si8_system.n[0],) * 3).transpose(2, 1, 0).ravel()) < 1e-12
    fin = tmp_path / "density-final"
    pinyfiles.density_write(None, fin, rho)
    assert rel(pinyfiles.density_read(fin, rho.size), rho) < 1e-15
    vol = 270.0
    ck = tmp_path / "density-checkpoint"
    pinyfiles.checkpoint_write(ck, 3, 4, vol, rho, [rho], [])
    got = pinyfiles.checkpoint_read(ck, 4, rho.size)
    assert got["iscf"] == 3 and rel(got["rho_up"], rho * vol) < 1e-15
    mo = tmp_path / "frag-MO-0"
    pinyfiles.frag_mo_write(mo, [C])
    assert rel(pinyfiles.frag_mo_read(mo, [C.shape])[0], C) < 1e-15


@pytest.mark.parametrize("which", [0, 1])
def test_spectral_bound_matches_oracle(hip, si8_ctx, si8_system, which):
    """sdft_hip_spectral_bound (genEnergyMax / genEnergyMin on the device)
    against the oracle's restatement, a fixed 25 iterations (conv = 0) from
    the same start vector: energy 1e-11, final vector 1e-10 relative; and a
    converged run (conv 1e-5) stops where the oracle stops."""
    from sdft_amd import systems
    say("case: spectral bound which=%d" % which)
    v0 = systems.noise_orbitals(si8_system, 1, seed=77)[0]
    emax = 6.0
    e_g, v_g, it_g = si8_ctx.spectral_bound(v0, which, emax=emax, conv=0.0, max_iter=25)
    e_r, v_r, it_r = O.gen_energy_bound(v0, si8_system, which, emax=emax, conv=0.0, max_iter=25)
    assert it_g == it_r == 25
    assert abs(e_g - e_r) < 1e-11 * abs(e_r)
    assert rel(v_g, v_r) < 1e-10
    if which == 0:
        e_g, _, it_g = si8_ctx.spectral_bound(v0, 0, conv=1e-5, max_iter=5000)
        e_r, _, it_r = O.gen_energy_bound(v0, si8_system, 0, conv=1e-5, max_iter=5000)
        assert it_g == it_r and abs(e_g - e_r) < 1e-9 * abs(e_r)


def piny_newton_set(L, emin, emax, mus, beta):
    """The sample points and coefficients PINY hands the Newton filter
    (genSampNewtonHermit / genCoeffNewtonHermit, stodft/coeff.c:199-533, as
    the oracle restates them) and the Newton map's scale and centre
    (gen-stodft-wf.c:216-220: (Smax - Smin)/energyDiff, energyMean)."""
    ns = 4.0 / (emax - emin)
    samp, unscale = O.newton_samp_points(L, emin, ns)
    co = O.newton_hermite_coeffs(samp, unscale, mus, beta)
    return co, samp, ns, 0.5 * (emin + emax)


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-4)])
def test_newton_filter_matches_oracle(hip, si8_ctx, si8_system, prec, tol):
    """sdft_hip_newton_filter (filterNewtonPolyHerm) against the oracle with
    PINY's own sample points and divided differences for three chemical
    potentials, L = 64."""
    from sdft_amd import systems
    say("case: newton filter prec=%d (PINY sample points, L=64)" % prec)
    emin, emax = bounds(si8_system)
    co, z, ns, emean = piny_newton_set(64, emin, emax, [0.1, 0.3, 0.5], 4.0)
    chi = systems.noise_orbitals(si8_system, 5, seed=12)
    out_g, pl_g = si8_ctx.newton_filter(chi, co, z[:-1], ns, emean, prec=prec)
    out_r, pl_r = O.newton_filter(chi, co, z[:-1], si8_system, ns, emean)
    for m in range(3):
        assert rel(out_g[m], out_r[m]) < tol, (m, rel(out_g[m], out_r[m]))
    assert rel(pl_g, pl_r) < tol


def test_newton_filter_si64_piny_points_l1024(hip, si64_system):
    """filterNewtonPolyHerm at 64^3 with the sample points and coefficients
    PINY generates (genSampNewtonHermit / genCoeffNewtonHermit) at L = 1024,
    kT = 0.01 Ha, two chemical-potential columns, one orbital pair, against
    the oracle: FP64 within 1e-10 relative per column, FP32 within 1e-4.  The
    expansion's own fit error (calcFitErrorNewton) at this degree is 5.1e-5
    (bounded here by 1e-4); PINY would raise L until it falls below
    poly_fit_error (default 1e-9), which changes the points, not the
    recurrence under test."""
    from sdft_amd import systems
    say("case: newton filter si64 L=1024 PINY points, 2 mu columns (oracle ~20 s)")
    ctx = hip.Context(si64_system)
    emin, emax = bounds(si64_system)
    mus, beta = [0.25, 0.35], 100.0
    L = 1024
    co, z, ns, emean = piny_newton_set(L, emin, emax, mus, beta)
    fit = O.newton_fit_error(co, z, emin, emax, ns, mus, beta)
    say("  fit error %.3e" % fit)
    assert fit < 1e-4
    chi = systems.noise_orbitals(si64_system, 2, seed=13)
    out_r, pl_r = O.newton_filter(chi, co, z[:-1], si64_system, ns, emean)
    for prec, tol in ((hip.PREC_FP64, 1e-10), (hip.PREC_FP32, 1e-4)):
        out_g, pl_g = ctx.newton_filter(chi.copy(), co, z[:-1], ns, emean, prec=prec)
        errs = [rel(out_g[m], out_r[m]) for m in range(len(mus))]
        say("  si64 newton L=1024 prec=%d: per-column rel %s, last P %.3g" % (
            prec, ["%.3g" % e for e in errs], rel(pl_g, pl_r)))
        assert max(errs) < tol, (prec, errs)
    ctx.close()


@pytest.fixture(scope="module")
def si64_system():
    from sdft_amd import systems
    return systems.build_system("si64", 10.26, 2, 64)


def test_plane_tiles_and_tuner_si64(hip, si64_system):
    """Every plane-kernel variant at 64^3 (sdft_hip_set_config: b-line tile,
    256 / 512 threads, radix-8 / radix-4 schedules of the c- and b-lines)
    against the oracle (FP64 1e-12, FP32 2e-5), with an odd orbital count,
    and the tuner returns one of its 24 candidates."""
    from sdft_amd import systems
    say("case: plane variants 0-5 + tuner si64")
    ctx = hip.Context(si64_system)
    for nst, seed in ((4, 23), (3, 29)):
        C = systems.noise_orbitals(si64_system, nst, seed=seed)
        ref = O.hpsi_raw(C, si64_system)
        for tile in range(6):
            for kt in (256, 2048):
                ctx.set_config(tile, kt)
                e64 = rel(ctx.hpsi_raw(C), ref)
                e32 = rel(ctx.hpsi_raw(C, hip.PREC_FP32), ref)
                assert e64 < 1e-12 and e32 < 2e-5, (nst, tile, kt, e64, e32)
    with pytest.raises(RuntimeError):
        ctx.set_config(6, 256)
    (tile, kt, fu), ms = ctx.tune(8, hip.PREC_FP32, steps=2)
    assert tile in range(6) and kt in (256, 1024) and fu in (0, 1)
    assert len(ms) == 24 and np.all(ms > 0)
    C = systems.noise_orbitals(si64_system, 4, seed=23)
    assert rel(ctx.hpsi_raw(C, hip.PREC_FP32), O.hpsi_raw(C, si64_system)) < 2e-5


@pytest.fixture(scope="module")
def si64_full_ref(si64_system):
    """The oracle's filter of one orbital pair (seed 1) of the headline
    workload at its full degree: L = 1024, kT = 0.01 Ha, mu = 0.2."""
    from sdft_amd import systems
    say("case: oracle filter si64 full degree L=1024, 2 orbitals (~30 s)")
    emin, emax = bounds(si64_system)
    L = 1024
    coeffs = O.cheby_coeffs(L, emin, emax, 0.2, 100.0, 2)
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    chi = systems.noise_orbitals(si64_system, 2, seed=1)
    ref, _ = O.filter_cheby(chi, coeffs, si64_system, scale, emean)
    return dict(chi=chi, coeffs=coeffs, scale=scale, emean=emean, ref=ref)


def test_filter_si64_full_degree(hip, si64_system, si64_full_ref):
    """The headline workload's filter at its full degree (Si64, 64^3, L = 1024,
    kT = 0.01 Ha) on one orbital pair of its noise (seed 1) against the
    oracle's filterChebyPolyHerm: FP64 filtered orbitals and density within
    1e-10 relative (the plain form's bar), the FP32 form within 1e-4 relative
    (orbitals) and 1e-6 L1 per electron (density)."""
    say("case: filter si64 full degree L=1024, 2 orbitals")
    ctx = hip.Context(si64_system)
    r = si64_full_ref
    chi, coeffs, scale, emean, ref = r["chi"], r["coeffs"], r["scale"], r["emean"], r["ref"]
    rho_ref = O.density_real(ref, si64_system, 1.0)
    ne = rho_ref.sum()
    out64, _ = ctx.cheby_filter(chi, coeffs, scale, emean, prec=hip.PREC_FP64)
    assert rel(out64, ref) < 1e-10, rel(out64, ref)
    rho64 = ctx.density(out64, 1.0)
    assert rel(rho64, rho_ref) < 1e-10, rel(rho64, rho_ref)
    out32, _ = ctx.cheby_filter(chi, coeffs, scale, emean, prec=hip.PREC_FP32)
    assert rel(out32, ref) < 1e-4, rel(out32, ref)
    rho32 = ctx.density(out32, 1.0)
    say("  si64 L=1024: FP64 orbitals %.3g density %.3g; FP32 orbitals %.3g density L1/e %.3g" % (
        rel(out64, ref), rel(rho64, rho_ref), rel(out32, ref),
        np.abs(rho32 - rho_ref).sum() / ne))
    assert np.abs(rho32 - rho_ref).sum() / ne < 1e-6
    ctx.close()


def energy_and_forces_device(ctx, sysm, C, occ, nsto):
    """Total energy and NL + local Hellmann-Feynman forces of filtered
    orbitals, on the device: E_kin + E_NL (calcEnergyChemPot,
    stodft/calc-energy.c:133-191) + E_H + E_ext + E_xc of their density
    (cpGetVksStodft's energies, assembled as calcTotEnergy :329-333 without
    the ion-ion term), forces = NL (calc-nuclei-force.c:302-489) + local
    (:254).  Replaces the context's v_KS."""
    from sdft_amd import systems
    dens = systems.density_sphere(sysm)
    vext = systems.vloc_density(sysm, dens)
    eke, enl = ctx.energies(C, occ, nsto)
    rho = ctx.density(C, occ / nsto)
    ctx.ks_init(dens)
    _, en = ctx.ks_potential(rho, vext, sysm.volume, xc=0)
    e = eke + enl + en["eh"] + en["eext"] + en["exc"]
    vloc = systems.vloc_form_factor(sysm, dens)
    f = ctx.nl_forces(C, occ, nsto) + ctx.local_forces(sysm.pos, np.zeros(sysm.natm, int), vloc,
                                                       dens.gvec)
    return e, f


def energy_and_forces_oracle(sysm, C, occ, nsto):
    from sdft_amd import systems
    dens = systems.density_sphere(sysm)
    vext = systems.vloc_density(sysm, dens)
    eke = O.energy_ke(C, sysm.ak2, occ, nsto)
    enl = O.energy_nl(C, sysm.proj, sysm.proj_energy, occ, nsto)
    rho = O.density_real(C, sysm, occ / nsto)
    _, out = O.ks_potential(rho, vext, dens, sysm.volume, 0)
    e = eke + enl + out["eh"] + out["eext"] + out["exc"]
    vloc = systems.vloc_form_factor(sysm, dens)
    f = O.nl_forces(C, sysm, occ, nsto) + O.local_forces(out["Q"], dens, sysm.pos, vloc,
                                                         np.zeros(sysm.natm, int))
    return e, f


def test_fp32_acceptance_energy_and_forces_si64(hip, si64_system, si64_full_ref):
    """The precision form's acceptance bar on the headline workload at its
    full degree (Si64, L = 1024, one orbital pair of seed 1): the total
    energy (kinetic + nonlocal + Hartree + external + LDA exchange-correlation
    of the filtered orbitals and their density) within 1e-4 Ha/atom of the
    FP64 oracle's, and the nonlocal + local Hellmann-Feynman forces within
    1e-3 Ha/bohr; the FP64 form meets the same quantities to 1e-9."""
    say("case: FP32 acceptance energy/forces si64 L=1024")
    r = si64_full_ref
    s = si64_system
    e_r, f_r = energy_and_forces_oracle(s, r["ref"], 2.0, 2.0)
    out = {}
    for prec in (hip.PREC_FP64, hip.PREC_FP32):
        ctx = hip.Context(s)
        C, _ = ctx.cheby_filter(r["chi"], r["coeffs"], r["scale"], r["emean"], prec=prec)
        out[prec] = energy_and_forces_device(ctx, s, C, 2.0, 2.0)
        ctx.close()
    (e64, f64), (e32, f32) = out[hip.PREC_FP64], out[hip.PREC_FP32]
    de32 = abs(e32 - e_r) / s.natm
    df32 = float(np.max(np.abs(f32 - f_r)))
    say("  E oracle %.10f, FP64 dE %.3g, FP32 dE/atom %.3g Ha; |F| max %.3g, FP32 dF %.3g Ha/bohr"
        % (e_r, abs(e64 - e_r), de32, np.abs(f_r).max(), df32))
    assert abs(e64 - e_r) < 1e-9 * abs(e_r)
    assert np.max(np.abs(f64 - f_r)) < 1e-9 * max(1.0, np.abs(f_r).max())
    assert de32 < 1e-4, de32
    assert df32 < 1e-3, df32


def test_moments_and_mu_si64_full_degree(hip, si64_system):
    """The moments pass and the mu search at the headline degree (Si64,
    L = 1024, two orbitals of seed 1): the L + 1 moments against the oracle's
    calcChebyMoments within 1e-10 of their largest magnitude, and mu from the
    device's bisection against the oracle's from the same moments within
    1e-9 (electron count 2 x 2 orbitals scaled to the two-orbital sample)."""
    from sdft_amd import systems
    say("case: moments + mu si64 L=1024, 2 orbitals (oracle ~15 s)")
    ctx = hip.Context(si64_system)
    emin, emax = bounds(si64_system)
    L = 1024
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    chi = systems.noise_orbitals(si64_system, 2, seed=1)
    m_g = ctx.cheby_moments(chi, L, scale, emean)
    m_r = O.cheby_moments(chi, si64_system, L, scale, emean)
    err = np.max(np.abs(m_g - m_r)) / np.max(np.abs(m_r))
    say("  si64 L=1024 moments max rel %.3g" % err)
    assert err < 1e-10
    beta, ne = 100.0, float(si64_system.nelec)
    mu_r, n_r = O.chem_pot_cheby(m_r, 2, L, emin, emax, beta, ne, 0.3, 0.4)
    mu_g, n_g = ctx.chem_pot(L, emin, emax, beta, m_r, 2, ne, 0.3, 0.4)
    assert abs(n_g - ne) <= 1e-11 * ne * 1.0001
    assert abs(mu_g - mu_r) < 1e-9, (mu_g, mu_r)
    ctx.close()


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-4)])
def test_filter_fused_and_separate_expand(hip, si8_system, si8_ctx, prec, tol):
    """The filter with the contract kernel writing the next step's
    a-expansion (fused) and with an expand kernel per step give the oracle's
    filterChebyPolyHerm; in FP64 the two forms agree to rounding."""
    from sdft_amd import stodft, systems
    say("case: filter fused vs separate expand prec=%d" % prec)
    emin, emax = stodft.spectral_bounds(si8_system)
    L = 17
    co = O.cheby_coeffs(L, emin, emax, 0.3, 100.0, 2)
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    chi = systems.noise_orbitals(si8_system, 6, seed=31)
    ref, _ = O.filter_cheby(chi, co, si8_system, scale, emean)
    out = {}
    for fu in (0, 1):
        si8_ctx.set_config(0, 512, fu)
        out[fu], _ = si8_ctx.cheby_filter(chi, co, scale, emean, prec=prec)
        assert rel(out[fu], ref) < tol, (fu, rel(out[fu], ref))
    si8_ctx.set_config(0, 512, 1)
    if prec == 0:
        assert rel(out[0], out[1]) < 1e-13


def test_config1_si8_full_workload(hip, si8_ctx, si8_system):
    """The CPU-runnable configuration in full: Si8 (32^3, ncoef 370, 8 atoms,
    32 G-space projectors), 64 Rademacher orbitals from seed 0, Chebyshev
    degree 256, kT = 0.01 Ha -- the moments pass, mu for N_e = 32 (relative
    tolerance 1e-11, calc-chempot-chebyshev.c:78), the sqrt(f)
    coefficients, the filter and the density, each on the device against the
    oracle's chain: moments, filtered orbitals, T_{L-1} and density 1e-10
    relative, mu 1e-9; the FP32 form's filtered orbitals 1e-4."""
    from sdft_amd import systems
    say("case: config 1 in full (Si8, 64 orbitals, L=256; oracle ~15 s)")
    s = si8_system
    B, L, beta = 64, 256, 100.0
    emin, emax = bounds(s)
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    chi = systems.noise_orbitals(s, B, seed=0)
    ne = s.nelec
    mom_r = O.cheby_moments(chi, s, L, scale, emean)
    mu_r, n_r = O.chem_pot_cheby(mom_r, B, L, emin, emax, beta, ne, 0.3, 0.4)
    co_r = O.cheby_coeffs(L, emin, emax, mu_r, beta, 2)
    acc_r, tl_r = O.filter_cheby(chi, co_r, s, scale, emean)
    rho_r = O.density_real(acc_r, s, 2.0 / B)
    mom_g = si8_ctx.cheby_moments(chi, L, scale, emean)
    mu_g, n_g = si8_ctx.chem_pot(L, emin, emax, beta, mom_g, B, ne, 0.3, 0.4)
    co_g = si8_ctx.cheby_coeffs(L, emin, emax, mu_g, beta, 2)
    acc_g, tl_g = si8_ctx.cheby_filter(chi, co_g, scale, emean)
    rho_g = si8_ctx.density(acc_g, 2.0 / B)
    acc32, _ = si8_ctx.cheby_filter(chi, co_g, scale, emean, prec=hip.PREC_FP32)
    say("  config 1: moments %.3g, mu %.12f / %.12f, filter %.3g, T_L-1 %.3g, density %.3g, "
        "N %.10f, FP32 filter %.3g" % (rel(mom_g, mom_r), mu_g, mu_r, rel(acc_g, acc_r),
                                       rel(tl_g, tl_r), rel(rho_g, rho_r),
                                       rho_g.sum() / s.ngrid, rel(acc32, acc_r)))
    assert rel(mom_g, mom_r) < 1e-10
    assert abs(mu_g - mu_r) < 1e-9 and abs(n_g - ne) <= 1e-11 * ne
    assert rel(co_g, co_r) < 1e-12
    assert rel(acc_g, acc_r) < 1e-10 and rel(tl_g, tl_r) < 1e-10
    assert rel(rho_g, rho_r) < 1e-10
    assert rel(acc32, acc_r) < 1e-4


def test_hpsi_real_space_kb_dense_overlap(hip, si8_system):
    """Per-atom real-space KB with 300 synthetic projectors crowded into the
    first 2048 grid points (8 point blocks of the apply kernel, each touched
    by more projectors than it stages in LDS, so the entries past the staged
    set read Z directly), distinct point lists (no channel groups), random
    values and energies of both signs: H.psi against the oracle, FP64 1e-12,
    FP32 2e-5 relative, and the NL energy 1e-12."""
    from sdft_amd import systems
    say("case: real-space KB, 300 crowded synthetic projectors")
    s = si8_system
    rng = np.random.default_rng(5)
    nproj, npts = 300, 300
    offs, idx, vals = [0], [], []
    for _ in range(nproj):
        ix = np.sort(rng.choice(2048, npts, replace=False)).astype(np.int32)
        idx.append(ix)
        vals.append(0.05 * rng.standard_normal(npts))
        offs.append(offs[-1] + npts)
    energy = rng.uniform(0.5, 2.0, nproj) * np.where(rng.random(nproj) < 0.2, -1.0, 1.0)
    rs = systems.RealSpaceProjectors(np.array(offs, dtype=np.int64), np.concatenate(idx),
                                     np.concatenate(vals), energy, s.volume / s.ngrid,
                                     np.zeros(nproj, dtype=np.int64))
    ctx = hip.Context(s, with_nl=False)
    ctx.set_projectors_real(rs)
    C = systems.noise_orbitals(s, 7, seed=29)
    ref = O.hpsi_raw(C, s, nl_real=rs)
    e64 = rel(ctx.hpsi_raw(C), ref)
    e32 = rel(ctx.hpsi_raw(C, hip.PREC_FP32), ref)
    _, enl = ctx.energies(C, 2.0, 7.0)
    enl_ref = O.energy_nl_real(C, s, rs, 2.0, 7.0)
    say("  crowded: hpsi fp64 %.3g fp32 %.3g, enl %.12g / %.12g" % (e64, e32, enl, enl_ref))
    assert e64 < 1e-12 and e32 < 2e-5, (e64, e32)
    assert abs(enl - enl_ref) < 1e-12 * abs(enl_ref)
    ctx.close()


def test_nl_forces_real_space_per_atom(hip, si8_system):
    """Hellmann-Feynman NL forces in the per-atom real-space form
    (sdft_hip_set_projector_gradients, then sdft_hip_nl_forces) against the
    oracle's nlppKBRealEnergyForce restatement (itself pinned by finite
    differences of the real-space energy): 1e-10 relative to the largest
    component, for an odd orbital count (a zero-padded pair), then on the
    same context for 24 and 3 orbitals."""
    from sdft_amd import systems
    say("case: per-atom real-space NL forces si8")
    s = si8_system
    rs = systems.real_space_projectors(s, 4.0)
    grad = systems.real_space_projector_gradients(s, rs)
    ctx = hip.Context(s, with_nl=False)
    ctx.set_projectors_real(rs)
    ctx.set_projector_gradients(grad)
    C = systems.noise_orbitals(s, 5, seed=37)
    F = ctx.nl_forces(C, 2.0, 5.0)
    F_ref = O.nl_forces_real(C, s, rs, grad, 2.0, 5.0)
    say("  per-atom NL forces: max |F| %.4g, rel %.3g" % (np.abs(F_ref).max(), rel(F, F_ref)))
    assert np.abs(F_ref).max() > 0.1
    assert rel(F, F_ref) < 1e-10
    # the same context with more, then fewer states: its force partials grow
    # (5 -> 24) and are reused (24 -> 3)
    for nstate, seed in ((24, 38), (3, 39)):
        say("case: per-atom real-space NL forces si8 nstate=%d (same context)" % nstate)
        C = systems.noise_orbitals(s, nstate, seed=seed)
        F = ctx.nl_forces(C, 2.0, float(nstate))
        F_ref = O.nl_forces_real(C, s, rs, grad, 2.0, float(nstate))
        assert rel(F, F_ref) < 1e-10, nstate
    ctx.close()
