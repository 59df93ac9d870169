"""The C host shim (libsdft_host.so, include/sdft_stodft.h) driven with the
reference's conventions -- 1-based arrays, SDFT_STODFT fields named after the
PINY struct members -- against the oracle."""
import ctypes as C
import sys

import numpy as np
import pytest

from oracle import sdft_oracle as O

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)


class SdftStodft(C.Structure):
    _fields_ = [("ctx", C.c_void_p), ("prec", C.c_int),
                ("numStateUpProc", C.c_int), ("numCoeff", C.c_int), ("polynormLength", C.c_int),
                ("numStateStoUp", C.c_int),
                ("energyMin", C.c_double), ("energyMax", C.c_double), ("energyMean", C.c_double),
                ("energyDiff", C.c_double), ("scale", C.c_double), ("beta", C.c_double),
                ("occNumber", C.c_double), ("numElecTrue", C.c_double),
                ("chemPotInit", C.c_double), ("gapInit", C.c_double),
                ("cre_up", _dp), ("cim_up", _dp), ("fcre_up", _dp), ("fcim_up", _dp),
                ("wfUpRe1", _dp), ("wfUpIm1", _dp),
                ("stoWfUpRe", C.POINTER(_dp)), ("stoWfUpIm", C.POINTER(_dp)),
                ("expanCoeff", _dp), ("chebyMomentsUp", _dp), ("chemPot", _dp), ("rhoUp", _dp)]


def one_based(a):
    """Pointer p with p[1] = a[0] (the reference's cmalloc(n)-1 idiom)."""
    addr = a.ctypes.data - 8
    return C.cast(C.c_void_p(addr), _dp)


@pytest.fixture(scope="module")
def shim():
    from sdft_amd import build
    lib = C.CDLL(build.host_lib_path())
    for f in ("filterChebyPolyHerm", "calcCoefForceWrapSCF", "calcChebyMoments", "calcChemPotCheby",
              "calcRhoStoHybridCheby"):
        getattr(lib, f).argtypes = [C.POINTER(SdftStodft)]
        getattr(lib, f).restype = None
    lib.normHCheby.argtypes = [C.POINTER(SdftStodft), C.c_int, C.c_int]
    lib.normHCheby.restype = None
    return lib


def test_shim_filter_normh_mu_density(shim, si8_system):
    from sdft_amd import hip, stodft, systems
    print("case: host shim si8 B=4 L=24", flush=True)
    s = si8_system
    ctx = hip.Context(s)
    B, L, nc = 4, 24, s.ncoef
    chi = systems.noise_orbitals(s, B, 3)
    emin, emax = stodft.spectral_bounds(s)
    bufs = {k: np.zeros(B * nc) for k in ("cre", "cim", "fre", "fim", "wre", "wim", "ore", "oim")}
    bufs["cre"][:] = chi.real.ravel()
    bufs["cim"][:] = chi.imag.ravel()
    coeffs = O.cheby_coeffs(L, emin, emax, 0.3, 100.0, 2)
    mom = np.zeros(L + 1)
    chem = np.zeros(1)
    rho = np.zeros(s.ngrid)
    ore_p = (_dp * 1)(one_based(bufs["ore"]))
    oim_p = (_dp * 1)(one_based(bufs["oim"]))
    st = SdftStodft(ctx=ctx.h, prec=0, numStateUpProc=B, numCoeff=nc, polynormLength=L,
                    numStateStoUp=B, energyMin=emin, energyMax=emax,
                    energyMean=0.5 * (emin + emax), energyDiff=emax - emin,
                    scale=2.0 / (emax - emin), beta=100.0, occNumber=2.0, numElecTrue=16.0,
                    chemPotInit=0.3, gapInit=0.4,
                    cre_up=one_based(bufs["cre"]), cim_up=one_based(bufs["cim"]),
                    fcre_up=one_based(bufs["fre"]), fcim_up=one_based(bufs["fim"]),
                    wfUpRe1=one_based(bufs["wre"]), wfUpIm1=one_based(bufs["wim"]),
                    stoWfUpRe=ore_p, stoWfUpIm=oim_p,
                    expanCoeff=coeffs.ctypes.data_as(_dp), chebyMomentsUp=mom.ctypes.data_as(_dp),
                    chemPot=chem.ctypes.data_as(_dp), rhoUp=rho.ctypes.data_as(_dp))
    scale, emean = st.scale, st.energyMean
    acc_r, tl_r = O.filter_cheby(chi, coeffs, s, scale, emean)
    shim.filterChebyPolyHerm(C.byref(st))
    acc = (bufs["ore"] + 1j * bufs["oim"]).reshape(B, nc)
    tl = (bufs["cre"] + 1j * bufs["cim"]).reshape(B, nc)
    assert np.max(np.abs(acc - acc_r)) < 1e-10 * np.max(np.abs(acc_r))
    assert np.max(np.abs(tl - tl_r)) < 1e-10 * np.max(np.abs(tl_r))
    # normHCheby loop on the host arrays (reference call pattern, filter.c:560-633)
    bufs["cre"][:] = chi.real.ravel()
    bufs["cim"][:] = chi.imag.ravel()
    bufs["wre"][:] = chi.real.ravel()
    bufs["wim"][:] = chi.imag.ravel()
    acc2 = coeffs[0] * chi
    for k in range(1, L):
        shim.normHCheby(C.byref(st), k, 1)
        acc2 = acc2 + coeffs[k] * (bufs["cre"] + 1j * bufs["cim"]).reshape(B, nc)
    assert np.max(np.abs(acc2 - acc_r)) < 1e-10 * np.max(np.abs(acc_r))
    # chemical potential from the moments of chi
    bufs["cre"][:] = chi.real.ravel()
    bufs["cim"][:] = chi.imag.ravel()
    shim.calcChemPotCheby(C.byref(st))
    m_r = O.cheby_moments(chi, s, L, scale, emean)
    mu_r, _ = O.chem_pot_cheby(m_r, B, L, emin, emax, 100.0, 16.0, 0.3, 0.4)
    assert abs(chem[0] - mu_r) < 1e-9
    # density of the filtered orbitals
    bufs["ore"][:] = acc_r.real.ravel()
    bufs["oim"][:] = acc_r.imag.ravel()
    shim.calcRhoStoHybridCheby(C.byref(st))
    rho_r = O.density_real(acc_r, s, 2.0 / B)
    assert np.max(np.abs(rho - rho_r)) < 1e-12 * np.max(np.abs(rho_r))
    ctx.close()
