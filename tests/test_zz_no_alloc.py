"""sdft_hip_chem_pot_noalloc (the chemical-potential search on the library's
fixed per-device scratch array, no device allocation) against
sdft_hip_chem_pot through the C ABI: the same kernels on the same inputs, so
mu and N(mu) agree bit for bit.  Two calls in a row, the second at a smaller
degree (the shared array keeps the first call's contents), then a degree
above the array's limit, which must be refused before any launch."""
import ctypes as C
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)

from oracle import sdft_oracle as O  # noqa: E402

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def ctx():
    from sdft_amd import hip, systems
    sysm = systems.build_system("si8", 10.26, 1, 32)
    c = hip.Context(sysm)
    yield c
    c.close()


def _search(ctx, fn, L, moments, nsto, ne, beta, emin, emax):
    mu, nn = C.c_double(), C.c_double()
    m = np.ascontiguousarray(moments, dtype=np.float64)
    rc = fn(ctx.h, L, emin, emax, beta, m.ctypes.data_as(C.POINTER(C.c_double)), nsto, ne, 0.3,
            0.4, 2.0, 1e-11 * ne, C.byref(mu), C.byref(nn))
    return rc, mu.value, nn.value


def _moments(L, nsto, seed):
    """Moments of a synthetic spectrum (a complete-basis trace of T_k over
    eigenvalues spread over [emin, emax]), scaled as a sum over nsto
    orbitals, so that N(mu) is monotone and the search has a root."""
    rng = np.random.default_rng(seed)
    emin, emax = -0.3, 2.5
    ev = np.sort(rng.uniform(emin, emax, 40))
    y = (ev - 0.5 * (emin + emax)) * 2.0 / (emax - emin)
    k = np.arange(L)
    mom = nsto * np.cos(np.outer(k, np.arccos(np.clip(y, -1, 1)))).sum(axis=1)
    return mom, emin, emax


def test_chem_pot_noalloc_matches_allocating_entry(ctx):
    lib = ctx.lib
    beta, nsto = 40.0, 4.0
    for L, ne, seed in ((1024, 32.0, 0), (256, 20.0, 1)):
        say("case: chem_pot_noalloc L=%d ne=%g seed=%d" % (L, ne, seed))
        mom, emin, emax = _moments(L, nsto, seed)
        rc_a, mu_a, n_a = _search(ctx, lib.sdft_hip_chem_pot, L, mom, nsto, ne, beta, emin, emax)
        rc_n, mu_n, n_n = _search(ctx, lib.sdft_hip_chem_pot_noalloc, L, mom, nsto, ne, beta,
                                  emin, emax)
        assert rc_a == 0 and rc_n == 0, (rc_a, rc_n)
        say("  mu %.15f / %.15f, N %.12f / %.12f" % (mu_a, mu_n, n_a, n_n))
        assert mu_n == mu_a and n_n == n_a
        assert abs(n_n - ne) <= 1e-11 * ne
        # the oracle's N(mu) at the returned mu meets the search's tolerance
        n_r = O.num_elec_cheby(mu_n, mom / nsto, L, emin, emax, beta, 2.0)
        assert abs(n_r - ne) <= 1e-9 * ne, (n_r, ne)


def test_chem_pot_noalloc_refuses_degree_above_scratch(ctx):
    say("case: chem_pot_noalloc L=8193 refused")
    mom, emin, emax = _moments(8193, 4.0, 2)
    rc, _, _ = _search(ctx, ctx.lib.sdft_hip_chem_pot_noalloc, 8193, mom, 4.0, 20.0, 40.0,
                       emin, emax)
    assert rc == -1
    assert b"exceeds the fixed scratch" in ctx.lib.sdft_hip_last_error()
