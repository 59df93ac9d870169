"""ctypes binding of libsdft_hip.so (include/sdft_hip.h).

The library must have been built (sdft_amd.build.build()); importing this
module never falls back to a CPU path: a missing library or a failing call
raises.  Orbital blocks are numpy complex arrays (nstate, ncoef) with G=0
last, split into the separate FP64 real / imaginary arrays the C ABI takes.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

PREC_FP64 = 0
PREC_FP32 = 1
RHO_FFTW = 0
RHO_GENERIC = 1
XC_PZ_LDA, XC_PW_LDA, XC_PADE_LDA = 0, 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class SdftError(RuntimeError):
    pass


class GridDesc(C.Structure):
    _fields_ = [("n1", C.c_int), ("n2", C.c_int), ("n3", C.c_int), ("ncoef", C.c_int),
                ("map", _ip), ("map_conj", _ip), ("ak2_kinetic", _dp)]


EXPORTS = [
    "sdft_hip_init", "sdft_hip_finalize", "sdft_hip_last_error", "sdft_hip_set_vks",
    "sdft_hip_set_projectors", "sdft_hip_hpsi", "sdft_hip_cheby_coeffs", "sdft_hip_cheby_filter",
    "sdft_hip_cheby_moments", "sdft_hip_chem_pot", "sdft_hip_chem_pot_noalloc",
    "sdft_hip_density", "sdft_hip_energies",
    "sdft_hip_block_create", "sdft_hip_block_destroy", "sdft_hip_block_set",
    "sdft_hip_block_filter", "sdft_hip_block_get", "sdft_hip_set_profiling",
    "sdft_hip_kernel_times", "sdft_hip_selftest_fft", "sdft_hip_selftest_copy",
    "sdft_hip_nl_forces", "sdft_hip_noise", "sdft_hip_block_noise",
    "sdft_hip_batch_create", "sdft_hip_batch_destroy", "sdft_hip_batch_filter",
    "sdft_hip_batch_autotune", "sdft_hip_batch_config", "sdft_hip_set_projectors_real",
    "sdft_hip_set_nl_form", "sdft_hip_set_projector_gradients",
    "sdft_hip_spectral_bound", "sdft_hip_newton_filter", "sdft_hip_tune", "sdft_hip_set_config",
    "sdft_hip_ks_init", "sdft_hip_ks_potential", "sdft_hip_local_forces",
]

_LIB = None


def load(path: str | None = None):
    """Load libsdft_hip.so (once).  Raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    p = path or _build.lib_path()
    if not os.path.exists(p):
        raise SdftError("libsdft_hip.so not built at %s: run __graft_entry__.build()" % p)
    lib = C.CDLL(p)
    vp = C.c_void_p
    sig = {
        "sdft_hip_init": ([C.POINTER(GridDesc), C.POINTER(vp)], C.c_int),
        "sdft_hip_finalize": ([vp], C.c_int),
        "sdft_hip_last_error": ([], C.c_char_p),
        "sdft_hip_set_vks": ([vp, _dp], C.c_int),
        "sdft_hip_set_projectors": ([vp, C.c_int, _dp, _dp, _dp], C.c_int),
        "sdft_hip_hpsi": ([vp, C.c_int, _dp, _dp, _dp, _dp, C.c_int], C.c_int),
        "sdft_hip_cheby_coeffs": ([vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                   C.c_int, _dp], C.c_int),
        "sdft_hip_cheby_filter": ([vp, C.c_int, C.c_int, _dp, C.c_double, C.c_double, _dp, _dp,
                                   _dp, _dp, _dp, C.c_int], C.c_int),
        "sdft_hip_cheby_moments": ([vp, C.c_int, C.c_int, C.c_double, C.c_double, _dp, _dp, _dp,
                                    C.c_int], C.c_int),
        "sdft_hip_chem_pot": ([vp, C.c_int, C.c_double, C.c_double, C.c_double, _dp, C.c_double,
                               C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _dp,
                               _dp], C.c_int),
        "sdft_hip_chem_pot_noalloc": ([vp, C.c_int, C.c_double, C.c_double, C.c_double, _dp,
                                       C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                       C.c_double, _dp, _dp], C.c_int),
        "sdft_hip_density": ([vp, C.c_int, _dp, _dp, C.c_double, _dp, C.c_int, C.c_int], C.c_int),
        "sdft_hip_energies": ([vp, C.c_int, _dp, _dp, C.c_double, C.c_double, _dp, _dp], C.c_int),
        "sdft_hip_block_create": ([vp, C.c_int, C.c_int, C.POINTER(vp)], C.c_int),
        "sdft_hip_block_destroy": ([vp], C.c_int),
        "sdft_hip_block_set": ([vp, _dp, _dp], C.c_int),
        "sdft_hip_block_filter": ([vp, C.c_int, _dp, C.c_double, C.c_double, _dp], C.c_int),
        "sdft_hip_block_get": ([vp, C.c_int, _dp, _dp], C.c_int),
        "sdft_hip_set_profiling": ([vp, C.c_int], C.c_int),
        "sdft_hip_kernel_times": ([vp, C.c_int, C.POINTER(C.c_char_p), _dp, _ip], C.c_int),
        "sdft_hip_selftest_fft": ([C.c_int, C.c_int, C.c_int, C.c_int, _dp, _dp], C.c_int),
        "sdft_hip_selftest_copy": ([C.c_longlong, C.c_int], C.c_int),
        "sdft_hip_nl_forces": ([vp, C.c_int, _dp, _dp, _dp, _ip, C.c_int, C.c_double, C.c_double,
                                _dp], C.c_int),
        "sdft_hip_noise": ([vp, C.c_int, C.c_longlong, C.c_longlong, _dp, _dp], C.c_int),
        "sdft_hip_block_noise": ([vp, C.c_longlong, C.c_longlong], C.c_int),
        "sdft_hip_batch_create": ([C.c_int, C.POINTER(vp), C.POINTER(vp)], C.c_int),
        "sdft_hip_batch_destroy": ([vp], C.c_int),
        "sdft_hip_batch_filter": ([vp, C.c_int, _dp, _dp, _dp, C.c_int], C.c_int),
        "sdft_hip_batch_autotune": ([vp, C.c_int, _ip, C.c_int, _dp], C.c_int),
        "sdft_hip_batch_config": ([vp, C.c_int, C.c_int, C.c_int, C.c_int], C.c_int),
        "sdft_hip_set_projectors_real": ([vp, C.c_int, C.POINTER(C.c_longlong), _ip, _dp, _dp,
                                          C.c_double], C.c_int),
        "sdft_hip_set_nl_form": ([vp, C.c_int], C.c_int),
        "sdft_hip_set_projector_gradients": ([vp, C.c_int, _dp, _dp, _dp], C.c_int),
        "sdft_hip_spectral_bound": ([vp, C.c_int, C.c_double, C.c_double, C.c_int, _dp, _dp, _dp,
                                     _ip, C.c_int], C.c_int),
        "sdft_hip_newton_filter": ([vp, C.c_int, C.c_int, C.c_int, _dp, _dp, C.c_double,
                                    C.c_double, _dp, _dp, _dp, _dp, C.c_int], C.c_int),
        "sdft_hip_tune": ([vp, C.c_int, C.c_int, C.c_int, _ip, C.c_int, _dp], C.c_int),
        "sdft_hip_set_config": ([vp, C.c_int, C.c_int, C.c_int], C.c_int),
        "sdft_hip_ks_init": ([vp, C.POINTER(GridDesc)], C.c_int),
        "sdft_hip_local_forces": ([vp, C.c_int, _dp, _ip, C.c_int, _dp, _dp, _dp], C.c_int),
        "sdft_hip_ks_potential": ([vp, _dp, C.c_int, _dp, _dp, C.c_double, C.c_int, _dp, _dp],
                                  C.c_int),
    }
    for name, (args, res) in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = res
    _LIB = lib
    return lib


def _ptr(a):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def _check(rc):
    if rc != 0:
        msg = load().sdft_hip_last_error().decode(errors="replace")
        raise SdftError("sdft_hip error %d: %s" % (rc, msg))


def _split(Cz):
    Cz = np.ascontiguousarray(Cz)
    return np.ascontiguousarray(Cz.real, dtype=np.float64), np.ascontiguousarray(Cz.imag, dtype=np.float64)


def selftest_fft(x: np.ndarray, direction: int, prec: int = PREC_FP64) -> np.ndarray:
    """Transform the rows of complex x (nlines, n) with the library's LDS FFT."""
    lib = load()
    x = np.ascontiguousarray(x, dtype=np.complex128)
    nl, n = x.shape
    buf = x.view(np.float64).copy()
    out = np.empty_like(buf)
    _check(lib.sdft_hip_selftest_fft(n, nl, direction, prec, _ptr(buf), _ptr(out)))
    return out.view(np.complex128).reshape(nl, n)


def selftest_copy(nbytes: int, reps: int = 1) -> None:
    """Run the library's plain copy kernel (byte-counter calibration)."""
    _check(load().sdft_hip_selftest_copy(int(nbytes), int(reps)))


class Context:
    """One system on the current HIP device (grid, G-set, v_KS, KB table)."""

    def __init__(self, sysm, with_nl: bool = True):
        self.lib = load()
        self.sys = sysm
        self._map = np.ascontiguousarray(sysm.map, dtype=np.int32)
        self._mapc = np.ascontiguousarray(sysm.map_conj, dtype=np.int32)
        self._ak2 = np.ascontiguousarray(sysm.ak2, dtype=np.float64)
        d = GridDesc(sysm.n[0], sysm.n[1], sysm.n[2], sysm.ncoef,
                     self._map.ctypes.data_as(_ip), self._mapc.ctypes.data_as(_ip),
                     self._ak2.ctypes.data_as(_dp))
        h = C.c_void_p()
        _check(self.lib.sdft_hip_init(C.byref(d), C.byref(h)))
        self.h = h
        self.set_vks(sysm.vks)
        if with_nl and sysm.nproj > 0:
            self.set_projectors(sysm.proj, sysm.proj_energy)

    def close(self):
        if getattr(self, "h", None):
            self.lib.sdft_hip_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_vks(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        _check(self.lib.sdft_hip_set_vks(self.h, _ptr(v)))

    def set_projectors(self, proj, energy):
        pr = np.ascontiguousarray(proj.real, dtype=np.float64)
        pi = np.ascontiguousarray(proj.imag, dtype=np.float64)
        e = np.ascontiguousarray(energy, dtype=np.float64)
        self._rs_atom = None
        _check(self.lib.sdft_hip_set_projectors(self.h, pr.shape[0], _ptr(pr), _ptr(pi), _ptr(e)))

    def set_projectors_real(self, rs):
        """Real-space KB projectors (systems.RealSpaceProjectors)."""
        off = np.ascontiguousarray(rs.offsets, dtype=np.int64)
        idx = np.ascontiguousarray(rs.index, dtype=np.int32)
        val = np.ascontiguousarray(rs.values, dtype=np.float64)
        e = np.ascontiguousarray(rs.energy, dtype=np.float64)
        self._rs_energy = e
        self._rs_atom = np.ascontiguousarray(rs.atom, dtype=np.int32)
        _check(self.lib.sdft_hip_set_projectors_real(
            self.h, len(e), off.ctypes.data_as(C.POINTER(C.c_longlong)),
            idx.ctypes.data_as(_ip), _ptr(val), _ptr(e), float(rs.vol_elem)))

    def set_projector_gradients(self, grad):
        """d beta / d r tables (3, nent) aligned with the real-space projector
        values (systems.real_space_projector_gradients)."""
        g = [np.ascontiguousarray(grad[j], dtype=np.float64) for j in range(3)]
        self._grad = g
        _check(self.lib.sdft_hip_set_projector_gradients(self.h, len(self._rs_energy),
                                                         _ptr(g[0]), _ptr(g[1]), _ptr(g[2])))

    def set_nl_form(self, form):
        """Real-space projectors: 0 per-atom form on the grid, 1 G-space table."""
        _check(self.lib.sdft_hip_set_nl_form(self.h, int(form)))

    def clear_projectors(self):
        self._rs_atom = None
        _check(self.lib.sdft_hip_set_projectors(self.h, 0, None, None, None))

    def hpsi_raw(self, Cz, prec=PREC_FP64):
        re, im = _split(Cz)
        fre = np.empty_like(re)
        fim = np.empty_like(im)
        _check(self.lib.sdft_hip_hpsi(self.h, re.shape[0], _ptr(re), _ptr(im), _ptr(fre),
                                      _ptr(fim), prec))
        return fre + 1j * fim

    def tune(self, nstate, prec=PREC_FP64, steps=6):
        """sdft_hip_tune: returns ((plane_tile, kb_wg_target, fused_expand),
        per-candidate ms per filter step)."""
        cfg = np.zeros(3, dtype=np.int32)
        ms = np.zeros(64)
        n = self.lib.sdft_hip_tune(self.h, nstate, prec, steps, cfg.ctypes.data_as(_ip), 64,
                                   _ptr(ms))
        if n < 0:
            _check(n)
        return (int(cfg[0]), int(cfg[1]), int(cfg[2])), ms[:min(n, 64)].copy()

    def ks_init(self, dens):
        """sdft_hip_ks_init: the density plane-wave sphere (systems.DensitySphere)."""
        self._dmap = np.ascontiguousarray(dens.map, dtype=np.int32)
        self._dmapc = np.ascontiguousarray(dens.map_conj, dtype=np.int32)
        self._dak2 = np.ascontiguousarray(dens.ak2, dtype=np.float64)
        d = GridDesc(dens.n[0], dens.n[1], dens.n[2], dens.ncoef,
                     self._dmap.ctypes.data_as(_ip), self._dmapc.ctypes.data_as(_ip),
                     self._dak2.ctypes.data_as(_dp))
        _check(self.lib.sdft_hip_ks_init(self.h, C.byref(d)))

    def ks_potential(self, rho, vext, volume, xc=0, layout=RHO_FFTW):
        """sdft_hip_ks_potential: v_KS from the density (becomes the context's
        v_KS).  Returns (v_KS in `layout`, dict(eh, eext, exc, muxc))."""
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        if rho.size != self.sys.ngrid or np.size(vext) != len(getattr(self, "_dak2", ())):
            raise SdftError("ks_potential: rho must have n^3 points and vext one value per "
                            "density-sphere coefficient (after ks_init)")
        vr = np.ascontiguousarray(np.real(vext), dtype=np.float64)
        vi = np.ascontiguousarray(np.imag(vext), dtype=np.float64)
        en = np.zeros(4)
        out = np.empty_like(rho)
        _check(self.lib.sdft_hip_ks_potential(self.h, _ptr(rho), layout, _ptr(vr), _ptr(vi),
                                              float(volume), xc, _ptr(en), _ptr(out)))
        return out, dict(eh=en[0], eext=en[1], exc=en[2], muxc=en[3])

    def local_forces(self, pos, atom_type, vloc, gvec):
        """sdft_hip_local_forces on the density of the last ks_potential call:
        (natm, 3) local pseudopotential forces."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        at = np.ascontiguousarray(atom_type, dtype=np.int32)
        vl = np.ascontiguousarray(vloc, dtype=np.float64)
        gv = np.ascontiguousarray(gvec, dtype=np.float64)
        if vl.ndim != 2 or vl.shape[1] != len(getattr(self, "_dak2", ())) or gv.shape != (
                vl.shape[1], 3) or pos.shape != (at.shape[0], 3):
            raise SdftError("local_forces: shapes (natm,3), (natm,), (ntype,ncoef), (ncoef,3)")
        out = np.zeros_like(pos)
        _check(self.lib.sdft_hip_local_forces(self.h, pos.shape[0], _ptr(pos),
                                              at.ctypes.data_as(_ip), vl.shape[0], _ptr(vl),
                                              _ptr(gv), _ptr(out)))
        return out

    def set_config(self, plane_tile, kb_wg_target=512, fused_expand=1):
        _check(self.lib.sdft_hip_set_config(self.h, plane_tile, kb_wg_target, fused_expand))

    def newton_filter(self, chi, coeffs, samp_point, scale, emean, prec=PREC_FP64):
        """filterNewtonPolyHerm: coeffs (L, nmu), samp_point (L-1,).  Returns
        (out (nmu, nstate, ncoef), P_{L-1})."""
        re, im = _split(chi)
        co = np.ascontiguousarray(coeffs, dtype=np.float64)
        if co.ndim == 1:
            co = co[:, None]
        L, nmu = co.shape
        z = np.ascontiguousarray(samp_point, dtype=np.float64)
        assert z.shape[0] >= L - 1
        ns, nc = re.shape
        ore = np.empty((nmu, ns, nc))
        oim = np.empty((nmu, ns, nc))
        _check(self.lib.sdft_hip_newton_filter(self.h, ns, L, nmu, _ptr(co), _ptr(z), scale, emean,
                                               _ptr(re), _ptr(im), _ptr(ore), _ptr(oim), prec))
        return ore + 1j * oim, re + 1j * im

    def spectral_bound(self, v0, which, emax=0.0, conv=1e-5, max_iter=100000,
                       prec=PREC_FP64):
        """genEnergyMax (which 0) / genEnergyMin (which 1) on the device from
        the start vector v0 (ncoef complex): returns (E, final vector, iterations)."""
        re, im = _split(np.asarray(v0).reshape(1, -1))
        e = C.c_double()
        it = C.c_int()
        _check(self.lib.sdft_hip_spectral_bound(self.h, which, emax, conv, max_iter, _ptr(re),
                                                _ptr(im), C.byref(e), C.byref(it), prec))
        return e.value, (re + 1j * im)[0], it.value

    def cheby_coeffs(self, L, emin, emax, mu, beta, fun_flag=2):
        c = np.empty(L)
        _check(self.lib.sdft_hip_cheby_coeffs(self.h, L, emin, emax, mu, beta, fun_flag, _ptr(c)))
        return c

    def cheby_filter(self, chi, coeffs, scale, emean, prec=PREC_FP64, moments=False):
        re, im = _split(chi)
        ore = np.empty_like(re)
        oim = np.empty_like(im)
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        mom = np.zeros(len(coeffs)) if moments else None
        _check(self.lib.sdft_hip_cheby_filter(self.h, re.shape[0], len(coeffs), _ptr(coeffs),
                                              scale, emean, _ptr(re), _ptr(im), _ptr(ore),
                                              _ptr(oim), _ptr(mom), prec))
        out = (ore + 1j * oim, re + 1j * im)
        return out + (mom,) if moments else out

    def cheby_moments(self, chi, L, scale, emean, prec=PREC_FP64):
        re, im = _split(chi)
        mom = np.zeros(L + 1)
        _check(self.lib.sdft_hip_cheby_moments(self.h, re.shape[0], L, scale, emean, _ptr(re),
                                               _ptr(im), _ptr(mom), prec))
        return mom

    def chem_pot(self, L, emin, emax, beta, moments, nsto, ne_true, mu_init, gap, occ=2.0,
                 tol=None):
        m = np.ascontiguousarray(moments, dtype=np.float64)
        tol = tol if tol is not None else 1e-11 * ne_true
        mu = C.c_double()
        ne = C.c_double()
        _check(self.lib.sdft_hip_chem_pot(self.h, L, emin, emax, beta, _ptr(m), nsto, ne_true,
                                          mu_init, gap, occ, tol, C.byref(mu), C.byref(ne)))
        return mu.value, ne.value

    def density(self, Cz, occ_over_nsto, layout=RHO_FFTW, accumulate_into=None):
        """rho in FFTW or generic order; with accumulate_into (an FP64 array of
        ngrid) the density is added to it in place and it is returned."""
        re, im = _split(Cz)
        if accumulate_into is None:
            rho = np.zeros(self.sys.ngrid)
        else:
            rho = accumulate_into
            assert rho.dtype == np.float64 and rho.flags.c_contiguous and rho.size == self.sys.ngrid
        _check(self.lib.sdft_hip_density(self.h, re.shape[0], _ptr(re), _ptr(im), occ_over_nsto,
                                         _ptr(rho), layout, 0 if accumulate_into is None else 1))
        return rho

    def energies(self, Cz, occ, nsto):
        re, im = _split(Cz)
        eke = C.c_double()
        enl = C.c_double()
        _check(self.lib.sdft_hip_energies(self.h, re.shape[0], _ptr(re), _ptr(im), occ, nsto,
                                          C.byref(eke), C.byref(enl)))
        return eke.value, enl.value

    def noise(self, nstate, seed, first=0):
        """Stochastic orbitals generated on the device (genNoiseOrbitalReal)."""
        nc = self.sys.ncoef
        re = np.empty((nstate, nc))
        im = np.empty((nstate, nc))
        _check(self.lib.sdft_hip_noise(self.h, nstate, seed, first, _ptr(re), _ptr(im)))
        return re + 1j * im

    def nl_forces(self, Cz, occ, nsto):
        """Hellmann-Feynman nonlocal forces (natm, 3) of the orbitals Cz."""
        s = self.sys
        re, im = _split(Cz)
        g = np.ascontiguousarray(s.gvec, dtype=np.float64)
        # the atom of each projector the context holds: the real-space ones
        # when set, else the system's G-space table
        pa = self._rs_atom if getattr(self, "_rs_atom", None) is not None else \
            np.ascontiguousarray(s.proj_atom, dtype=np.int32)
        out = np.zeros((s.natm, 3))
        _check(self.lib.sdft_hip_nl_forces(self.h, re.shape[0], _ptr(re), _ptr(im), _ptr(g),
                                           pa.ctypes.data_as(_ip), s.natm, occ, nsto, _ptr(out)))
        return out

    def set_profiling(self, on: bool):
        _check(self.lib.sdft_hip_set_profiling(self.h, 1 if on else 0))

    def kernel_times(self):
        names = (C.c_char_p * 32)()
        ms = np.zeros(32)
        la = np.zeros(32, dtype=np.int32)
        n = self.lib.sdft_hip_kernel_times(self.h, 32, names, _ptr(ms), la.ctypes.data_as(_ip))
        if n < 0:
            _check(n)
        return {names[i].decode(): (float(ms[i]), int(la[i])) for i in range(min(n, 32))}


class Block:
    """A device-resident block of orbitals (sdft_hip_block_*)."""

    def __init__(self, ctx: Context, nstate: int, prec: int = PREC_FP64):
        self.ctx = ctx
        self.nstate = nstate
        h = C.c_void_p()
        _check(ctx.lib.sdft_hip_block_create(ctx.h, nstate, prec, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.ctx.lib.sdft_hip_block_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, chi):
        re, im = _split(chi)
        _check(self.ctx.lib.sdft_hip_block_set(self.h, _ptr(re), _ptr(im)))

    def noise(self, seed, first=0):
        _check(self.ctx.lib.sdft_hip_block_noise(self.h, seed, first))

    def filter(self, coeffs, scale, emean, moments=False):
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        mom = np.zeros(len(coeffs)) if moments else None
        _check(self.ctx.lib.sdft_hip_block_filter(self.h, len(coeffs), _ptr(coeffs), scale, emean,
                                                  _ptr(mom)))
        return mom

    def get(self, which=0):
        ncoef = self.ctx.sys.ncoef
        re = np.empty((self.nstate, ncoef))
        im = np.empty((self.nstate, ncoef))
        _check(self.ctx.lib.sdft_hip_block_get(self.h, which, _ptr(re), _ptr(im)))
        return re + 1j * im


class Batch:
    """A fragment batch (sdft_hip_batch_*): resident blocks of independent
    systems filtered together."""

    def __init__(self, blocks):
        self.lib = load()
        self.blocks = list(blocks)
        arr = (C.c_void_p * len(self.blocks))(*[b.h.value for b in self.blocks])
        h = C.c_void_p()
        _check(self.lib.sdft_hip_batch_create(len(self.blocks), arr, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.sdft_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def filter(self, coeffs, scale, emean, use_graph=True):
        """coeffs (nfrag, L); scale / emean (nfrag,).  Results in each block's
        accumulator (Block.get(0)).  use_graph: False / 0 direct launches,
        True / 1 the captured pair of steps replayed, 2 the whole filter as
        one graph launch."""
        coeffs = np.ascontiguousarray(coeffs, dtype=np.float64)
        nf, L = coeffs.shape
        assert nf == len(self.blocks)
        sc = np.ascontiguousarray(scale, dtype=np.float64)
        em = np.ascontiguousarray(emean, dtype=np.float64)
        assert sc.shape == (nf,) and em.shape == (nf,)
        _check(self.lib.sdft_hip_batch_filter(self.h, L, _ptr(coeffs), _ptr(sc), _ptr(em),
                                              int(use_graph)))

    def autotune(self, steps=8):
        """Returns ((split, overlap, plane_tile, fused_expand), per-candidate
        ms/step)."""
        cfg = np.zeros(4, dtype=np.int32)
        ms = np.zeros(32)
        n = self.lib.sdft_hip_batch_autotune(self.h, steps, cfg.ctypes.data_as(_ip), 32, _ptr(ms))
        if n < 0:
            _check(n)
        return tuple(int(x) for x in cfg), ms[:n].copy()

    def config(self, split, overlap, plane_tile=0, fused_expand=1):
        """Launch configuration by hand; fused_expand=1 needs every fragment's
        sphere columns to be ka intervals (true of spherical cutoffs)."""
        _check(self.lib.sdft_hip_batch_config(self.h, split, overlap, plane_tile, fused_expand))
