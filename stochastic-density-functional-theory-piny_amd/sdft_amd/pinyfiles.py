"""ctypes binding of the reference's on-disk text formats (include/sdft_io.h,
host C in libsdft_host.so): the SCF density checkpoint
(stodft/checkpointIO.c:32-505), the density log / final density
(stodft/density.c:1656-1745) and the fragment orbital files
(stodft/fragment/frag-scf.c:254-265).

Grid arrays are 1-D float64 (the caller's real-space order); orbital blocks
are complex (nstate, ncoef) arrays with G=0 last, as in sdft_amd.hip.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

EINVAL, EOPEN, EFORMAT, EWRITE = -1, -10, -11, -12

EXPORTS = [
    "sdft_io_last_error", "sdft_io_checkpoint_nbank", "sdft_io_checkpoint_write",
    "sdft_io_checkpoint_header", "sdft_io_checkpoint_read", "sdft_io_density_write",
    "sdft_io_density_log_reset", "sdft_io_density_read", "sdft_io_frag_mo_write",
    "sdft_io_frag_mo_read",
]

_dp = C.POINTER(C.c_double)
_dpp = C.POINTER(_dp)
_ip = C.POINTER(C.c_int)
_LIB = None


class PinyFileError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__("sdft_io error %d: %s" % (code, msg))
        self.code = code


def load():
    """Load libsdft_host.so (once); raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    p = _build.host_lib_path()
    if not os.path.exists(p):
        raise PinyFileError(EINVAL, "libsdft_host.so not built at %s" % p)
    lib = C.CDLL(p)
    cp, ll, d, i = C.c_char_p, C.c_longlong, C.c_double, C.c_int
    sig = {
        "sdft_io_last_error": ([], cp),
        "sdft_io_checkpoint_nbank": ([i, i], i),
        "sdft_io_checkpoint_write": ([cp, i, i, i, ll, d, i, _dp, _dp, _dpp, _dpp, _dpp, _dpp], i),
        "sdft_io_checkpoint_header": ([cp, _ip, _ip], i),
        "sdft_io_checkpoint_read": ([cp, i, ll, i, _ip, _ip, _dp, _dp, _dpp, _dpp, _dpp, _dpp], i),
        "sdft_io_density_write": ([cp, cp, ll, i, _dp, _dp, d], i),
        "sdft_io_density_log_reset": ([cp], i),
        "sdft_io_density_read": ([cp, ll, i, _dp, _dp], i),
        "sdft_io_frag_mo_write": ([cp, i, _ip, _ip, _dpp, _dpp], i),
        "sdft_io_frag_mo_read": ([cp, i, _ip, _ip, _dpp, _dpp], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _LIB = lib
    return lib


def _check(rc: int):
    if rc != 0:
        raise PinyFileError(rc, load().sdft_io_last_error().decode(errors="replace"))


def _path(p):
    return None if p is None else os.fsencode(p)


def _f64(a, n=None):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if n is not None and a.size != n:
        raise ValueError("expected %d values, got %d" % (n, a.size))
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_dp)


def _ptrs(arrs):
    """A C array of double* over the given float64 arrays (None -> NULL)."""
    if arrs is None:
        return None
    out = (_dp * max(len(arrs), 1))()
    for k, a in enumerate(arrs):
        out[k] = _ptr(a)
    return C.cast(out, _dpp), out


def checkpoint_nbank(ndiis_now: int, ndiis: int) -> int:
    return load().sdft_io_checkpoint_nbank(ndiis_now, ndiis)


def checkpoint_write(path, iscf, ndiis, vol, rho_up, bank_up, err_up, rho_dn=None,
                     bank_dn=None, err_dn=None):
    """checkpointOutput: len(err_up) stored DIIS steps, bank_up holding
    checkpoint_nbank(len(err_up), ndiis) densities."""
    rho_up = _f64(rho_up)
    n = rho_up.size
    lsda = rho_dn is not None
    bu = [_f64(b, n) for b in bank_up]
    eu = [_f64(e, n) for e in err_up]
    # the C writer reads exactly these many pointers from each array
    nbank = checkpoint_nbank(len(eu), ndiis) if len(eu) <= ndiis else -1
    if nbank < 0 or len(bu) != nbank:
        raise ValueError("checkpoint: %d error entries and %d bank entries do not fit a bank "
                         "of %d (need %d stored steps <= %d and nbank entries)"
                         % (len(eu), len(bu), ndiis, len(eu), ndiis))
    if lsda and (bank_dn is None or err_dn is None or len(bank_dn) != nbank
                 or len(err_dn) != len(eu)):
        raise ValueError("checkpoint: the down-spin banks must match the up-spin counts")
    keep = [rho_up, bu, eu]
    args = [_ptrs(bu), _ptrs(eu)]
    rd = None
    if lsda:
        rd = _f64(rho_dn, n)
        bd = [_f64(b, n) for b in bank_dn]
        ed = [_f64(e, n) for e in err_dn]
        keep += [rd, bd, ed]
        args += [_ptrs(bd), _ptrs(ed)]
    else:
        args += [None, None]
    cargs = [a[0] if a is not None else None for a in args]
    _check(load().sdft_io_checkpoint_write(_path(path), iscf, len(eu), ndiis, n, vol, int(lsda),
                                           _ptr(rho_up), _ptr(rd), *cargs))
    del keep, args


def checkpoint_header(path):
    it, nn = C.c_int(), C.c_int()
    _check(load().sdft_io_checkpoint_header(_path(path), C.byref(it), C.byref(nn)))
    return it.value, nn.value


def checkpoint_read(path, ndiis, ngrid, lsda=False):
    """checkpointInput: returns dict(iscf, ndiis_now, rho_up, bank_up, err_up
    [, rho_dn, bank_dn, err_dn]) with the stored counts of bank entries."""
    rho_up = np.zeros(ngrid)
    rho_dn = np.zeros(ngrid) if lsda else None
    banks = [[np.zeros(ngrid) for _ in range(ndiis)] for _ in range(4 if lsda else 2)]
    ptrs = [_ptrs(b) for b in banks] + ([None, None] if not lsda else [])
    it, nn = C.c_int(), C.c_int()
    cargs = [p[0] if p is not None else None for p in ptrs]
    _check(load().sdft_io_checkpoint_read(_path(path), ndiis, ngrid, int(lsda), C.byref(it),
                                          C.byref(nn), _ptr(rho_up), _ptr(rho_dn), *cargs))
    nb = checkpoint_nbank(nn.value, ndiis)
    out = dict(iscf=it.value, ndiis_now=nn.value, rho_up=rho_up, bank_up=banks[0][:nb],
               err_up=banks[1][:nn.value])
    if lsda:
        out.update(rho_dn=rho_dn, bank_dn=banks[2][:nb], err_dn=banks[3][:nn.value])
    return out


def density_log_reset(log_path):
    _check(load().sdft_io_density_log_reset(_path(log_path)))


def density_write(log_path, final_path, rho_up, rho_dn=None, log_scale=1.0):
    """outputDensity: append rho * log_scale to the log, rewrite the final file."""
    rho_up = _f64(rho_up)
    rd = None if rho_dn is None else _f64(rho_dn, rho_up.size)
    _check(load().sdft_io_density_write(_path(log_path), _path(final_path), rho_up.size,
                                        int(rd is not None), _ptr(rho_up), _ptr(rd), log_scale))


def density_read(final_path, ngrid, lsda=False):
    up = np.zeros(ngrid)
    dn = np.zeros(ngrid) if lsda else None
    _check(load().sdft_io_density_read(_path(final_path), ngrid, int(lsda), _ptr(up), _ptr(dn)))
    return (up, dn) if lsda else up


def frag_mo_write(path, blocks):
    """frag-MO-<rank>: blocks is a list of complex (nstate, ncoef) arrays."""
    re = [_f64(np.real(b)) for b in blocks]
    im = [_f64(np.imag(b)) for b in blocks]
    nc = (C.c_int * max(len(blocks), 1))(*[np.shape(b)[1] for b in blocks])
    ns = (C.c_int * max(len(blocks), 1))(*[np.shape(b)[0] for b in blocks])
    pr, pi = _ptrs(re), _ptrs(im)
    _check(load().sdft_io_frag_mo_write(_path(path), len(blocks), nc, ns, pr[0], pi[0]))


def frag_mo_read(path, shapes):
    """Read fragments of the given (nstate, ncoef) shapes; complex arrays."""
    re = [np.zeros(int(s[0]) * int(s[1])) for s in shapes]
    im = [np.zeros(int(s[0]) * int(s[1])) for s in shapes]
    nc = (C.c_int * max(len(shapes), 1))(*[int(s[1]) for s in shapes])
    ns = (C.c_int * max(len(shapes), 1))(*[int(s[0]) for s in shapes])
    pr, pi = _ptrs(re), _ptrs(im)
    _check(load().sdft_io_frag_mo_read(_path(path), len(shapes), nc, ns, pr[0], pi[0]))
    return [(r + 1j * i).reshape(int(s[0]), int(s[1])) for r, i, s in zip(re, im, shapes)]
