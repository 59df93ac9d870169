"""The reference's on-disk text formats (include/sdft_io.h, host C): the
density checkpoint (stodft/checkpointIO.c:32-505), the density log / final
density (stodft/density.c:1656-1745) and frag-MO files
(stodft/fragment/frag-scf.c:254-265).

The reference ships no such files, so the layout is pinned line by line
against the reference's fprintf formats ("%i %i", "%.16lg", "%.10lg",
"%.16lg %.16lg"), rendered independently here by Python's printf-style
formatting (the same C99 conversion, correctly rounded), plus round trips
and the error paths.  No GPU is used."""
import os

import numpy as np
import pytest


@pytest.fixture(scope="module")
def pf():
    from sdft_amd import build, pinyfiles
    if not os.path.exists(build.host_lib_path()):
        build.build()
    pinyfiles.load()
    return pinyfiles


def lines(path):
    with open(path) as f:
        return f.read().split("\n")


def g16(x):
    return "%.16g" % x


def test_checkpoint_nbank(pf):
    # checkpointIO.c:133-134: one bank entry ahead while the bank fills
    assert pf.checkpoint_nbank(0, 5) == 1
    assert pf.checkpoint_nbank(3, 5) == 4
    assert pf.checkpoint_nbank(5, 5) == 5
    assert pf.checkpoint_nbank(0, 0) == 0


@pytest.mark.parametrize("ndiis_now,ndiis", [(0, 4), (2, 4), (4, 4)])
def test_checkpoint_layout_unpolarised(pf, tmp_path, ndiis_now, ndiis):
    rng = np.random.default_rng(ndiis_now)
    n, vol = 37, 270.0117
    rho = rng.random(n) * 1e-2
    nb = pf.checkpoint_nbank(ndiis_now, ndiis)
    bank = [rng.standard_normal(n) for _ in range(nb)]
    err = [rng.standard_normal(n) * 1e-7 for _ in range(ndiis_now)]
    p = tmp_path / "density-checkpoint"
    pf.checkpoint_write(p, 7, ndiis, vol, rho, bank, err)
    want = ["%i %i" % (7, ndiis_now)] + [g16(x * vol) for x in rho]
    for b in bank:
        want += [g16(x) for x in b]
    for e in err:
        want += [g16(x) for x in e]
    assert lines(p) == want + [""]
    assert pf.checkpoint_header(p) == (7, ndiis_now)
    got = pf.checkpoint_read(p, ndiis, n)
    assert got["iscf"] == 7 and got["ndiis_now"] == ndiis_now
    # the density comes back as stored (times the volume), 16 digits
    assert np.allclose(got["rho_up"], rho * vol, rtol=2e-16 * 8, atol=0)
    assert len(got["bank_up"]) == nb and len(got["err_up"]) == ndiis_now
    for a, b in zip(got["bank_up"] + got["err_up"], bank + err):
        assert np.all(np.abs(a - b) <= 1e-15 * np.abs(b))


def test_checkpoint_layout_lsda(pf, tmp_path):
    rng = np.random.default_rng(5)
    n, vol, ndiis = 11, 3.5, 3
    up, dn = rng.random(n), rng.random(n)
    bu = [rng.random(n) for _ in range(3)]
    eu = [rng.random(n) for _ in range(2)]
    bd = [rng.random(n) + 10 for _ in range(3)]
    ed = [rng.random(n) + 20 for _ in range(2)]
    p = tmp_path / "density-checkpoint-0"
    pf.checkpoint_write(p, 2, ndiis, vol, up, bu, eu, dn, bd, ed)
    want = ["2 2"] + [g16(x * vol) for x in up] + [g16(x * vol) for x in dn]
    for blk in bu + eu + bd + ed:
        want += [g16(x) for x in blk]
    assert lines(p) == want + [""]
    got = pf.checkpoint_read(p, ndiis, n, lsda=True)
    assert np.allclose(got["rho_dn"], dn * vol, rtol=1e-15, atol=0)
    for a, b in zip(got["bank_dn"] + got["err_dn"], bd + ed):
        assert np.allclose(a, b, rtol=1e-15, atol=0)


def test_checkpoint_errors(pf, tmp_path):
    n = 8
    p = tmp_path / "ck"
    with pytest.raises(pf.PinyFileError) as e:
        pf.checkpoint_read(tmp_path / "missing", 2, n)
    assert e.value.code == pf.EOPEN
    # more stored steps than the reader's bank
    pf.checkpoint_write(p, 1, 3, 1.0, np.ones(n), [np.ones(n)] * 3, [np.ones(n)] * 3)
    with pytest.raises(pf.PinyFileError) as e:
        pf.checkpoint_read(p, 2, n)
    assert e.value.code == pf.EFORMAT
    # a file cut short
    text = open(p).read().split("\n")
    (tmp_path / "short").write_text("\n".join(text[: len(text) // 2]))
    with pytest.raises(pf.PinyFileError) as e:
        pf.checkpoint_read(tmp_path / "short", 3, n)
    assert e.value.code == pf.EFORMAT
    # no header
    (tmp_path / "bad").write_text("x y\n1\n")
    with pytest.raises(pf.PinyFileError) as e:
        pf.checkpoint_header(tmp_path / "bad")
    assert e.value.code == pf.EFORMAT
    # the writer refuses a stored count above the bank size, and bank lists
    # that do not match the stored count (before any pointer is read in C)
    with pytest.raises(ValueError):
        pf.checkpoint_write(p, 1, 1, 1.0, np.ones(n), [np.ones(n)] * 2, [np.ones(n)] * 2)
    with pytest.raises(ValueError):
        pf.checkpoint_write(p, 1, 4, 1.0, np.ones(n), [np.ones(n)] * 2, [np.ones(n)] * 2)
    # the C entry point itself rejects a stored count above the bank size
    import ctypes
    lib = pf.load()
    arr = (ctypes.POINTER(ctypes.c_double) * 2)()
    x = np.ones(n)
    rc = lib.sdft_io_checkpoint_write(str(p).encode(), 1, 2, 1, n, 1.0, 0,
                                      x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), None,
                                      ctypes.cast(arr, ctypes.POINTER(ctypes.POINTER(ctypes.c_double))),
                                      ctypes.cast(arr, ctypes.POINTER(ctypes.POINTER(ctypes.c_double))),
                                      None, None)
    assert rc == pf.EINVAL


def test_density_log_and_final(pf, tmp_path):
    rng = np.random.default_rng(9)
    n, vol = 23, 123.25
    log, fin = tmp_path / "dens.log", tmp_path / "dens.final"
    pf.density_log_reset(log)
    assert lines(log) == [""]
    r1, r2 = rng.random(n), rng.random(n)
    pf.density_write(log, fin, r1, log_scale=1.0 / vol)
    pf.density_write(log, fin, r2, log_scale=1.0 / vol)
    # the log appends every call (%.10lg of rho / vol); the final file holds
    # the last density (%.16lg)
    want_log = ["%.10g" % (x * (1.0 / vol)) for x in np.concatenate([r1, r2])]
    assert lines(log) == want_log + [""]
    assert lines(fin) == [g16(x) for x in r2] + [""]
    back = pf.density_read(fin, n)
    assert np.all(np.abs(back - r2) <= 1e-15 * r2)
    # spin-polarised: up then down
    pf.density_write(None, fin, r1, r2)
    up, dn = pf.density_read(fin, n, lsda=True)
    assert np.allclose(up, r1, rtol=1e-15, atol=0) and np.allclose(dn, r2, rtol=1e-15, atol=0)
    with pytest.raises(pf.PinyFileError) as e:
        pf.density_read(fin, 3 * n)
    assert e.value.code == pf.EFORMAT


def test_frag_mo_round_trip(pf, tmp_path):
    rng = np.random.default_rng(3)
    shapes = [(3, 17), (0, 5), (2, 40)]
    blocks = [rng.standard_normal(s) + 1j * rng.standard_normal(s) for s in shapes]
    p = tmp_path / "frag-MO-0"
    pf.frag_mo_write(p, blocks)
    want = []
    for b in blocks:
        for z in b.ravel():
            want.append("%s %s" % (g16(z.real), g16(z.imag)))
    assert lines(p) == want + [""]
    back = pf.frag_mo_read(p, shapes)
    for a, b in zip(back, blocks):
        assert a.shape == b.shape
        assert np.all(np.abs(a - b) <= 1e-15 * np.abs(b) + 1e-300)
    with pytest.raises(pf.PinyFileError) as e:
        pf.frag_mo_read(p, shapes + [(1, 4)])
    assert e.value.code == pf.EFORMAT
