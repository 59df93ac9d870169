"""The C ABI library loads without a GPU and exports every entry point that
include/sdft_hip.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(REPO, "include", "sdft_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(sdft_hip_\w+)\s*\(", src)))


@pytest.fixture(scope="module")
def lib_path():
    from sdft_amd import build
    p = build.lib_path()
    if not os.path.exists(p):
        build.build()
    return p


def test_header_declares_the_filter_path():
    names = declared_functions()
    for must in ("sdft_hip_init", "sdft_hip_hpsi", "sdft_hip_cheby_filter", "sdft_hip_cheby_moments",
                 "sdft_hip_chem_pot", "sdft_hip_cheby_coeffs", "sdft_hip_density",
                 "sdft_hip_energies", "sdft_hip_finalize"):
        assert must in names


def test_library_exports_all_declared(lib_path):
    lib = ctypes.CDLL(lib_path)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_python_binding_covers_all_declared():
    from sdft_amd import hip
    assert sorted(hip.EXPORTS) == declared_functions()


def test_host_shim_exports_reference_entry_points(lib_path):
    from sdft_amd import build
    src = open(os.path.join(REPO, "include", "sdft_stodft.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(?:void|double)\s+(\w+)\s*\(SDFT_STODFT", src)))
    assert names == sorted(["filterChebyPolyHerm", "normHCheby", "calcCoefForceWrapSCF",
                            "calcChebyCoeff", "calcChebyMoments", "calcNumElecCheby",
                            "calcChemPotCheby", "calcRhoStoHybridCheby", "calcKSPot",
                            "filterNewtonPolyHerm", "genEnergyMax", "genEnergyMin"])
    host = ctypes.CDLL(build.host_lib_path())
    assert all(hasattr(host, n) for n in names)


def test_last_error_without_gpu(lib_path):
    lib = ctypes.CDLL(lib_path)
    lib.sdft_hip_last_error.restype = ctypes.c_char_p
    assert isinstance(lib.sdft_hip_last_error(), bytes)
    # argument validation happens before any device call
    lib.sdft_hip_init.restype = ctypes.c_int
    assert lib.sdft_hip_init(None, None) == -1
    assert b"null" in lib.sdft_hip_last_error()


def test_host_library_exports_file_formats(lib_path):
    """include/sdft_io.h: every declared sdft_io_* entry point is exported by
    libsdft_host.so and bound in sdft_amd.pinyfiles."""
    from sdft_amd import build, pinyfiles
    src = open(os.path.join(REPO, "include", "sdft_io.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(sdft_io_\w+)\s*\(", src)))
    assert names == sorted(pinyfiles.EXPORTS)
    host = ctypes.CDLL(build.host_lib_path())
    missing = [n for n in names if not hasattr(host, n)]
    assert not missing, missing
