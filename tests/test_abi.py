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


# The reference's filter-path prototypes (stodft/proto_stodft_local.h:14-195)
# that the shim stands in for.  PINY defines these names itself with its
# CP / CLASS / GENERAL_DATA signatures, so the shim must not export them.
REFERENCE_NAMES = ["filterChebyPolyHerm", "normHCheby", "calcCoefForceWrapSCF", "calcChebyCoeff",
                   "calcChebyMoments", "calcNumElecCheby", "calcChemPotCheby",
                   "calcRhoStoHybridCheby", "calcKSPot", "filterNewtonPolyHerm", "genEnergyMax",
                   "genEnergyMin", "normHNewtonHerm", "genNewtonHermit", "genChebyHermit",
                   "calcEnergyChemPot", "calcEnergyForce", "genNoiseOrbitalReal",
                   "genStoOrbitalCheby", "rhoCalcRealStoHybrid"]


def _dynamic_exports(path):
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                         check=True).stdout
    return sorted(line.split()[-1] for line in out.splitlines() if line.strip())


def test_host_shim_exports_prefixed_entry_points(lib_path):
    """Every reference entry point the shim covers is declared and exported as
    sdft_<name>(SDFT_STODFT*, ...); the library exports nothing without the
    sdft_ prefix, so none of its symbols can clash with PINY's own
    definitions of the unprefixed names (whatever their signatures)."""
    from sdft_amd import build
    src = open(os.path.join(REPO, "include", "sdft_stodft.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    decl = sorted(set(re.findall(r"\b(?:void|double)\s+(\w+)\s*\(SDFT_STODFT", src)))
    covered = ["filterChebyPolyHerm", "normHCheby", "calcCoefForceWrapSCF", "calcChebyCoeff",
               "calcChebyMoments", "calcNumElecCheby", "calcChemPotCheby",
               "calcRhoStoHybridCheby", "calcKSPot", "filterNewtonPolyHerm", "genEnergyMax",
               "genEnergyMin"]
    assert decl == sorted("sdft_" + n for n in covered)
    exports = _dynamic_exports(build.host_lib_path())
    assert set(decl) <= set(exports)
    unprefixed = [e for e in exports if not e.startswith("sdft_")]
    assert not unprefixed, unprefixed
    assert not set(REFERENCE_NAMES) & set(exports)
    # the raw-force entry point is a C function of the declared arity
    host = ctypes.CDLL(build.host_lib_path())
    assert all(hasattr(host, n) for n in decl)
    assert not any(hasattr(host, n) for n in REFERENCE_NAMES)


def test_device_library_c_exports_are_prefixed(lib_path):
    """libsdft_hip.so: every C-linkage export is an sdft_hip_* entry point
    (C++ symbols are mangled inside namespace sdft), so none collides with a
    reference name either."""
    exports = _dynamic_exports(lib_path)
    c_names = [e for e in exports if not e.startswith("_Z") and not e.startswith("__hip_cuid_")
               and e not in ("_init", "_fini", "_edata", "_end")]
    assert c_names and all(e.startswith("sdft_hip_") for e in c_names), c_names
    # C++ exports: the library's own namespace or standard-library template
    # instances (weak, identical wherever they are instantiated)
    mangled = [e for e in exports if e.startswith("_Z")]
    odd = [e for e in mangled if "4sdft" not in e and not e.startswith(("_ZNSt", "_ZSt"))]
    assert not odd, odd[:5]
    assert not set(REFERENCE_NAMES) & set(exports)


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
