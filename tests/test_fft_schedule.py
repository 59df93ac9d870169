"""CPU checks of the LDS FFT's index arithmetic (csrc/fft_lds.h, host side).

* The radix schedules: for every supported line length and both schedules
  (0: radix 8 first, 1: radix 4 first) the radices multiply to n, the
  reversed order used by the forward transforms is the same set, a pruned
  first / last pass (ZIN / ZOUT) is radix 4 or 8 where the plane kernel uses
  it, and the two schedules differ exactly where the power of two in n is
  2^6 (64, 192; 128 = 2^7 is 8 x 4 x 4 either way).  Compiled
  with g++ against the header itself.
* The thread -> (line, butterfly) map for a runtime line count: the
  quotient through a single-precision reciprocal, trunc((idx + 1/2) rcp(n)),
  equals idx // n for every idx < 2^16 and n <= 1024, with rcp(n) the
  correctly rounded 1/n or one ulp off either way (v_rcp_f32's bound).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(REPO, "stochastic-density-functional-theory-piny_amd", "csrc")

SIZES = (24, 32, 36, 48, 64, 96, 128, 192)

PROGRAM = r"""
#include <cstdio>
// host compile of the header: its device pass templates are never
// instantiated here, only named
#define __syncthreads() ((void)0)
#define __builtin_amdgcn_rcpf(x) (1.0f / (x))
#include "fft_lds.h"
using namespace sdft;
int main() {
  const int sizes[] = {24, 32, 36, 48, 64, 96, 128, 192};
  for (int n : sizes)
    for (int s = 0; s < 2; ++s) {
      std::printf("%d %d %d", n, s, n_passes(n, s));
      for (int i = 0; i < n_passes(n, s); ++i)
        std::printf(" %d/%d/%d", radix_at(n, i, s), radix_sched(n, i, true, s),
                    ns_sched(n, i, false, s));
      std::printf(" | %d\n", fft_max_lines(n, 256, s));
    }
  return 0;
}
"""


@pytest.fixture(scope="module")
def schedules(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("fftsched")
    src = d / "sched.cpp"
    exe = d / "sched"
    src.write_text(PROGRAM)
    subprocess.run([gxx, "-std=c++17", "-O0", "-I", CSRC, str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    table = {}
    for line in out.strip().splitlines():
        head, ml = line.split("|")
        f = head.split()
        n, s, p = int(f[0]), int(f[1]), int(f[2])
        passes = [tuple(int(x) for x in t.split("/")) for t in f[3:]]
        assert len(passes) == p
        table[(n, s)] = (passes, int(ml))
    return table


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("s", (0, 1))
def test_radix_schedule_factors_n(schedules, n, s):
    passes, ml = schedules[(n, s)]
    radices = [r for r, _, _ in passes]
    assert int(np.prod(radices)) == n
    assert all(r in (2, 3, 4, 8) for r in radices)
    # the reversed schedule is the same radices in reverse order
    assert [rr for _, rr, _ in passes] == radices[::-1]
    # ns_sched: product of the radices before pass i
    assert [ns for _, _, ns in passes] == [int(np.prod(radices[:i])) for i in range(len(radices))]
    # powers of two before the radix-3 passes
    k3 = [i for i, r in enumerate(radices) if r == 3]
    assert k3 == list(range(len(radices) - len(k3), len(radices)))
    # the pruned first inverse pass and last forward pass (reversed order)
    # are radix 4 or 8 on every size the plane kernel prunes (n >= 32)
    if n >= 32:
        assert radices[0] in (4, 8)
    # every thread of a 256-thread workgroup has at most two butterflies of
    # the smallest radix
    assert ml * n // min(radices) <= 2 * 256 or ml == 1


@pytest.mark.parametrize("n", SIZES)
def test_schedules_differ_only_at_two_to_the_sixth(schedules, n):
    r0 = [r for r, _, _ in schedules[(n, 0)][0]]
    r1 = [r for r, _, _ in schedules[(n, 1)][0]]
    m = n
    e2 = 0
    while m % 2 == 0:
        m //= 2
        e2 += 1
    if e2 == 6:
        assert r0 != r1
        assert r1.count(4) == 3 and 8 not in r1[:3]
    else:
        assert r0 == r1


def test_reciprocal_slot_division_is_exact():
    idx = np.arange(0, 1 << 16, dtype=np.float32)
    truth = np.arange(1 << 16, dtype=np.int64)
    for n in range(1, 1025):
        r0 = np.float32(1.0) / np.float32(n)
        for rn in (np.nextafter(r0, np.float32(0)), r0, np.nextafter(r0, np.float32(1))):
            q = np.trunc((idx + np.float32(0.5)) * rn).astype(np.int64)
            assert np.array_equal(q, truth // n), n
    # a plausible mistake: without the half offset the quotient is one short
    # at exact multiples for some n
    bad = 0
    for n in range(1, 1025):
        rn = np.nextafter(np.float32(1.0) / np.float32(n), np.float32(0))
        q = np.trunc(idx * rn).astype(np.int64)
        bad += int(np.any(q != truth // n))
    assert bad > 0
