"""GPU parity of the fragment batch (sdft_hip_batch_*) against the CPU oracle.

Fragments are cut from the Si1728 cell exactly as for the 343-fragment
workload (seed 3) and filtered together; each fragment's accumulator must
match the oracle's filterChebyPolyHerm on that fragment alone (FP64: 1e-10
relative, FP32: 1e-4 relative, as for single systems), the graph replay must
be bit-identical to direct launches, and the batch must agree with the
single-system block filter.
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


@pytest.fixture(scope="module")
def frags():
    """Two fragments of each grid size (24, 32, 36, 48) with few orbitals
    (odd counts included, to exercise the pair padding)."""
    from sdft_amd import stodft, systems
    specs = systems.fragment_specs()
    picked = []
    for n in systems.FRAGMENT_GRIDS:
        picked += [s for s in specs if s.n == n][:2]
    out = []
    for i, s in enumerate(picked):
        sysm = systems.build_fragment(s)
        emin, emax = stodft.spectral_bounds(sysm)
        nstate = 3 + (i % 4)  # 3..6
        chi = systems.noise_orbitals(sysm, nstate, seed=11 + i)
        out.append((sysm, chi, emin, emax))
    return out


def _coeffs(frags, L):
    co = np.stack([O.cheby_coeffs(L, emin, emax, 0.1 + 0.02 * i, 100.0, 2)
                   for i, (_, _, emin, emax) in enumerate(frags)])
    scale = np.array([2.0 / (emax - emin) for (_, _, emin, emax) in frags])
    emean = np.array([0.5 * (emin + emax) for (_, _, emin, emax) in frags])
    return co, scale, emean


def _make_batch(hip, frags, prec):
    ctxs, blocks = [], []
    for sysm, chi, _, _ in frags:
        c = hip.Context(sysm)
        b = hip.Block(c, chi.shape[0], prec)
        b.set(chi)
        ctxs.append(c)
        blocks.append(b)
    return ctxs, blocks, hip.Batch(blocks)


@pytest.mark.parametrize("L", [2, 3, 17])
def test_batch_fp64_matches_oracle(hip, frags, L):
    say("case: batch fp64 8 fragments L=%d" % L)
    co, scale, emean = _coeffs(frags, L)
    ctxs, blocks, bat = _make_batch(hip, frags, hip.PREC_FP64)
    bat.filter(co, scale, emean, use_graph=True)
    for f, (sysm, chi, _, _) in enumerate(frags):
        acc_r, tl_r = O.filter_cheby(chi, co[f], sysm, scale[f], emean[f])
        acc_g = blocks[f].get(0)
        tl_g = blocks[f].get(1)
        assert rel(acc_g, acc_r) < 1e-10, (f, sysm.name, rel(acc_g, acc_r))
        assert rel(tl_g, tl_r) < 1e-10, (f, sysm.name, rel(tl_g, tl_r))
    bat.close()


def test_batch_graph_equals_direct_and_block(hip, frags):
    say("case: batch fp64 graph vs direct vs block filter, L=24")
    L = 24
    co, scale, emean = _coeffs(frags, L)
    ctxs, blocks, bat = _make_batch(hip, frags, hip.PREC_FP64)
    bat.filter(co, scale, emean, use_graph=True)
    g = [b.get(0) for b in blocks]
    bat.filter(co, scale, emean, use_graph=False)
    d = [b.get(0) for b in blocks]
    # the same kernels in the same order: identical bits
    for f in range(len(frags)):
        assert np.array_equal(g[f], d[f]), f
    # the single-system block filter (its KB split may differ: sums in
    # another order)
    for f, b in enumerate(blocks):
        b.filter(co[f], scale[f], emean[f])
        assert rel(b.get(0), g[f]) < 1e-12, f
    # a second batch call re-filters chi and gives the same result again
    bat.filter(co, scale, emean, use_graph=True)
    for f in range(len(frags)):
        assert np.array_equal(blocks[f].get(0), g[f]), f
    # the whole filter as one graph launch: identical bits, also after a
    # filter of another length (recaptured) and a configuration change
    bat.filter(co, scale, emean, use_graph=2)
    for f in range(len(frags)):
        assert np.array_equal(blocks[f].get(0), g[f]), f
    co16, sc16, em16 = _coeffs(frags, 16)
    bat.filter(co16, sc16, em16, use_graph=2)
    h16 = [b.get(0) for b in blocks]
    bat.filter(co16, sc16, em16, use_graph=False)
    for f in range(len(frags)):
        assert np.array_equal(blocks[f].get(0), h16[f]), f
    bat.filter(co, scale, emean, use_graph=2)
    for f in range(len(frags)):
        assert np.array_equal(blocks[f].get(0), g[f]), f
    bat.config(2, 0, 1, 1)
    bat.filter(co, scale, emean, use_graph=2)
    h2 = [b.get(0) for b in blocks]
    bat.filter(co, scale, emean, use_graph=False)
    for f in range(len(frags)):
        assert np.array_equal(blocks[f].get(0), h2[f]), f
        assert rel(h2[f], g[f]) < 1e-12, f
    with pytest.raises(RuntimeError):
        bat.filter(co, scale, emean, use_graph=3)
    bat.close()


def test_batch_fp32_matches_oracle(hip, frags):
    say("case: batch fp32 8 fragments L=40")
    L = 40
    co, scale, emean = _coeffs(frags, L)
    ctxs, blocks, bat = _make_batch(hip, frags, hip.PREC_FP32)
    bat.filter(co, scale, emean, use_graph=True)
    for f, (sysm, chi, _, _) in enumerate(frags):
        acc_r, _ = O.filter_cheby(chi, co[f], sysm, scale[f], emean[f])
        assert rel(blocks[f].get(0), acc_r) < 1e-4, (f, sysm.name)
    bat.close()


def test_batch_autotune_and_configs(hip, frags):
    say("case: batch autotune + every KB split / overlap / plane tile / fused-expand "
        "configuration, L=9")
    L = 9
    co, scale, emean = _coeffs(frags, L)
    ctxs, blocks, bat = _make_batch(hip, frags, hip.PREC_FP64)
    (split, ov, pv, fu), ms = bat.autotune(steps=4)
    assert split in (1, 2, 4, 8) and ov in (0, 1) and pv in (0, 1) and fu in (0, 1)
    assert len(ms) == 32 and np.all(ms > 0)
    refs = [O.filter_cheby(chi, co[f], sysm, scale[f], emean[f])[0]
            for f, (sysm, chi, _, _) in enumerate(frags)]
    for sp in (1, 2, 4, 8):
        for o in (0, 1):
            for t in (0, 1):
                for e in (0, 1):
                    bat.config(sp, o, t, e)
                    bat.filter(co, scale, emean, use_graph=True)
                    for f in range(len(frags)):
                        assert rel(blocks[f].get(0), refs[f]) < 1e-10, (sp, o, t, e, f)
    with pytest.raises(hip.SdftError):
        bat.config(3, 0, 0, 0)
    with pytest.raises(hip.SdftError):
        bat.config(1, 0, 0, 2)
    bat.close()


def test_batch_rejects_bad_input(hip, frags):
    say("case: batch argument checks")
    sysm, chi, _, _ = frags[0]
    c = hip.Context(sysm)
    b64 = hip.Block(c, 2, hip.PREC_FP64)
    b32 = hip.Block(c, 2, hip.PREC_FP32)
    with pytest.raises(hip.SdftError):
        hip.Batch([b64, b32])  # mixed precision
    with pytest.raises(hip.SdftError):
        hip.Batch([b64, b64])  # same block twice
    b64.set(chi[:2])
    bat = hip.Batch([b64])
    with pytest.raises(hip.SdftError):
        bat.filter(np.ones((1, 1)), np.ones(1), np.zeros(1))  # L < 2
    bat.close()


def test_batch_refuses_projector_change_after_create(hip, frags):
    """The batch sizes its KB buffers from each context's projector count at
    sdft_hip_batch_create; a context whose projectors change afterwards (more,
    or from none to some) makes the next batch filter return an error instead
    of launching with the old sizes, and re-creating the batch works."""
    say("case: batch projector change after create")
    sysm, chi, emin, emax = frags[0]
    L = 4
    co, sc, em = _coeffs([frags[0]], L)
    c = hip.Context(sysm, with_nl=False)
    b = hip.Block(c, 2, hip.PREC_FP64)
    b.set(chi[:2])
    bat = hip.Batch([b])
    bat.filter(co, sc, em)
    c.set_projectors(sysm.proj, sysm.proj_energy)  # 0 -> nproj projectors
    with pytest.raises(hip.SdftError, match="projectors changed"):
        bat.filter(co, sc, em)
    bat.close()
    bat = hip.Batch([b])
    bat.filter(co, sc, em)
    acc = b.get(0)
    ref, _ = O.filter_cheby(chi[:2], co[0], sysm, sc[0], em[0])
    assert rel(acc, ref) < 1e-10
    # more projectors than at create time
    c.set_projectors(np.concatenate([sysm.proj, sysm.proj]),
                     np.concatenate([sysm.proj_energy, sysm.proj_energy]))
    with pytest.raises(hip.SdftError, match="projectors changed"):
        bat.filter(co, sc, em)
    bat.close()
    b.close()
    c.close()


def test_batch_mixed_local_only_and_real_space_fragments(hip, frags):
    """A batch whose fragments mix the three KB forms: G-space table, no
    projectors (local potential only: the contract takes no NL force) and
    real-space projectors (through their device-built table, nl_form 1: the
    batch refuses the per-atom form); each against the oracle."""
    from sdft_amd import systems
    say("case: batch with local-only and real-space-KB fragments, L=9")
    L = 9
    sub = frags[:4]
    co, scale, emean = _coeffs(sub, L)
    ctxs, blocks = [], []
    kinds = ["gspace", "local", "real", "gspace"]
    rs = {}
    for i, (sysm, chi, _, _) in enumerate(sub):
        if kinds[i] == "local":
            c = hip.Context(sysm, with_nl=False)
        elif kinds[i] == "real":
            c = hip.Context(sysm, with_nl=False)
            rs[i] = systems.real_space_projectors(sysm, 0.45 * sysm.L)
            c.set_projectors_real(rs[i])
            b0 = hip.Block(c, chi.shape[0], hip.PREC_FP64)
            with pytest.raises(hip.SdftError):
                hip.Batch([b0])
            b0.close()
            c.set_nl_form(1)
        else:
            c = hip.Context(sysm)
        b = hip.Block(c, chi.shape[0], hip.PREC_FP64)
        b.set(chi)
        ctxs.append(c)
        blocks.append(b)
    bat = hip.Batch(blocks)
    bat.filter(co, scale, emean)
    for i, (sysm, chi, _, _) in enumerate(sub):
        if kinds[i] == "local":
            ref, _ = O.filter_cheby(chi, co[i], sysm, scale[i], emean[i], with_nl=False)
        elif kinds[i] == "real":
            ref, _ = O.filter_cheby(chi, co[i], sysm, scale[i], emean[i], nl_real=rs[i])
        else:
            ref, _ = O.filter_cheby(chi, co[i], sysm, scale[i], emean[i])
        assert rel(blocks[i].get(0), ref) < 1e-10, (i, kinds[i])
    bat.close()


def test_batch_long_filter_both_forms(hip, frags):
    """L = 128 on the 8 fragments with the autotuned configuration, in FP64
    (1e-10 relative to the oracle per fragment) and FP32 (1e-4): the graph
    replays 63 step pairs, so a step-parity or counter slip would show."""
    say("case: batch 8 fragments L=128, FP64 and FP32, autotuned")
    L = 128
    co, scale, emean = _coeffs(frags, L)
    refs = [O.filter_cheby(chi, co[f], sysm, scale[f], emean[f])[0]
            for f, (sysm, chi, _, _) in enumerate(frags)]
    for prec, tol in ((hip.PREC_FP64, 1e-10), (hip.PREC_FP32, 1e-4)):
        ctxs, blocks, bat = _make_batch(hip, frags, prec)
        cfg, _ = bat.autotune(steps=4)
        for b, (_, chi, _, _) in zip(blocks, frags):
            b.set(chi)
        bat.filter(co, scale, emean, use_graph=True)
        worst = max(rel(blocks[f].get(0), refs[f]) for f in range(len(frags)))
        say("  prec %d config %s: worst fragment rel %.3g" % (prec, cfg, worst))
        assert worst < tol, (prec, cfg, worst)
        bat.close()
        for b in blocks:
            b.close()
