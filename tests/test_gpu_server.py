"""Continuous-batching server: each request's results are bit-identical alone and
in a mix of 2-D and 3-D pairs of different sizes and algorithms, a budget that
holds only a few pairs at a time still drains every request, and a pair larger
than the whole budget is rejected at submit."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from multi_quickpiv_amd import PIV, setPIVParameters, synth
from multi_quickpiv_amd._lib import MQPIV_ERR_OOM, MqpivError
from multi_quickpiv_amd.piv import Server

pytestmark = pytest.mark.gpu


def _host_pair(shape, d, seed):
    n = synth.n_particles(shape, 0.05 if len(shape) == 2 else 0.01)
    a, b = synth.particle_pair(shape, n, 1.2, synth.uniform(d), seed, device="cpu", dtype=torch.float32)
    return a.cpu().numpy(), b.cpu().numpy()


def _specs():
    out = []
    rows = [
        ((256, 224), "nsqecc", 32, 16, 16, False),
        ((300, 200), "zncc", 24, 8, 12, False),
        ((256, 256), "fft", 16, 8, 8, False),
        ((384, 320), "nsqecc", 32, 16, 16, True),
        ((64, 64, 48), "nsqecc", 16, 6, 8, False),
        ((320, 256), "nsqecc", 64, 32, 32, False),
        ((48, 56, 40), "zncc", 16, 4, 8, False),
        ((512, 512), "zncc", 32, 16, 16, False),
    ]
    for i, (shape, alg, isz, sm, st, masked) in enumerate(rows):
        d = (1.5 + 0.3 * i, -2.0 + 0.5 * i, 0.7)[: len(shape)]
        a, b = _host_pair(shape, d, 300 + i)
        p = setPIVParameters(corr_alg=alg, interSize=isz, searchMargin=sm, step=st, computeSN=True)
        m = None
        if masked:
            m = synth.ellipse_mask(shape, n=5, cover=0.3, seed=i).numpy()
        out.append((a, b, p, m))
    return out


def _run(server, specs, order):
    tickets = {}
    for i in order:
        a, b, p, m = specs[i]
        tickets[i] = server.submit(a, b, p, mask=m, mode="auto", return_peaks=True)
    return {i: server.wait(t, timeout=60.0) for i, t in tickets.items()}


def _same(r1, r2):
    for x, y in zip(r1, r2):
        if x is None:
            assert y is None
        else:
            assert np.array_equal(x, y, equal_nan=True)


def test_server_alone_vs_mix_is_bit_identical():
    specs = _specs()
    sv = Server(hbm_budget_bytes=1 << 30, n_slots=3, max_batch=4)
    try:
        alone = {}
        for i in range(len(specs)):
            alone.update(_run(sv, specs, [i]))
        rng = np.random.default_rng(5)
        for rep in range(2):
            order = list(rng.permutation(len(specs))) * 2  # duplicates in flight together
            tickets = [(i, sv.submit(specs[i][0], specs[i][1], specs[i][2], mask=specs[i][3], mode="auto",
                                     return_peaks=True)) for i in order]
            for i, t in tickets:
                res = sv.wait(t, timeout=60.0)
                assert res is not None, (rep, i)
                print(rep, i, specs[i][0].shape, flush=True)
                _same(res, alone[i])
        # and the lone server results equal a direct call
        for i, (a, b, p, m) in enumerate(specs):
            args = (m, p) if m is not None else (p,)
            _same(alone[i], PIV(a, b, *args, return_peaks=True, mode="auto"))
        st = sv.stats()
        assert st["completed"] == len(specs) * 5 and st["pending"] == 0
        assert 0 < st["p50_ms"] <= st["p99_ms"] <= st["max_ms"]
    finally:
        sv.close()


def test_server_small_budget_drains_and_rejects_oversize():
    specs = _specs()
    a, b, p, _ = specs[7]  # 512x512 f32 ZNCC: 2 MB of inputs + 4.2 MB of integral arrays
    sv = Server(hbm_budget_bytes=16 << 20, n_slots=2, max_batch=8)
    try:
        ref = PIV(a, b, p, mode="auto")
        tickets = [sv.submit(a, b, p) for _ in range(12)]
        for t in tickets:
            res = sv.wait(t, timeout=60.0)
            assert res is not None
            _same(res, ref)
        big = np.zeros((1024, 1024), dtype=np.float64)  # 16 MiB of inputs + 8 MiB of arrays
        with pytest.raises(MqpivError) as ei:
            sv.submit(big, big, setPIVParameters(interSize=32, searchMargin=8, step=32))
        assert ei.value.code == MQPIV_ERR_OOM
    finally:
        sv.close()
