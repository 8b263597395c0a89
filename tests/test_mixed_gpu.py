"""Mixed-size transient batches: one launch (or one hipGraph replay) over
classes of n = 8 .. 256, each on its own kernel shape, against the oracle
and against the single-class path."""
import numpy as np
import pytest
import torch

from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E

pytestmark = pytest.mark.gpu

SIZES = (8, 16, 32, 64, 128, 256)


def _classes(B=24, steps=60):
    return [E.TranClass(CK.ladder(n, seed=100 + n), B, 1e-12, steps, probes=[0, n // 2, n - 2],
                        seed=11 + n) for n in SIZES]


def _cfg():
    return E.nr_config(limiter=E.Limiter.Pnjlim)


def _close(a, b, rtol=1e-9, atol=1e-12):
    return np.all(np.abs(a - b) <= np.maximum(atol, rtol * np.abs(b)))


def test_mixed_eager_matches_oracle(oracle):
    cls = _classes()
    kinds = E.transient_mixed_device(cls, _cfg(), graph=False)
    E.Device.get(0).synchronize()
    assert kinds == [E.KernelKind.Dense] * 4 + [E.KernelKind.Sparse256] * 2
    for c in cls:
        got = c.host_result()
        want = oracle.transient(c.circuit, got.params,
                                A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM), c.tstep,
                                c.steps, probes=c.probes, nthreads=16)
        np.testing.assert_array_equal(got.status, want["status"])
        np.testing.assert_array_equal(got.iterations, want["iterations"])
        assert _close(got.waveforms, want["waveforms"]) and _close(got.solutions, want["x"])
        assert np.all(got.accepted_mask())


def test_graph_replays_equal_eager():
    cls = _classes(B=32, steps=40)
    E.transient_mixed_device(cls, _cfg(), graph=False)
    E.Device.get(0).synchronize()
    ref = [{k: v.clone() for k, v in c.buffers.items()} for c in cls]
    for rep in range(3):
        for c in cls:  # the replay must rewrite every output
            for v in c.buffers.values():
                v.fill_(-7)
        torch.cuda.synchronize()  # torch's stream is not the library's
        E.transient_mixed_device(cls, _cfg(), graph=True)
        E.Device.get(0).synchronize()
        torch.cuda.synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()  # eager work between replays
        for c, r in zip(cls, ref):
            for k in r:
                assert torch.equal(c.buffers[k], r[k]), (rep, c.circuit.name, k)


@pytest.mark.parametrize("kind", [E.KernelKind.Dense, E.KernelKind.Sparse64,
                                  E.KernelKind.Sparse256, E.KernelKind.Lanes])
def test_every_kernel_kind_matches_oracle(oracle, kind):
    c = E.TranClass(CK.ladder(48, seed=4), 100, 1e-12, 50, probes=[0, 10, 46], seed=3)
    got_kinds = E.transient_mixed_device([c], _cfg(), graph=False, kernels=[kind])
    E.Device.get(0).synchronize()
    assert got_kinds == [kind]
    got = c.host_result()
    want = oracle.transient(c.circuit, got.params, A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM),
                            c.tstep, c.steps, probes=c.probes, nthreads=16)
    np.testing.assert_array_equal(got.iterations, want["iterations"])
    assert _close(got.waveforms, want["waveforms"])


def test_autotune_picks_a_valid_kernel_and_keeps_results():
    cls = _classes(B=64, steps=30)
    E.transient_mixed_device(cls, _cfg(), graph=False)
    E.Device.get(0).synchronize()
    ref = [c.buffers["waveforms"].clone() for c in cls]
    kinds = E.transient_mixed_device(cls, _cfg(), graph=False, autotune=True)
    E.Device.get(0).synchronize()
    for n, k in zip(SIZES, kinds):
        assert k != E.KernelKind.Auto and (n <= 64 or k != E.KernelKind.Dense)
    for c, r in zip(cls, ref):
        w = c.buffers["waveforms"]
        assert torch.allclose(w, r, rtol=1e-9, atol=1e-12)


def test_mixed_ir_matches_oracle(oracle):
    # FP32+IR on the register-LU classes (n <= 64); the sparse classes factor
    # in FP64.  Both against the FP64 oracle; graph replay equals eager.
    cls = _classes(B=24, steps=60)
    cfg = E.nr_config(limiter=E.Limiter.Pnjlim, precision=E.Precision.F32_IR)
    kinds = E.transient_mixed_device(cls, cfg, graph=False)
    E.Device.get(0).synchronize()
    assert kinds == [E.KernelKind.Dense] * 4 + [E.KernelKind.Sparse256] * 2
    ref = [{k: v.clone() for k, v in c.buffers.items()} for c in cls]
    for c in cls:
        got = c.host_result()
        want = oracle.transient(c.circuit, got.params,
                                A.default_nr_config(limiter=A.SPX_LIMIT_PNJLIM), c.tstep,
                                c.steps, probes=c.probes, nthreads=16)
        np.testing.assert_array_equal(got.status, want["status"])
        # every instance is compared: those whose Newton counts equal the
        # oracle's to 1e-9, any other (a borderline dx test on an FP32+IR
        # solve) at the SPICE tolerance, reltol 1e-3 and vntol 1 uV
        same = got.iterations == want["iterations"]
        assert same.mean() >= 0.999, (c.circuit.name, np.count_nonzero(~same))
        assert _close(got.waveforms[same], want["waveforms"][same])
        assert np.all(np.abs(got.waveforms[~same] - want["waveforms"][~same])
                      <= 1e-3 * np.abs(want["waveforms"][~same]) + 1e-6)
        assert np.all(got.accepted_mask())
    for rep in range(2):
        for c in cls:
            for v in c.buffers.values():
                v.fill_(-7)
        torch.cuda.synchronize()
        E.transient_mixed_device(cls, cfg, graph=True)
        E.Device.get(0).synchronize()
        _ = torch.ones(1000, device="cuda").sum().item()
        for c, r in zip(cls, ref):
            for k in r:
                assert torch.equal(c.buffers[k], r[k]), (rep, c.circuit.name, k)
