"""GPU checks of the learned warm start: the engine's BF16-MFMA forward and
backward against the NumPy oracle on the same weights and inputs, and what
the trained network buys the Newton sweep (parity against the FP64 oracle
started from the same warm start)."""
import numpy as np
import pytest

import mlp_oracle as M
from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E

pytestmark = pytest.mark.gpu

N_IN, H, N_OUT = 32, 256, 64


@pytest.fixture(scope="module")
def net():
    s = E.Surrogate(N_IN, H, N_OUT, seed=3)
    return s


def block_errors(g, ref):
    out = []
    o = 0
    for size in (N_IN * H, H, H * H, H, H * N_OUT, N_OUT):
        a, b = g[o:o + size], ref[o:o + size]
        out.append(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))
        o += size
    return out


def test_gradient_matches_oracle(net):
    rng = np.random.default_rng(1)
    w = net.get_weights()
    B = 384
    X = rng.standard_normal((B, N_IN)).astype(np.float32)
    T = rng.standard_normal((B, N_OUT)).astype(np.float32)
    loss, g = net.gradient(X, T)
    rloss, rg = M.loss_and_grad(w, X, T, N_IN, H, N_OUT, bf=True)
    assert abs(loss - rloss) <= 1e-4 * rloss
    errs = block_errors(g.astype(np.float64), rg)
    # FP32 accumulation (engine) vs float64 accumulation (oracle) of the same
    # BF16 operands
    assert max(errs) < 2e-3, errs


def test_gradient_ragged_batch_and_weights_roundtrip(net):
    rng = np.random.default_rng(2)
    w = (rng.standard_normal(net.num_weights) * 0.05).astype(np.float32)
    net.set_weights(w)
    np.testing.assert_array_equal(net.get_weights(), w)
    B = 77   # not a multiple of the 64-row tile
    X = rng.standard_normal((B, N_IN)).astype(np.float32)
    T = rng.standard_normal((B, N_OUT)).astype(np.float32)
    loss, g = net.gradient(X, T)
    rloss, rg = M.loss_and_grad(w, X, T, N_IN, H, N_OUT, bf=True)
    assert abs(loss - rloss) <= 1e-4 * rloss
    assert max(block_errors(g.astype(np.float64), rg)) < 2e-3


def test_predict_matches_oracle_forward(net):
    import torch
    rng = np.random.default_rng(4)
    c = CK.inverter_chain()
    P = c.num_params
    B = 200
    params = rng.standard_normal((B, P))
    slots = np.arange(N_IN, dtype=np.int32)
    net.set_inputs(slots, P)
    im, isd = rng.standard_normal(N_IN), 0.5 + rng.random(N_IN)
    om, osd = rng.standard_normal(N_OUT), 0.5 + rng.random(N_OUT)
    net.set_normalization(im, isd, om, osd)
    dp = torch.as_tensor(params, device="cuda")
    x0 = torch.empty(B, 70, dtype=torch.float64, device="cuda")
    net.predict(dp, 70, x0)
    E.Device.get(0).synchronize()
    got = x0.cpu().numpy()
    Xn = ((params[:, :N_IN] - im.astype(np.float32)) / isd.astype(np.float32)).astype(np.float32)
    Y, _ = M.forward(net.get_weights(), Xn, N_IN, H, N_OUT, bf=True)
    want = Y * osd.astype(np.float32) + om.astype(np.float32)
    np.testing.assert_allclose(got[:, :N_OUT], want, rtol=2e-3, atol=2e-3 * np.abs(want).max())
    assert np.all(got[:, N_OUT:] == 0)


def test_trained_surrogate_saves_newton_iterations(oracle):
    import torch
    c = CK.inverter_chain()
    s = E.train_surrogate(c, list(range(32)), n_train=32768, epochs=12, batch=512, lr=3e-3)
    hist = s.loss_history
    assert hist[-1] < 0.1 * hist[0], hist
    s.attach()
    B = 2048
    base = E.sweep_dcop(c, B, seed=42, config=E.nr_config())
    cfg = E.nr_config(precision=E.Precision.F32_IR)
    cfg.warm_start = A.SPX_WARM_SURROGATE
    warm = E.sweep_dcop(c, B, seed=42, config=cfg)
    assert warm.converged_count == B and np.all(warm.accepted_mask())
    assert np.all(warm.flags & A.SPX_FLAG_WARM_START)
    # fewer Newton iterations than from zero
    assert warm.iterations.mean() < base.iterations.mean() - 0.5, (
        warm.iterations.mean(), base.iterations.mean())
    # parity: the FP64 oracle started from the same prediction
    x0 = torch.empty(B, c.size, dtype=torch.float64, device="cuda")
    s.predict(torch.as_tensor(warm.params, device="cuda"), c.size, x0)
    E.Device.get(0).synchronize()
    x0 = x0.cpu().numpy()
    want = oracle.sweep_dcop(c, warm.params, A.default_nr_config(warm_start=A.SPX_WARM_GIVEN),
                             x0=x0, nthreads=8)
    np.testing.assert_array_equal(warm.status, want["status"])
    same = warm.iterations == want["iterations"]
    assert same.mean() >= 0.999, np.count_nonzero(~same)
    err = np.abs(warm.solutions - want["x"])
    assert np.all(err[same] <= np.maximum(1e-12, 1e-9 * np.abs(want["x"][same])))
    # any other instance (a borderline dx test) at the SPICE tolerance
    assert np.all(err[~same] <= 1e-3 * np.abs(want["x"][~same]) + 1e-6)
    # and the warm-started result agrees with the zero-start one at SPICE tolerance
    assert np.all(np.abs(warm.solutions - base.solutions)
                  <= 2e-3 * np.abs(base.solutions) + 2e-6)


@pytest.mark.parametrize("n_in,h,n_out", [(20, 128, 48), (7, 64, 5), (32, 320, 64)])
def test_predict_other_shapes_match_oracle(n_in, h, n_out):
    # widths the fused inference kernel takes (hidden a multiple of 64 up to
    # 256, padded inputs / outputs) and one it does not (layered GEMMs)
    import torch
    rng = np.random.default_rng(n_in + h)
    s = E.Surrogate(n_in, h, n_out, seed=5)
    w = (rng.standard_normal(s.num_weights) * 0.2).astype(np.float32)
    s.set_weights(w)
    P, B, nx = 40, 333, n_out + 3
    params = rng.standard_normal((B, P))
    slots = rng.choice(P, n_in, replace=False).astype(np.int32)
    s.set_inputs(slots, P)
    im, isd = rng.standard_normal(n_in), 0.5 + rng.random(n_in)
    om, osd = rng.standard_normal(n_out), 0.5 + rng.random(n_out)
    s.set_normalization(im, isd, om, osd)
    x0 = torch.full((B, nx), -1.0, dtype=torch.float64, device="cuda")
    s.predict(torch.as_tensor(params, device="cuda"), nx, x0)
    E.Device.get(0).synchronize()
    got = x0.cpu().numpy()
    Xn = ((params[:, slots] - im.astype(np.float32)) / isd.astype(np.float32)).astype(np.float32)
    Y, _ = M.forward(w, Xn, n_in, h, n_out, bf=True)
    want = Y * osd.astype(np.float32) + om.astype(np.float32)
    np.testing.assert_allclose(got[:, :n_out], want, rtol=2e-3, atol=2e-3 * np.abs(want).max())
    assert np.all(got[:, n_out:] == 0)
