"""GPU parity of controlled sources (VCVS, VCCS, CCCS, CCVS;
controlled.rs:58-375) against the oracle on identical seeded parameters,
through the C ABI: DC sweeps of the reference's golden netlists and of
finite-gain op-amp amplifiers (checked against their closed forms too), and
transients of a ladder that uses all four kinds, with every integration
method, adaptive stepping, every kernel kind and the FP32+IR form.  Bar:
waveforms and solutions within 1e-9 relative (1e-12 absolute); status,
flags and iteration counts equal; FP32+IR normwise (its refinement's stopping
test) and componentwise on the large nodes."""
import json
import os

import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E
from spicier_amd.netlist import parse_netlist

from test_transient_methods_gpu import _check_adaptive, _check_fixed, close

pytestmark = pytest.mark.gpu

M = E.IntegrationMethod
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "dc_controlled_sources.json")


def _probes(c):
    return [c.node_names[k] - 1 for k in ("a1", "b1", "h1", "q1")] + [c.size - 1]


def _check_dc(got, want):
    np.testing.assert_array_equal(got.status, want["status"])
    np.testing.assert_array_equal(got.iterations, want["iterations"])
    np.testing.assert_array_equal(got.flags, want["flags"])
    assert np.all(close(got.solutions, want["x"]))


def test_golden_netlists_on_gpu(oracle):
    with open(GOLDEN) as f:
        cases = json.load(f)["circuits"]
    for case in cases:
        text = case["netlist"]
        if text.split("\n")[0].split()[0].upper() == "V1":
            text = case["name"] + "\n" + text
        c = parse_netlist(text).circuit
        got = E.sweep_dcop(c, 4, seed=1)
        _check_dc(got, oracle.sweep_dcop(c, got.params))
        for var, want in case["results"].items():
            col = c.node_names[var[2:-1]] - 1
            np.testing.assert_allclose(got.solutions[:, col], want,
                                       atol=max(case["voltage_tol"], 1e-9))


def test_opamp_monte_carlo_closed_form(oracle):
    # inverting amplifier with Rin, Rf and the open-loop gain A varied per instance
    c = CK.Circuit("inv_mc")
    c.vsource("1", "0", 1.0)
    c.resistor("1", "2", c.gauss_rel(1e3, 0.05, name="RIN"))
    c.resistor("2", "3", c.gauss_rel(10e3, 0.05, name="RF"))
    c.vcvs("3", "0", "0", "2", c.loguniform(1e3, 1e6, name="A"))
    got = E.sweep_dcop(c, 4096, seed=11)
    _check_dc(got, oracle.sweep_dcop(c, got.params, nthreads=16))
    p = {n: got.params[:, c.slot_names.index(n)] for n in ("RIN", "RF", "A")}
    want = -(p["RF"] / p["RIN"]) / (1 + (1 + p["RF"] / p["RIN"]) / p["A"])
    np.testing.assert_allclose(got.solutions[:, c.node_names["3"] - 1], want, rtol=1e-12)


@pytest.mark.parametrize("method", [M.BackwardEuler, M.Trapezoidal, M.TrBdf2])
@pytest.mark.parametrize("sections", [1, 7, 16])
def test_controlled_ladder_methods_match_oracle(oracle, sections, method):
    # 10, 58 (register LU) and 130 (shared-symbolic sparse) unknowns
    c = CK.controlled_ladder(sections, seed=sections)
    B = 32 if sections <= 7 else 12
    probes = _probes(c)
    got = E.transient_sweep(c, B, 1e-12, 150, probes=probes, seed=7, method=method)
    want = oracle.transient_ex(c, got.params, A.default_nr_config(), 1e-12, 150, probes=probes,
                               method=int(method), nthreads=16)
    _check_fixed(got, want)
    assert np.all(got.accepted_mask())


@pytest.mark.parametrize("kind", [E.KernelKind.Dense, E.KernelKind.Sparse64,
                                  E.KernelKind.Sparse256, E.KernelKind.Lanes])
def test_every_kernel_kind_with_controlled_sources(oracle, kind):
    ad = E.AdaptiveTransientParams(tstop=2e-10, h_init=1e-13, h_min=1e-16, h_max=2e-11,
                                   reltol=1e-3, abstol=1e-6, max_points=2000)
    c0, c1 = CK.controlled_ladder(5, seed=4), CK.controlled_ladder(5, seed=6)
    cfg = A.default_nr_config()
    cls = [E.TranClass(c0, 64, 1e-12, 60, probes=_probes(c0), seed=3, method=M.BackwardEuler),
           E.TranClass(c1, 64, 0.0, 0, probes=_probes(c1), seed=9, adaptive=ad)]
    got_kinds = E.transient_mixed_device(cls, E.nr_config(), graph=False, kernels=[kind, kind])
    E.Device.get(0).synchronize()
    assert got_kinds == [kind, kind]
    g0, g1 = cls[0].host_result(), cls[1].host_result()
    _check_fixed(g0, oracle.transient_ex(c0, g0.params, cfg, 1e-12, 60, probes=cls[0].probes,
                                         method=int(M.BackwardEuler), nthreads=16))
    _check_adaptive(g1, oracle.transient_ex(c1, g1.params, cfg, probes=cls[1].probes,
                                            adaptive=ad, nthreads=16))


def test_fp32_ir_with_controlled_sources(oracle):
    c = CK.controlled_ladder(6, seed=2)
    cfg = E.nr_config(precision=E.Precision.F32_IR)
    probes = _probes(c)
    got = E.transient_sweep(c, 48, 1e-12, 100, probes=probes, seed=7, config=cfg,
                            method=M.Trapezoidal)
    want = oracle.transient_ex(c, got.params, A.default_nr_config(), 1e-12, 100, probes=probes,
                               method=int(M.Trapezoidal), nthreads=16)
    np.testing.assert_array_equal(got.status, want["status"])
    # FP32 factors refined in FP64 stop on dsgesv's normwise test, so the bar
    # is normwise: 1e-9 of the instance's largest probed value (the VCCS
    # output nodes sit near -3 mV beside 1.8 V nodes and carry ~4e-12 V)
    w = want["waveforms"]
    scale = np.abs(w).max(axis=(1, 2), keepdims=True)
    err = np.abs(got.waveforms - w)
    assert np.all(err <= 1e-9 * scale), (err / scale).max()
    assert np.all(close(got.waveforms[:, :, :3], w[:, :, :3]))  # the large nodes componentwise


def test_compile_rejects_bad_control():
    # a CCCS whose controlling element is itself or a resistor is rejected
    c = CK.Circuit("bad")
    c.vsource("1", "0", 1.0)
    r = c.resistor("1", "0", 1e3)
    c.cccs("2", "0", 0, 1.0)
    c.resistor("2", "0", 1e3)
    c.ctrl[2] = r
    with pytest.raises(E.InvalidArgument):
        E.sweep_dcop(c, 2)
    c.ctrl[2] = 2
    with pytest.raises(E.InvalidArgument):
        E.sweep_dcop(c, 2)
