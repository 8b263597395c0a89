"""Pins the oracle's controlled sources (Vcvs, Vccs, Cccs, Ccvs stamps,
spicier-devices/src/controlled.rs:58-375) and the netlist reader's E/G/F/H
(parser/elements.rs:716-830) with the reference's ngspice-43 golden DC
points (tests/golden/dc_controlled_sources.json, written from the
reference's test data by tests/golden/make_controlled_sources.py), its
ngspice_validation tests (ngspice_validation.rs:927-1105) and parser tests
(parser/mod.rs:903-960), and with closed forms: the finite-gain op-amp
amplifiers, and a VCCS charging a capacitor, whose Backward-Euler and
trapezoidal ramps are exact."""
import json
import os

import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E
from spicier_amd.netlist import parse_netlist

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "dc_controlled_sources.json")


def _dcop(oracle, c):
    out = oracle.sweep_dcop(c, CK.nominal_values(c).reshape(1, -1))
    assert out["status"][0] == A.SPX_CONVERGED
    return out["x"][0]


def _v(c, x, node):
    return x[c.node_names[str(node)] - 1] if str(node) != "0" else 0.0


with open(GOLDEN) as f:
    _CASES = json.load(f)["circuits"]


@pytest.mark.parametrize("case", _CASES, ids=[c["name"] for c in _CASES])
def test_ngspice_golden_dc_points(oracle, case):
    text = case["netlist"]
    if text.split("\n")[0].split()[0].upper() == "V1":  # two fixtures start without a title
        text = case["name"] + "\n" + text
    c = parse_netlist(text).circuit
    x = _dcop(oracle, c)
    for var, want in case["results"].items():
        node = var[2:-1]
        assert _v(c, x, node) == pytest.approx(want, abs=max(case["voltage_tol"], 1e-9)), var


@pytest.mark.parametrize("text,node,want", [
    ("VCVS Gain Test\nV1 1 0 DC 2\nR1 1 0 1k\nR2 2 0 1k\nE1 2 0 1 0 5\n.end\n", 2, 10.0),
    ("VCCS\nV1 1 0 DC 2\nR1 1 0 1k\nR2 2 0 1k\nG1 2 0 1 0 1m\n.end\n", 2, 2.0),
    ("CCCS\nV1 1 0 DC 10\nR1 1 2 1k\nVsense 2 0 DC 0\nR2 3 0 1k\nF1 3 0 Vsense 5\n.end\n",
     3, -50.0),
    ("CCVS\nV1 1 0 DC 5\nR1 1 2 1k\nVsense 2 0 DC 0\nR2 3 0 10k\nH1 3 0 Vsense 2k\n.end\n",
     3, 10.0),
])
def test_reference_validation_circuits(oracle, text, node, want):
    # ngspice_validation.rs:927-1105
    c = parse_netlist(text).circuit
    assert _v(c, _dcop(oracle, c), node) == pytest.approx(want, rel=1e-12)


@pytest.mark.parametrize("text,ndev,ncur", [
    ("VCVS Test\nV1 1 0 10\nR1 1 2 1k\nR2 3 0 1k\nE1 3 0 1 2 2.0\n.end\n", 4, 2),
    ("VCCS Test\nV1 1 0 10\nR1 2 0 1k\nG1 2 0 1 0 0.001\n.end\n", 3, 1),
    ("CCCS Test\nV1 1 0 10\nR1 1 0 1k\nR2 2 0 1k\nF1 2 0 V1 3.0\n.end\n", 4, 1),
    ("CCVS Test\nV1 1 0 10\nR1 1 0 1k\nR2 2 0 1k\nH1 2 0 V1 100.0\n.end\n", 4, 2),
])
def test_reference_parser_counts(text, ndev, ncur):
    # parser/mod.rs:903-960: devices and current variables
    c = parse_netlist(text).circuit
    assert len(c.elements) == ndev and c.size - c.num_nodes == ncur


@pytest.mark.parametrize("gain", [10.0, 1e3, 1e5, 1e9])
def test_finite_gain_opamp_closed_forms(oracle, gain):
    # inverting: v3 = -(Rf/Rin) vin / (1 + (1 + Rf/Rin)/A); non-inverting:
    # v3 = (1 + Rf/R1) vin / (1 + (1 + Rf/R1)/A)
    rin, rf, vin = 1e3, 10e3, 1.0
    c = CK.Circuit("inv")
    c.vsource("1", "0", vin)
    c.resistor("1", "2", rin)
    c.resistor("2", "3", rf)
    c.vcvs("3", "0", "0", "2", gain)
    x = _dcop(oracle, c)
    want = -(rf / rin) * vin / (1 + (1 + rf / rin) / gain)
    assert _v(c, x, 3) == pytest.approx(want, rel=1e-12)
    c = CK.Circuit("noninv")
    c.vsource("1", "0", vin)
    c.resistor("2", "0", rin)
    c.resistor("2", "3", rf)
    c.vcvs("3", "0", "1", "2", gain)
    x = _dcop(oracle, c)
    k = 1 + rf / rin
    assert _v(c, x, 3) == pytest.approx(k * vin / (1 + k / gain), rel=1e-12)


def test_ccvs_controlled_by_vcvs_branch_and_cccs_by_ccvs(oracle):
    # chains of current control: H1 reads E1's branch current, F1 reads H1's
    c = CK.Circuit("chain")
    c.vsource("1", "0", 2.0)
    e1 = c.vcvs("2", "0", "1", "0", 3.0)       # v2 = 6 V
    c.resistor("2", "0", 1e3)                   # E1 sources 6 mA: branch current -6 mA
    h1 = c.ccvs("3", "0", e1, 500.0)            # v3 = 500 * (-6 mA) = -3 V
    c.resistor("3", "0", 1e3)                   # H1 branch current +3 mA
    c.cccs("4", "0", h1, 2.0)                   # KCL at 4: v4/1k + 2 * 3 mA = 0
    c.resistor("4", "0", 1e3)
    x = _dcop(oracle, c)
    np.testing.assert_allclose([_v(c, x, k) for k in (1, 2, 3, 4)], [2.0, 6.0, -3.0, -6.0],
                               rtol=1e-12)


@pytest.mark.parametrize("method", [A.SPX_TRAN_BE, A.SPX_TRAN_TRAP])
def test_vccs_switched_on_ramp(oracle, method):
    # the control steps from 0 to vin at t = 0+: C charges at gm vin / C
    gm, vin, cap, h = 1e-3, 0.5, 1e-9, 1e-9
    c = CK.Circuit("ramp_on")
    c.vpulse("in", "0", 0.0, vin, 0.0, 0.0, 0.0, 1.0, 0.0)
    c.resistor("in", "0", 1e3)
    c.vccs("out", "0", "in", "0", gm)
    c.capacitor("out", "0", cap)
    c.resistor("out", "0", 1e12)  # a DC path; its leak is 4e-11 of the ramp current
    p = CK.nominal_values(c).reshape(1, -1)
    cfg = A.default_nr_config()
    out = oracle.transient_ex(c, p, cfg, h, 40, probes=[c.node_names["out"] - 1], method=method)
    w = out["waveforms"][0, :, 0]
    k = np.arange(41)
    # BE: C (v_k - v_k-1)/h = gm vin; trapezoidal averages i_0 = 0 into the first step
    ramp = gm * vin * h / cap
    want = k * ramp if method == A.SPX_TRAN_BE else np.maximum(k - 0.5, 0) * ramp
    np.testing.assert_allclose(w, want, rtol=1e-9, atol=1e-9 * ramp)


def test_netlist_rejects_unknown_or_later_control():
    with pytest.raises(E.InvalidArgument):
        parse_netlist("t\nR1 1 0 1k\nF1 1 0 Vx 2\n")
    with pytest.raises(E.InvalidArgument):  # controlled by a later source
        parse_netlist("t\nR1 1 0 1k\nH1 1 0 V1 2\nV1 1 0 1\n")
    with pytest.raises(E.InvalidArgument):  # a resistor has no branch current
        parse_netlist("t\nR1 1 0 1k\nF1 1 0 R1 2\n")
    with pytest.raises(ValueError):
        CK.Circuit("x").cccs("1", "0", 0, 1.0)
