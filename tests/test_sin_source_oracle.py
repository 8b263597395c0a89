"""SIN sources (waveforms.rs:34-49, 185-201) in the oracle's transient,
pinned by the reference's waveform test (waveforms.rs:276-292) and the
closed form of a resistive divider, which follows the source exactly; and
the netlist front end's SIN syntax (parser/waveforms.rs:44-72)."""
import math

import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd.netlist import parse_netlist
from spicier_amd.engine import InvalidArgument


def _divider(vo, va, freq, td=0.0, theta=0.0, phase=0.0, r1=1e3, r2=3e3):
    c = CK.Circuit("sin_divider")
    c.vsin("in", "0", vo, va, freq, td, theta, phase)
    c.resistor("in", "out", r1)
    c.resistor("out", "0", r2)
    return c


def _run(oracle, c, h, steps, method=A.SPX_TRAN_BE):
    p = CK.nominal_values(c).reshape(1, -1)
    out = oracle.transient_ex(c, p, A.default_nr_config(), h, steps,
                              probes=[c.node_names["in"] - 1, c.node_names["out"] - 1],
                              method=method)
    assert out["status"][0] == A.SPX_CONVERGED
    return out["waveforms"][0]


def test_reference_sin_values(oracle):
    # waveforms.rs:276-292: SIN(0 1 1k) at 0, T/4, T/2, 3T/4
    w = _run(oracle, _divider(0.0, 1.0, 1e3), 0.25e-3, 3)
    np.testing.assert_allclose(w[:, 0], [0.0, 1.0, 0.0, -1.0], atol=1e-10)
    np.testing.assert_allclose(w[:, 1], 0.75 * w[:, 0], rtol=1e-15, atol=1e-18)


@pytest.mark.parametrize("td,theta,phase", [(0.0, 0.0, 30.0), (0.3e-3, 0.0, 0.0),
                                           (0.1e-3, 2e3, -45.0)])
def test_sin_closed_form(oracle, td, theta, phase):
    vo, va, f = 0.4, 1.3, 2.5e3
    h, steps = 1e-5, 300
    w = _run(oracle, _divider(vo, va, f, td, theta, phase), h, steps, A.SPX_TRAN_TRAP)
    t = np.arange(steps + 1) * h
    tr = t - td
    v = vo + va * np.exp(-theta * tr) * np.sin(2 * np.pi * f * tr + phase * np.pi / 180.0)
    v = np.where(t < td, vo, v)
    v[0] = vo  # the operating point takes the DC value
    np.testing.assert_allclose(w[:, 0], v, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(w[:, 1], 0.75 * v, rtol=1e-12, atol=1e-14)


def test_rc_lowpass_steady_state_amplitude(oracle):
    # trapezoidal steps of T/2000 give the analytic low-pass gain 1/sqrt(1+(wRC)^2)
    f, R, C = 1e3, 1e3, 0.3e-6
    c = CK.Circuit("sin_rc")
    c.vsin("in", "0", 0.0, 1.0, f)
    c.resistor("in", "out", R)
    c.capacitor("out", "0", C)
    h, steps = 1 / f / 2000, 20000
    w = _run(oracle, c, h, steps, A.SPX_TRAN_TRAP)[:, 1]
    amp = 0.5 * (w[-2000:].max() - w[-2000:].min())
    assert amp == pytest.approx(1 / math.sqrt(1 + (2 * math.pi * f * R * C) ** 2), rel=1e-5)


def test_netlist_sin_source():
    p = parse_netlist("""sine source
V1 in 0 SIN(0.5 1 1k 1m 10 90)
R1 in 0 1k
.end
""")
    kinds = [e[0] for e in p.circuit.elements]
    assert A.SPX_EL_VSIN in kinds
    vals = CK.nominal_values(p.circuit)
    el = p.circuit.elements[kinds.index(A.SPX_EL_VSIN)]
    assert [vals[s] for s in el[2][:6]] == [0.5, 1.0, 1e3, 1e-3, 10.0, 90.0]
    p = parse_netlist("title\nV1 in 0 SIN(0 2 50)\nR1 in 0 1\n", name="x")
    el = p.circuit.elements[0]
    assert [CK.nominal_values(p.circuit)[s] for s in el[2][:6]] == [0.0, 2.0, 50.0, 0.0, 0.0, 0.0]
    with pytest.raises(InvalidArgument):
        parse_netlist("title\nV1 in 0 SIN(0 1)\nR1 in 0 1\n", name="x")
    with pytest.raises(InvalidArgument):  # PWL takes pairs
        parse_netlist("title\nV1 in 0 PWL(0 0 1m)\nR1 in 0 1\n", name="x")
