"""Subcircuit flattening in the netlist reader (parser/subcircuit.rs), pinned
with the reference's own subcircuit tests (parser/mod.rs:1067-1700,
tests/parse_and_simulate.rs:494-590) run through the oracle, and with
hand-solved dividers for parameter precedence, nesting, MOSFET instance
overrides and current control inside a definition."""
import numpy as np
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E
from spicier_amd.netlist import parse_netlist


def _x(oracle, c):
    out = oracle.sweep_dcop(c, CK.nominal_values(c).reshape(1, -1))
    assert out["status"][0] == A.SPX_CONVERGED
    return out["x"][0]


def _v(c, x, node):
    return x[c.node_names[node] - 1]


def _rvals(c):
    vals = CK.nominal_values(c)
    return sorted(vals[e[2][0]] for e in c.elements if e[0] == A.SPX_EL_RESISTOR)


def test_reference_divider_and_nested(oracle):
    # parse_and_simulate.rs:494-538 and 540-590
    c = parse_netlist("Subcircuit Test\n.SUBCKT VDIV in out\nR1 in out 1k\nR2 out 0 1k\n"
                      ".ENDS VDIV\nV1 1 0 DC 10\nX1 1 2 VDIV\n.end\n").circuit
    assert len(c.elements) == 3
    x = _x(oracle, c)
    assert _v(c, x, "1") == pytest.approx(10.0) and _v(c, x, "2") == pytest.approx(5.0)
    c = parse_netlist("Nested\n.SUBCKT RES a b\nR1 a b 1k\n.ENDS RES\n.SUBCKT TWORES in out\n"
                      "X1 in mid RES\nX2 mid out RES\n.ENDS TWORES\nV1 1 0 DC 10\n"
                      "X1 1 2 TWORES\nR3 2 0 2k\n.end\n").circuit
    assert len(c.elements) == 4
    x = _x(oracle, c)
    assert _v(c, x, "2") == pytest.approx(5.0, rel=1e-12)
    assert "X1_mid" in c.node_names  # the internal node of instance X1


@pytest.mark.parametrize("text,want", [
    # parser/mod.rs:1588-1700: defaults, overrides, several parameters,
    # nested expressions, global parameters, precedence
    (".SUBCKT RES in out PARAMS: R=1k\nR1 in out {R}\n.ENDS RES\nV1 1 0 10\nX1 1 0 RES\n",
     [1e3]),
    (".SUBCKT RES in out PARAMS: R=1k\nR1 in out {R}\n.ENDS RES\nV1 1 0 10\n"
     "X1 1 0 RES PARAMS: R=2k\n", [2e3]),
    (".SUBCKT INNER a b PARAMS: R=1k\nR1 a b {R}\n.ENDS INNER\n.SUBCKT OUTER in out PARAMS: W=1k\n"
     "X1 in out INNER PARAMS: R={W*2}\n.ENDS OUTER\nV1 1 0 10\nX1 1 0 OUTER PARAMS: W=500\n",
     [1e3]),
    (".PARAM SCALE=2\n.SUBCKT RES in out PARAMS: R=1k\nR1 in out {R*SCALE}\n.ENDS RES\n"
     "V1 1 0 10\nX1 1 0 RES\n", [2e3]),
    (".PARAM R=500\n.SUBCKT TEST in out PARAMS: R=1k\nR1 in out {R}\n.ENDS TEST\nV1 1 0 10\n"
     "X1 1 0 TEST PARAMS: R=2k\n", [2e3]),
])
def test_reference_parameter_cases(text, want):
    c = parse_netlist("title\n" + text + ".end\n").circuit
    assert len(c.elements) == 2
    assert _rvals(c) == pytest.approx(want)


def test_rc_filter_params_and_transient_names(oracle):
    # parser/mod.rs:1623-1638 (R=2k, C=100n), then the RC step through the oracle
    c = parse_netlist("t\n.SUBCKT RCFILT in out PARAMS: R=1k C=1u\nR1 in mid {R}\nC1 mid out {C}\n"
                      ".ENDS RCFILT\nV1 1 0 10\nX1 1 2 RCFILT PARAMS: R=2k C=100n\n"
                      "R2 2 0 1k\n.end\n").circuit
    vals = CK.nominal_values(c)
    caps = [vals[e[2][0]] for e in c.elements if e[0] == A.SPX_EL_CAPACITOR]
    assert caps == pytest.approx([100e-9]) and _rvals(c) == pytest.approx([1e3, 2e3])


def test_mosfet_override_key_is_not_a_parameter(oracle):
    # a definition parameter named W must not replace the W= key of an instance override
    c = parse_netlist("t\n.model nch NMOS(VTO=0.5 KP=100u)\n.SUBCKT INV in out vdd PARAMS: W=4u\n"
                      "M1 out in 0 0 nch W={W} L=1u\nR1 vdd out 10k\n.ENDS INV\n"
                      "VDD 3 0 1.8\nVIN 1 0 1.8\nX1 1 2 3 INV PARAMS: W=8u\n.end\n").circuit
    m = [e for e in c.elements if e[0] == A.SPX_EL_NMOS][0]
    vals = CK.nominal_values(c)
    assert vals[m[2][3]] == pytest.approx(8e-6) and vals[m[2][4]] == pytest.approx(1e-6)
    x = _x(oracle, c)
    # saturation is excluded at vds well below vgs - vt; check KCL by hand: the
    # drain current equals (1.8 - v) / 10k in the linear region of level 1
    v = _v(c, x, "2")
    beta = 100e-6 * 8.0
    ids = beta * ((1.8 - 0.5) * v - 0.5 * v * v)
    assert (1.8 - v) / 10e3 == pytest.approx(ids, rel=1e-9)


def test_current_control_inside_a_definition(oracle):
    # F inside the definition reads the definition's own series sense source:
    # each instance drives 2 x its sense current into its output's 1 kOhm
    c = parse_netlist("t\n.SUBCKT MIR in sink out\nVs in sink 0\nF1 0 out Vs 2\n.ENDS MIR\n"
                      "V1 1 0 1\nR1 1 2 1k\nX1 2 0 3 MIR\nR2 3 0 1k\n"
                      "V2 4 0 2\nR5 4 5 1k\nX2 5 0 6 MIR\nR6 6 0 1k\n.end\n").circuit
    x = _x(oracle, c)
    assert _v(c, x, "3") == pytest.approx(2.0, rel=1e-12)
    assert _v(c, x, "6") == pytest.approx(4.0, rel=1e-12)


@pytest.mark.parametrize("text", [
    "t\nX1 1 0 NOPE\nR1 1 0 1k\n",                                   # unknown definition
    "t\n.SUBCKT A p q\nR1 p q 1k\n.ENDS\nX1 1 A\nR1 1 0 1k\n",        # port count
    "t\n.SUBCKT A p q\nR1 p q 1k\nR2 1 0 1k\n",                      # no .ENDS
    "t\n.SUBCKT A p\nX1 p A\n.ENDS\nX1 1 A\nR1 1 0 1k\n",            # recursion
])
def test_subckt_errors(text):
    with pytest.raises(E.InvalidArgument):
        parse_netlist(text)


def test_transmission_line_inside_a_definition(oracle):
    # two instances of a subcircuit holding a T line: each line's ports take
    # the instance's connections, its internal section nodes are named after
    # the flattened line (TX1_1_int1), so the lines short 1 -> 2 -> 3 at DC
    # and R1 sees V1 directly
    c = parse_netlist("t\n.SUBCKT line a b\nT1 a 0 b 0 Z0=50 TD=1n NL=2\n.ENDS\n"
                      "V1 1 0 DC 10\nX1 1 2 line\nX2 2 3 line\nR1 3 0 100\n.end\n").circuit
    assert set(c.node_names) == {"0", "1", "2", "3", "TX1_1_int1", "TX2_1_int1"}
    assert c.size - c.num_nodes == 1 + 2 * 2  # V1's branch and two sections per line
    x = _x(oracle, c)
    for node in ("1", "2", "3", "TX1_1_int1", "TX2_1_int1"):
        assert _v(c, x, node) == pytest.approx(10.0, rel=1e-12)
    assert x[c.num_nodes] == pytest.approx(-0.1, rel=1e-12)  # V1 supplies 10 V / 100 Ohm


def test_couplings_inside_a_definition_and_before_their_inductors(oracle):
    # a K may name inductors that come later in the deck; inside a definition
    # it names that instance's own inductors
    p = parse_netlist("t\nK1 L1 L2 0.5\nV1 1 0 DC 1\nL1 1 2 1m\nR1 2 0 1k\nL2 3 0 1m\n"
                      "R2 3 0 1k\n.end\n")
    assert p.couplings == [("K1", ["L1", "L2"], 0.5)]
    p = parse_netlist("t\n.SUBCKT xf a b\nLa a 0 1m\nLb b 0 1m\nK1 La Lb 0.9\n.ENDS\n"
                      "V1 1 0 DC 1\nR1 1 2 1k\nX1 2 3 xf\nX2 3 4 xf\nR2 4 0 1k\n.end\n")
    assert p.couplings == [("KX1_1", ["LX1_A", "LX1_B"], 0.9),
                           ("KX2_1", ["LX2_A", "LX2_B"], 0.9)]
    with pytest.raises(E.InvalidArgument):
        parse_netlist("t\nK1 L1 L9 0.5\nL1 1 0 1m\nR1 1 0 1k\n.end\n")
