"""The SPICE netlist front end (spicier_amd/netlist.py) on the CPU: value
suffixes against the reference's own unit tests (spicier-core/src/units.rs:90-117),
node numbering (spicier-parser/src/parser/mod.rs:518-552), models, sources,
Monte Carlo functions and the errors for what the engine does not model."""
import pytest

from spicier_amd import _abi as A
from spicier_amd import circuit as CK
from spicier_amd import engine as E
from spicier_amd.netlist import parse_netlist, parse_value


def test_parse_value_reference_cases():
    # units.rs:91-116
    assert parse_value("1.5") == 1.5 and parse_value("-2.5") == -2.5 and parse_value("1e-3") == 1e-3
    for s, v in (("1k", 1e3), ("4.7K", 4.7e3), ("10M", 10e-3), ("10MEG", 10e6), ("2.2u", 2.2e-6),
                 ("3n", 3e-9), ("10p", 10e-12), ("5f", 5e-15), ("1G", 1e9), ("2T", 2e12),
                 ("1mil", 25.4e-6)):
        assert abs(parse_value(s) - v) <= abs(v) * 1e-10 + 1e-20, s
    assert parse_value("abc") is None and parse_value("") is None
    assert parse_value("10kohm") is None    # units are not suffixes (units.rs:37-49)


def test_node_numbering_follows_the_reference():
    p = parse_netlist("""numbering
V1 3 0 1
R1 3 out 1k
R2 out 1 1k
R3 1 2 1k
R4 2 GND 1k
""")
    # numeric names keep their number, 'out' takes max + 1
    assert p.circuit.node_names == {"0": 0, "3": 3, "out": 4, "1": 1, "2": 2}
    assert p.circuit.size == 5 and p.circuit.num_vsources == 1
    with pytest.raises(E.InvalidArgument):
        parse_netlist("gap\nR1 1 3 1k\nR2 3 0 1k\n")


def test_title_comments_continuation_and_constant_params():
    p = parse_netlist("""* this first line is the title
* a comment
.PARAM FREQ=1MEG PERIOD=1/FREQ  ; parser/mod.rs:1380-1390
R1 1 0
+ {PERIOD*1e9}
V1 1 0 DC 2
.end
""")
    assert p.title == "* this first line is the title"
    assert abs(p.params["PERIOD"] - 1e-6) < 1e-18
    kind, nodes, slots = p.circuit.elements[0]
    assert kind == A.SPX_EL_RESISTOR and abs(p.circuit.specs[slots[0]][2] - 1000.0) < 1e-9


def test_models_sources_and_defaults():
    p = parse_netlist("""models
.model dx D(IS=2e-15 N=1.5)
.model nch NMOS (VTO=0.6 KP=120u LAMBDA=0.01)
.model pch PMOS(VTO=-0.65 KP=40u)
D1 a 0 dx
D2 b 0
M1 d g 0 0 nch W=20u L=2u
M2 d g vdd vdd pch
M3 d2 g 0 0
I1 0 a 1m
V1 g 0 PULSE(0 1.8 0 10p)
V2 vdd 0 1.8
R1 vdd d 1k
R2 vdd d2 1k
R3 b vdd 1k
""")
    c = p.circuit
    sp = c.specs

    def vals(e):
        return [sp[s][2] for s in e[2] if s >= 0]

    el = c.elements
    assert el[0][0] == A.SPX_EL_DIODE and vals(el[0]) == [2e-15, 1.5]
    assert vals(el[1]) == [1e-14, 1.0]                              # diode.rs:27-38
    assert el[2][0] == A.SPX_EL_NMOS and vals(el[2]) == pytest.approx([0.6, 120e-6, 0.01, 20e-6, 2e-6])
    assert el[3][0] == A.SPX_EL_PMOS and vals(el[3]) == pytest.approx([-0.65, 40e-6, 0.0, 10e-6, 1e-6])
    assert el[4][0] == A.SPX_EL_NMOS and vals(el[4]) == pytest.approx([0.7, 2e-5, 0.0, 10e-6, 1e-6])
    assert el[5][0] == A.SPX_EL_ISOURCE and vals(el[5]) == [1e-3]
    assert el[6][0] == A.SPX_EL_VPULSE and vals(el[6]) == pytest.approx([0, 1.8, 0, 10e-12, 0, 0, 0])


def test_monte_carlo_functions():
    p = parse_netlist("""mc
.param vt=agauss(0.7, 0.03, 1) kp = gauss(110u, 0.05, 1)
R1 1 0 {aunif(1k, 10)}
R2 1 2 {unif(2k, 0.01)}
V1 2 0 {agauss(1.8, 0.09, 3)}
M1 1 2 0 0 nch
.model nch NMOS(VTO={vt} KP={kp})
""")
    sp = p.circuit.specs
    kinds = {p.circuit.slot_names[i]: s for i, s in enumerate(sp)}
    assert kinds["vt"][0] == A.SPX_P_GAUSS and kinds["vt"][2:4] == pytest.approx((0.7, 0.03))
    assert kinds["kp"][2:4] == pytest.approx((110e-6, 5.5e-6))
    assert kinds["R1"][0] == A.SPX_P_UNIFORM and kinds["R1"][2:4] == pytest.approx((990, 1010))
    assert kinds["R2"][2:4] == pytest.approx((1980, 2020))
    assert kinds["V1"][2:4] == pytest.approx((1.8, 0.03))
    # the model's VTO / KP are the .PARAM slots themselves (one shared draw)
    m = p.circuit.elements[-1]
    assert m[2][0] == p.params["VT"].slot and m[2][1] == p.params["KP"].slot


@pytest.mark.parametrize("text", [
    "t\nL1 1 0 1n\nR1 1 0 1\nK1 L1 R1 0.9\n",            # K naming a resistor
    "t\nL1 1 0 1n\nL2 2 0 1n\nK1 L1 L2 1.5\n",          # coupling above 1
    "t\nX1 1 0 sub\n",                                  # subcircuit
    "t\n.param a=agauss(1,0.1,1)\nR1 1 0 {2*a}\n",      # arithmetic on a sampled value
    "t\nV1 1 0 PULSE 0 1\nR1 1 0 1k\n",                 # PULSE without parentheses
    "t\n.model q1 NPN(BF=100)\nR1 1 0 1k\n",            # model outside the engine
    "t\nR1 1 0 1kohm\n",                                # unit, not a suffix
])
def test_errors_for_what_the_engine_does_not_model(text):
    with pytest.raises(E.InvalidArgument):
        parse_netlist(text)


def test_inductor_parses_as_reference_test():
    # parser/mod.rs:640-652: V1 and L1 each carry a branch current
    p = parse_netlist("Inductor Test\nV1 1 0 10\nL1 1 2 1m\nR1 2 0 100\n.end\n")
    c = p.circuit
    assert len(c.elements) == 3 and c.num_nodes == 2 and c.size == 4
    kind, nodes, slots = c.elements[1]
    assert kind == A.SPX_EL_INDUCTOR and nodes[:2] == [1, 2]
    assert CK.nominal_values(c)[slots[0]] == pytest.approx(1e-3)


def test_mutual_inductance_read_without_effect():
    # a coupled LC filter like parser/mod.rs:2066's: K is read and checked, and
    # the circuit program is the one without it (mutual.rs:108-114 stamps nothing)
    text = ("t\nV1 1 0 DC 1\nL1 1 2 1m\nC1 2 0 1u\nL2 3 0 1m\nR1 3 0 50\nK1 L1 L2 0.99\n"
            ".end\n")
    p = parse_netlist(text)
    assert p.couplings == [("K1", ["L1", "L2"], 0.99)]
    q = parse_netlist(text.replace("K1 L1 L2 0.99\n", ""))
    assert p.circuit.elements == q.circuit.elements
