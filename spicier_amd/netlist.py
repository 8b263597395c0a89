"""SPICE netlist front end: a netlist in the subset the engine's device models
cover becomes a Circuit (the element table the library compiles into its
circuit program), so Monte Carlo, corner and parameter sweeps run from .sp
text instead of programmatic fixtures.

Follows spicier-parser's conventions (crates/spicier-parser/src):
  * the first line is the title unless it is a dot command; '*' starts a
    comment line after it, ';' an inline comment, '+' continues the previous
    line (lexer.rs:56-160);
  * values carry SPICE suffixes T G MEG K M MIL U N P F, case-insensitive
    (spicier-core/src/units.rs:15-50);
  * E, G (4 nodes and a gain), F, H (2 nodes, the name of an earlier V, L, E or H
    element and a gain) — parser/elements.rs:716-830;
  * T (Z0= TD= [NL=], the reference's LC-section expansion, tline.rs);
  * R, C, L (parse_inductor, parser/elements.rs:188-202), V (DC value, PULSE(v1 v2 td tr tf pw per) or SIN(vo va freq td
    theta phase), missing trailing arguments 0, or PWL(t1 v1 t2 v2 ...) with
    constant points — parser/waveforms.rs:12-106), I (current flows from n+
    through the source to n-), D (optional model), M (d g s b model, W= L=
    instance overrides) (parser/elements.rs);
  * .MODEL name D(IS= N=) / NMOS(...) / PMOS(...) level 1 with VTO KP LAMBDA
    W L and the reference's defaults (parser/commands.rs:612-824,
    mosfet/level1.rs:42-65, diode.rs:27-38);
  * .PARAM name=value, including constant arithmetic.
Sampled parameters use the ngspice Monte Carlo functions, in a .PARAM or in
braces in an element or model value:
  agauss(nom, avar, sigma)  -> N(nom, avar/sigma)
  gauss(nom, rvar, sigma)   -> N(nom, nom*rvar/sigma)
  aunif(nom, avar)          -> U(nom - avar, nom + avar)
  unif(nom, rvar)           -> U(nom*(1 - rvar), nom*(1 + rvar))
A .PARAM is one sampled parameter shared by every use (correlated); a
function written in a value is a parameter of its own.  Arithmetic on sampled
parameters, and elements outside the engine's models (Q, J, B), are
errors, not approximations.  K couplings are read and checked and, as in the
reference's DC and transient analyses, change nothing (ParsedNetlist.couplings).  .SUBCKT definitions with PARAMS: defaults and X
instances (nested, with PARAMS: overrides) are flattened first
(parser/subcircuit.rs).
"""
from __future__ import annotations

import ast
import operator
import re
from dataclasses import dataclass, field

from ._abi import SPX_EL_INDUCTOR as A_EL_INDUCTOR
from .circuit import BRANCH_KINDS, Circuit, Param
from .engine import InvalidArgument

# units.rs:37-49
_SUFFIX = {"T": 1e12, "G": 1e9, "MEG": 1e6, "K": 1e3, "": 1.0, "M": 1e-3, "MIL": 25.4e-6,
           "U": 1e-6, "N": 1e-9, "P": 1e-12, "F": 1e-15}
_PLAIN = re.compile(r"^[+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?$|^[+-]?(INF|INFINITY|NAN)$")


def parse_value(s: str):
    """A SPICE number with optional SI suffix (units.rs:15-50), None if not one."""
    s = s.strip().upper()
    if _PLAIN.match(s):
        return float(s)
    end = 0
    while end < len(s) and (s[end].isdigit() or s[end] in ".-+E"):
        end += 1
    if end == 0:
        return None
    num, suffix = s[:end], s[end:]
    if not _PLAIN.match(num) or suffix not in _SUFFIX:
        return None
    return float(num) * _SUFFIX[suffix]


@dataclass
class ParsedNetlist:
    circuit: Circuit
    title: str = ""
    tran: tuple | None = None        # (tstep, tstop) of a .TRAN line
    params: dict = field(default_factory=dict)   # .PARAM name -> value or Param
    # K couplings (name, inductors, k): read and checked but without effect
    # here, as in the reference's DC and transient analyses
    couplings: list = field(default_factory=list)
    # element name (upper case) -> index in circuit.elements (a T line: its
    # last section's capacitor)
    elements: dict = field(default_factory=dict)

    def unknown(self, quantity: str) -> int:
        """Unknown index of a SPICE output name: V(node) or I(source) of a V,
        L, E or H element (MNA branch current, ngspice's sign)."""
        q = quantity.strip()
        kind, arg = q[0].upper(), q[2:-1].strip()
        if q[1] != "(" or q[-1] != ")" or kind not in "VI":
            raise InvalidArgument(f"expected V(node) or I(source), got {quantity!r}")
        if kind == "V":
            if arg not in self.circuit.node_names or self.circuit.node_names[arg] == 0:
                raise InvalidArgument(f"no node {arg!r}")
            return self.circuit.node_names[arg] - 1
        i = self.elements.get(arg.upper())
        if i is None:
            raise InvalidArgument(f"no element {arg!r}")
        try:
            return self.circuit.branch_unknown(i)
        except ValueError:
            raise InvalidArgument(f"{arg} carries no branch current") from None


# ------------------------------------------------------------------ lines
def _logical_lines(text: str):
    """(line number, content) after title, comment and continuation handling."""
    raw = text.splitlines()
    out = []
    first = True
    title = ""
    for no, line in enumerate(raw, 1):
        s = line.strip()
        if first:
            if not s:
                continue
            first = False
            if not s.startswith("."):
                title = s
                continue
        if not s or s.startswith("*"):
            continue
        s = s.split(";", 1)[0].strip()
        if not s:
            continue
        if s.startswith("+"):
            if not out:
                raise InvalidArgument(f"line {no}: continuation without a line to continue")
            out[-1] = (out[-1][0], out[-1][1] + " " + s[1:].strip())
        else:
            out.append((no, s))
    return title, out


_TOK = re.compile(r"\{[^}]*\}|[A-Za-z_][\w.]*\s*\([^()]*\)|[()=,]|[^\s()=,]+")


def _tokens(s: str):
    return [t for t in _TOK.findall(s) if t != ","]


# ------------------------------------------------------------ subcircuits
# node tokens after the element name, per element letter (subcircuit
# flattening and the element reader share it)
_NODE_COUNT = {"R": 2, "C": 2, "L": 2, "V": 2, "I": 2, "D": 2, "M": 4, "E": 4, "G": 4, "F": 2,
               "H": 2, "T": 4}
_IDENT = re.compile(r"(?<![\w.])[A-Za-z_]\w*")


def _split_params(toks, line):
    """(tokens before PARAMS:, {NAME: value text}) of a .SUBCKT or X line."""
    for i, t in enumerate(toks):
        if t.upper().startswith("PARAMS:"):
            rest = ([t[7:]] if len(t) > 7 else []) + toks[i + 1:]
            kv, j = {}, 0
            while j < len(rest):
                if j + 2 < len(rest) and rest[j + 1] == "=":
                    kv[rest[j].upper()] = rest[j + 2]
                    j += 3
                else:
                    raise InvalidArgument(f"line {line}: expected name=value after PARAMS:")
            return toks[:i], kv
    return toks, {}


def _subst(text, ctx):
    """Replaces subcircuit parameter names in a value by their values."""
    def rep(m):
        v = ctx.get(m.group(0).upper())
        if v is None:
            return m.group(0)
        return v if re.fullmatch(r"[\w.+-]+", v) else f"({v})"
    return _IDENT.sub(rep, text)


def _flatten(lines):
    """Expands X instances of .SUBCKT definitions (parser/subcircuit.rs): an
    element R1 of instance X1 becomes RX1_1, an internal node n becomes
    X1_n, ports take the instance's connections and 0 / gnd stay ground;
    parameters resolve global .PARAM < the definition's PARAMS: defaults <
    the instance's PARAMS:, innermost instance last.  Nested instances
    expand recursively.  A current-controlled source names its control
    inside the definition first (renamed with it), else outside."""
    defs, top, cur = {}, [], None
    for no, s in lines:
        toks = _tokens(s)
        cmd = toks[0].upper()
        if cmd == ".SUBCKT":
            if cur is not None:
                raise InvalidArgument(f"line {no}: .SUBCKT inside .SUBCKT {cur['name']}")
            head, params = _split_params(toks[1:], no)
            if not head:
                raise InvalidArgument(f"line {no}: .SUBCKT needs a name")
            cur = {"name": head[0].upper(), "ports": [p.upper() for p in head[1:]],
                   "params": params, "body": [], "line": no}
        elif cmd == ".ENDS":
            if cur is None:
                raise InvalidArgument(f"line {no}: .ENDS without .SUBCKT")
            defs[cur["name"]] = cur
            cur = None
        elif cur is not None:
            cur["body"].append((no, s))
        else:
            top.append((no, s))
    if cur is not None:
        raise InvalidArgument(f"line {cur['line']}: .SUBCKT {cur['name']} has no .ENDS")

    def expand(no, toks, ctx, depth):
        inst = toks[0]
        head, params = _split_params(toks[1:], no)
        if len(head) < 1:
            raise InvalidArgument(f"line {no}: {inst} needs a subcircuit name")
        conns, sub = head[:-1], head[-1].upper()
        d = defs.get(sub)
        if d is None:
            raise InvalidArgument(f"line {no}: unknown subcircuit {head[-1]}")
        if len(conns) != len(d["ports"]):
            raise InvalidArgument(f"line {no}: subcircuit {head[-1]} expects {len(d['ports'])} "
                                  f"ports but {len(conns)} are given")
        if depth > 32:
            raise InvalidArgument(f"line {no}: subcircuits nested more than 32 deep")
        local = dict(ctx)
        for k, v in d["params"].items():
            local[k] = _subst(v.strip("{}"), ctx)
        for k, v in params.items():
            local[k] = _subst(v.strip("{}"), ctx)
        ports = dict(zip(d["ports"], conns))
        names = {t[0].upper() for _, b in d["body"] for t in [_tokens(b)] if t}

        def node(t):
            if t.upper() in ports:
                return ports[t.upper()]
            if t in ("0", "gnd", "GND"):
                return t
            return f"{inst}_{t}"

        def val(t):
            if t.startswith("{") and t.endswith("}"):
                return "{" + _subst(t[1:-1], local) + "}"
            if t.upper() in local:
                return "{" + local[t.upper()] + "}"
            m = re.match(r"^([A-Za-z_][\w.]*\s*\()(.*)\)$", t)
            if m:
                return m.group(1) + _subst(m.group(2), local) + ")"
            return t

        out = []
        for bno, b in d["body"]:
            bt = _tokens(b)
            if bt[0].startswith("."):
                out.append((bno, b))
                continue
            kind = bt[0][0].upper()
            name = f"{bt[0][0]}{inst}_{bt[0][1:]}"
            if kind == "X":
                bh, bp = _split_params(bt[1:], bno)
                sub_t = [name] + [node(t) for t in bh[:-1]] + [bh[-1]]
                if bp:
                    sub_t += ["PARAMS:"] + [x for k, v in bp.items()
                                            for x in (k, "=", "{" + _subst(v.strip("{}"), local) + "}")]
                out.extend(expand(no, sub_t, local, depth + 1))
                continue
            nn = _NODE_COUNT.get(kind, 0)
            toks2 = [name]
            rest = bt[1:]
            for i, t in enumerate(rest):
                if i < nn:
                    toks2.append(node(t))
                elif kind in "FH" and i == nn and t.upper() in names:
                    toks2.append(f"{t[0]}{inst}_{t[1:]}")
                elif kind == "K" and i < len(rest) - 1 and t.upper() in names:
                    toks2.append(f"{t[0]}{inst}_{t[1:]}")  # an inductor of this definition
                elif i + 1 < len(rest) and rest[i + 1] == "=":
                    toks2.append(t)  # the key of an instance override such as W=...
                else:
                    toks2.append(val(t))
            out.append((bno, " ".join(toks2)))
        return out

    flat = []
    for no, s in top:
        if s[0].upper() == "X":
            flat.extend(expand(no, _tokens(s), {}, 0))
        else:
            flat.append((no, s))
    return flat


# ------------------------------------------------------------ expressions
_MC = {"AGAUSS", "GAUSS", "AUNIF", "UNIF"}
_OPS = {ast.Add: operator.add, ast.Sub: operator.sub, ast.Mult: operator.mul,
        ast.Div: operator.truediv, ast.Pow: operator.pow, ast.USub: operator.neg,
        ast.UAdd: operator.pos}


class _Scope:
    def __init__(self, circuit: Circuit):
        self.c = circuit
        self.params: dict = {}

    def _const(self, node, line):
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            return float(node.value)
        if isinstance(node, ast.Name):
            key = node.id.upper()
            if key in self.params:
                v = self.params[key]
                if isinstance(v, Param):
                    raise InvalidArgument(f"line {line}: arithmetic on the sampled parameter "
                                          f"{node.id} is not supported")
                return v
            raise InvalidArgument(f"line {line}: unknown parameter {node.id}")
        if isinstance(node, ast.BinOp) and type(node.op) in _OPS:
            return _OPS[type(node.op)](self._const(node.left, line), self._const(node.right, line))
        if isinstance(node, ast.UnaryOp) and type(node.op) in _OPS:
            return _OPS[type(node.op)](self._const(node.operand, line))
        raise InvalidArgument(f"line {line}: unsupported expression")

    def value(self, tok: str, line: int, name: str | None = None):
        """A float, or a Param for a sampled value."""
        t = tok.strip()
        if t.startswith("{") and t.endswith("}"):
            t = t[1:-1].strip()
        v = parse_value(t)
        if v is not None:
            return v
        m = re.match(r"^([A-Za-z_]\w*)\s*\((.*)\)$", t)
        if m and m.group(1).upper() in _MC:
            args = [a for a in re.split(r"[,\s]+", m.group(2).strip()) if a]
            vals = []
            for a in args:
                x = self.value(a, line)
                if isinstance(x, Param):
                    raise InvalidArgument(f"line {line}: {m.group(1)} takes constant arguments")
                vals.append(x)
            return self._mc(m.group(1).upper(), vals, line, name)
        if t.upper() in self.params:
            return self.params[t.upper()]
        # constant arithmetic, SPICE numbers with suffixes turned into floats
        expr = re.sub(r"(?<![\w.])(\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+)([A-Za-z]+)",
                      lambda mm: repr(parse_value(mm.group(0))) if parse_value(mm.group(0))
                      is not None else mm.group(0), t)
        try:
            tree = ast.parse(expr, mode="eval").body
        except SyntaxError:
            raise InvalidArgument(f"line {line}: cannot read value {tok!r}") from None
        return self._const(tree, line)

    def _mc(self, fn, a, line, name):
        need = {"AGAUSS": 3, "GAUSS": 3, "AUNIF": 2, "UNIF": 2}[fn]
        if len(a) != need:
            raise InvalidArgument(f"line {line}: {fn.lower()} takes {need} arguments")
        if fn in ("AGAUSS", "GAUSS"):
            nom, var, sig = a
            if not sig > 0:
                raise InvalidArgument(f"line {line}: {fn.lower()} needs sigma > 0")
            sd = var / sig if fn == "AGAUSS" else abs(nom) * var / sig
            return self.c.gauss(nom, sd, name=name)
        nom, var = a
        lo, hi = (nom - var, nom + var) if fn == "AUNIF" else (nom * (1 - var), nom * (1 + var))
        return self.c.uniform(min(lo, hi), max(lo, hi), name=name)


# ------------------------------------------------------------------ nodes
class _Nodes:
    """Node numbering of parser/mod.rs:178-183,518-552: 0 / gnd / GND are
    ground, a numeric name is that node number unless another name holds it,
    any other name takes the highest number so far plus one."""

    def __init__(self):
        self.map = {"0": 0, "gnd": 0, "GND": 0}

    def id(self, name: str) -> int:
        if name in self.map:
            return self.map[name]
        if re.fullmatch(r"\+?\d+", name) and int(name) < 2 ** 32:
            num = int(name)
            if num == 0:
                return 0
            if not any(v == num and k != name for k, v in self.map.items()):
                self.map[name] = num
                return num
        nid = max([v for v in self.map.values() if v != 0], default=0) + 1
        self.map[name] = nid
        return nid


# ------------------------------------------------------------------ parser
_MODEL_DEFAULTS = {
    "D": {"IS": 1e-14, "N": 1.0},
    "NMOS": {"VTO": 0.7, "KP": 2e-5, "LAMBDA": 0.0, "W": 10e-6, "L": 1e-6},
    "PMOS": {"VTO": -0.7, "KP": 1e-5, "LAMBDA": 0.0, "W": 10e-6, "L": 1e-6},
}


def _kv(tokens, line, scope, prefix=""):
    """NAME = value pairs (anything else is skipped, as the reference does)."""
    out = {}
    i = 0
    while i < len(tokens):
        if i + 2 < len(tokens) and tokens[i + 1] == "=":
            key = tokens[i].upper()
            out[key] = scope.value(tokens[i + 2], line, name=f"{prefix}{key}")
            i += 3
        else:
            i += 1
    return out


def parse_netlist(text: str, name: str | None = None) -> ParsedNetlist:
    """Compile a SPICE netlist into a Circuit (see the module docstring)."""
    title, lines = _logical_lines(text)
    lines = _flatten(lines)
    c = Circuit(name or (title or "netlist"))
    scope = _Scope(c)
    models: dict = {}
    pending = []
    tran = None
    # dot commands first: .PARAM and .MODEL may follow the elements using them
    for no, s in lines:
        toks = _tokens(s)
        cmd = toks[0].upper()
        if cmd == ".PARAM":
            body = s[len(toks[0]):].strip()
            for m in re.finditer(r"([A-Za-z_]\w*)\s*=\s*(\{[^}]*\}|[A-Za-z_]\w*\s*\([^()]*\)|[^\s]+)",
                                 body):
                key = m.group(1).upper()
                scope.params[key] = scope.value(m.group(2), no, name=m.group(1))
        elif cmd == ".MODEL":
            if len(toks) < 3:
                raise InvalidArgument(f"line {no}: .MODEL needs a name and a type")
            mname = toks[1].upper()
            mtype = toks[2].upper()
            rest = toks[3:]
            m = re.match(r"^([A-Za-z]+)\s*\((.*)\)$", toks[2])
            if m:  # "D(IS=1e-14)" read as one token
                mtype, rest = m.group(1).upper(), _tokens(m.group(2)) + rest
            if mtype not in _MODEL_DEFAULTS:
                raise InvalidArgument(f"line {no}: model type {mtype} is not supported")
            p = dict(_MODEL_DEFAULTS[mtype])
            level = None
            kv = _kv(rest, no, scope, prefix=f"{mname}.")
            if "LEVEL" in kv:
                level = kv.pop("LEVEL")
            if level not in (None, 1, 1.0):
                raise InvalidArgument(f"line {no}: only level-1 MOSFET models are supported")
            for k, v in kv.items():
                if k in p:
                    p[k] = v
            models[mname] = (mtype, p)
        elif cmd == ".TRAN":
            if len(toks) < 3:
                raise InvalidArgument(f"line {no}: .TRAN needs tstep and tstop")
            tran = (scope.value(toks[1], no), scope.value(toks[2], no))
        elif cmd.startswith("."):
            continue  # .END, .OP, analyses the sweep engine does not take
        else:
            pending.append((no, s, toks))
    nodes = _Nodes()
    by_name = {}  # element name -> index (find_vsource_branch_index, netlist.rs:412-420)
    couplings = []
    for no, s, toks in pending:
        kind = toks[0][0].upper()
        nm = toks[0]
        args = toks[1:]
        nn = _NODE_COUNT.get(kind, 0)
        args = [nodes.id(a) if k < nn else a for k, a in enumerate(args)]

        def need(k):
            if len(args) < k:
                raise InvalidArgument(f"line {no}: {nm} needs {k} fields")

        if kind in "RCL":  # parse_inductor (elements.rs:188-202): L name n+ n- value
            need(3)
            v = scope.value(args[2], no, name=nm)
            {"R": c.resistor, "C": c.capacitor, "L": c.inductor}[kind](args[0], args[1], v)
        elif kind == "V":
            need(3)
            rest = args[2:]
            pulse = next((t for t in rest if t.upper().startswith("PULSE")), None)
            sinw = next((t for t in rest if t.upper().startswith("SIN")), None)
            pwl = next((t for t in rest if t.upper().startswith("PWL")), None)
            if pwl is not None:  # parse_pwl_waveform (parser/waveforms.rs:74-106)
                if "(" not in pwl or not pwl.endswith(")"):
                    raise InvalidArgument(f"line {no}: expected PWL(t1 v1 t2 v2 ...)")
                inner = _tokens(pwl[pwl.index("(") + 1:-1])
                if not inner or len(inner) % 2:
                    raise InvalidArgument(f"line {no}: PWL needs time-value pairs")
                vals = [scope.value(t, no, name=f"{nm}.PWL{k}") for k, t in enumerate(inner)]
                if any(isinstance(v, Param) for v in vals):
                    raise InvalidArgument(f"line {no}: PWL points must be constants")
                c.vpwl(args[0], args[1], list(zip(vals[0::2], vals[1::2])))
            elif sinw is not None:
                if "(" not in sinw or not sinw.endswith(")"):
                    raise InvalidArgument(f"line {no}: expected SIN(vo va freq [td theta phase])")
                inner = _tokens(sinw[sinw.index("(") + 1:-1])
                if len(inner) < 3 or len(inner) > 6:
                    raise InvalidArgument(f"line {no}: SIN takes 3 to 6 values")
                vals = [scope.value(t, no, name=f"{nm}.SIN{k}") for k, t in enumerate(inner)]
                vals += [0.0] * (6 - len(vals))
                c.vsin(args[0], args[1], *vals)
            elif pulse is not None:
                if "(" not in pulse or not pulse.endswith(")"):
                    raise InvalidArgument(f"line {no}: expected PULSE(v1 v2 [td tr tf pw per])")
                inner = _tokens(pulse[pulse.index("(") + 1:-1])
                if len(inner) < 2 or len(inner) > 7:
                    raise InvalidArgument(f"line {no}: PULSE takes 2 to 7 values")
                vals = [scope.value(t, no, name=f"{nm}.PULSE{k}") for k, t in enumerate(inner)]
                vals += [0.0] * (7 - len(vals))
                c.vpulse(args[0], args[1], *vals)
            else:
                vtok = rest[1] if rest[0].upper() == "DC" and len(rest) > 1 else rest[0]
                c.vsource(args[0], args[1], scope.value(vtok, no, name=nm))
        elif kind in "EG":  # parse_vcvs / parse_vccs (elements.rs:716-760): 4 nodes, gain
            need(5)
            v = scope.value(args[4], no, name=nm)
            (c.vcvs if kind == "E" else c.vccs)(args[0], args[1], args[2], args[3], v)
        elif kind in "FH":  # parse_cccs / parse_ccvs (elements.rs:762-830): 2 nodes, source, gain
            need(4)
            ctrl = by_name.get(args[2].upper())
            if ctrl is None or c.elements[ctrl][0] not in BRANCH_KINDS:
                raise InvalidArgument(f"line {no}: {nm}: controlling source {args[2]} not found "
                                      "(it must be an earlier V, L, E or H element)")
            v = scope.value(args[3], no, name=nm)
            (c.cccs if kind == "F" else c.ccvs)(args[0], args[1], ctrl, v)
        elif kind == "K":  # parse_mutual_inductance (elements.rs:633-700)
            # Kname L1 L2 [L3 ...] k.  The reference's MutualInductance stamps
            # nothing at DC (mutual.rs:108-114) and its transient state builder
            # skips it (cli/stampers.rs:541-620): the coupling reaches only AC
            # analysis, so it changes no DC or transient result there either
            need(3)
            inds = args[:-1]
            if len(inds) < 2:
                raise InvalidArgument(f"line {no}: {nm} needs at least 2 inductors")
            kval = scope.value(args[-1], no, name=nm)
            if isinstance(kval, Param) or not (0.0 < kval <= 1.0):
                raise InvalidArgument(f"line {no}: {nm}: coupling must be a constant in (0, 1]")
            # the inductors may come later in the deck: checked after every element
            couplings.append((nm, [x.upper() for x in inds], kval, no))
            continue
        elif kind == "T":  # lossless line: elements.rs:963-1068, tline.rs:30-236
            need(4)
            kv = _kv(args[4:], no, scope, prefix=f"{nm}.")
            if "Z0" not in kv or "TD" not in kv:
                raise InvalidArgument(f"line {no}: {nm} needs Z0= and TD=")
            if any(isinstance(v, Param) for v in kv.values()):
                raise InvalidArgument(f"line {no}: {nm}: Z0, TD and NL must be constants")
            z0, td = kv["Z0"], kv["TD"]
            if not (z0 > 0 and td > 0):
                raise InvalidArgument(f"line {no}: {nm} needs Z0 > 0 and TD > 0")
            sections = max(int(kv.get("NL", 10)), 2)  # DEFAULT_NUM_SECTIONS, MIN_SECTIONS
            # port1+ -L- int1 -L- ... -L- port2+, C from each section's right
            # node to ground (build_transient_state, cli/stampers.rs:575-615):
            # the inductors take consecutive branch unknowns as the line's
            # current variables do; at DC they are the line's 0 V sources
            chain = ([args[0]] + [nodes.id(f"{nm}_int{k}") for k in range(1, sections)]
                     + [args[2]])
            for k in range(sections):
                c.inductor(chain[k], chain[k + 1], z0 * td / sections)
            for k in range(sections):
                c.capacitor(chain[k + 1], 0, td / (z0 * sections))
        elif kind == "I":
            need(3)
            rest = args[2:]
            vtok = rest[1] if rest[0].upper() == "DC" and len(rest) > 1 else rest[0]
            c.isource(args[0], args[1], scope.value(vtok, no, name=nm))
        elif kind == "D":
            need(2)
            p = dict(_MODEL_DEFAULTS["D"])
            if len(args) > 2:
                m = models.get(args[2].upper())
                if m and m[0] == "D":
                    p = m[1]
            c.diode(args[0], args[1], is_=p["IS"], n=p["N"])
        elif kind == "M":
            need(4)
            mtype, p = "NMOS", dict(_MODEL_DEFAULTS["NMOS"])
            rest = args[4:]
            if rest and "=" not in rest[0] and (len(rest) < 2 or rest[1] != "="):
                m = models.get(rest[0].upper())
                if m and m[0] in ("NMOS", "PMOS"):
                    mtype, p = m[0], dict(m[1])
                rest = rest[1:]
            inst = _kv(rest, no, scope, prefix=f"{nm}.")
            for k in ("W", "L"):
                if k in inst:
                    p[k] = inst[k]
            # the bulk node (args[3]) has no effect: level 1 without body effect
            (c.nmos if mtype == "NMOS" else c.pmos)(args[0], args[1], args[2], vto=p["VTO"],
                                                    kp=p["KP"], lam=p["LAMBDA"], w=p["W"],
                                                    l=p["L"])
        else:
            raise InvalidArgument(f"line {no}: element {nm} ({kind}) is not supported by the "
                                  "sweep engine")
        by_name[nm.upper()] = len(c.elements) - 1
    for nm, inds, kval, no in couplings:
        for ln in inds:
            i = by_name.get(ln)
            if i is None or c.elements[i][0] != A_EL_INDUCTOR:
                raise InvalidArgument(f"line {no}: {nm}: {ln} is not an inductor of the circuit")
    couplings = [(nm, inds, kval) for nm, inds, kval, _ in couplings]
    # unknown k - 1 is node k; a node number without elements would be an
    # empty (singular) row of every instance's system
    names = {k: v for k, v in nodes.map.items() if v != 0}
    used = sorted(set(names.values()))
    if used != list(range(1, len(used) + 1)):
        missing = sorted(set(range(1, (used[-1] if used else 0) + 1)) - set(used))
        raise InvalidArgument(f"node numbers {missing} have no element (every system would be "
                              "singular)")
    c.node_names = {"0": 0, **names}
    return ParsedNetlist(c, title, tran, dict(scope.params), couplings, dict(by_name))
