"""A small reader of include/spicier_hip.h for the ABI tests: function
prototypes, struct fields and enum constants, as plain Python data."""
import re

_COMMENT = re.compile(r"/\*.*?\*/|//[^\n]*", re.S)


def _clean(text):
    text = _COMMENT.sub(" ", text)
    return re.sub(r"^\s*#.*$", " ", text, flags=re.M)


def defines(text):
    """{name: int} of the header's integer #defines."""
    return {m.group(1): int(m.group(2), 0)
            for m in re.finditer(r"^\s*#define\s+(\w+)\s+\(?(-?(?:0x)?[0-9a-fA-F]+)\)?\s*$",
                                 _COMMENT.sub(" ", text), flags=re.M)}


_DEFINES = {}


def _ctype(decl):
    """('const double*', 'name', array_len or None) from a declarator."""
    decl = " ".join(decl.split())
    m = re.match(r"^(.*?)([A-Za-z_]\w*)\s*(\[(\w+)\])?$", decl)
    if not m:
        raise ValueError(decl)
    ty = m.group(1).replace(" *", "*").strip()
    ty = re.sub(r"\s+\*", "*", ty)
    n = m.group(4)
    if n is not None:
        n = int(n) if n.isdigit() else _DEFINES[n]
    return ty, m.group(2), n


def functions(text):
    """[(return type, name, [(arg type, arg name)])]"""
    t = _clean(text)
    t = re.sub(r"typedef\s+struct\s+\w+\s*\{.*?\}\s*\w+\s*;", ";", t, flags=re.S)
    t = re.sub(r"enum\s*\w*\s*\{.*?\}\s*\w*\s*;", ";", t, flags=re.S)
    t = t.replace("{", ";").replace("}", ";")
    out = []
    for stmt in t.split(";"):
        m = re.match(r"^\s*([A-Za-z_][\w \*]*?)\b(spx_\w+)\s*\((.*)\)\s*$", stmt, flags=re.S)
        if not m:
            continue
        ret = re.sub(r"\s+\*", "*", " ".join(m.group(1).split()))
        if ret.startswith("typedef"):
            continue
        args = " ".join(m.group(3).split())
        params = [] if args in ("", "void") else [_ctype(a)[:2] for a in args.split(",")]
        out.append((ret, m.group(2), params))
    return out


def structs(text):
    """{typedef name: [(field type, field name, array length or None)]}"""
    _DEFINES.update(defines(text))
    t = _clean(text)
    out = {}
    for m in re.finditer(r"typedef\s+struct\s+\w+\s*\{(.*?)\}\s*(\w+)\s*;", t, flags=re.S):
        fields = []
        for decl in m.group(1).split(";"):
            decl = decl.strip()
            if not decl:
                continue
            head = decl.split(",")
            ty, name, n = _ctype(head[0])
            fields.append((ty, name, n))
            base = re.sub(r"\*+$", "", ty)
            for extra in head[1:]:
                stars = len(extra) - len(extra.lstrip(" *"))
                e = extra.strip()
                ptr = "*" * e.count("*")
                fields.append((base + ptr, e.strip("* "), None))
        out[m.group(2)] = fields
    return out


def opaque(text):
    """Names of structs declared without a body (handles)."""
    return re.findall(r"typedef\s+struct\s+(\w+)\s+\1\s*;", _clean(text))


def constants(text):
    """{name: value} of every enumerator (explicit or counted on)."""
    t = _clean(text)
    out = {}
    for m in re.finditer(r"enum\s*\w*\s*\{(.*?)\}", t, flags=re.S):
        v = -1
        for item in m.group(1).split(","):
            item = item.strip()
            if not item:
                continue
            if "=" in item:
                name, val = (s.strip() for s in item.split("=", 1))
                v = int(val, 0)
            else:
                name, v = item, v + 1
            out[name] = v
    return out
