"""Writes dc_linear.json and transient.json beside this script: the netlists
and ngspice-43 results that the reference's parser tests hold in
crates/spicier-parser/tests/golden_data/ (read there by
ngspice_validation.rs:1674-1700 and 1890-1960).  Only data is kept: circuit
name, netlist text, analysis parameters, the probed quantity, the expected
values and their tolerances."""
import json
import os

SRC = "/root/reference/crates/spicier-parser/tests/golden_data/"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    for name in ("dc_linear.json", "transient.json"):
        with open(os.path.join(SRC, name)) as f:
            data = json.load(f)
        out = {"generator": data["generator"], "circuits": []}
        for c in data["circuits"]:
            a = c["analysis"]
            keep = {"name": c["name"], "netlist": c["netlist"], "type": a["type"],
                    "results": a["results"], "tolerances": a["tolerances"]}
            for k in ("params", "node", "period_us", "frequency_hz"):
                if k in a:
                    keep[k] = a[k]
            out["circuits"].append(keep)
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(out, f, indent=1)
            f.write("\n")


if __name__ == "__main__":
    main()
