"""Writes dc_controlled_sources.json beside this script: the netlists and
ngspice-43 DC operating points that the reference's parser tests hold in
crates/spicier-parser/tests/golden_data/dc_controlled_sources.json (read
there by ngspice_validation.rs:1756-1810).  Only data is kept: circuit name,
netlist text, expected node voltages and their tolerance."""
import json
import os

SRC = ("/root/reference/crates/spicier-parser/tests/golden_data/"
       "dc_controlled_sources.json")


def main():
    with open(SRC) as f:
        data = json.load(f)
    out = {"generator": data["generator"], "circuits": []}
    for c in data["circuits"]:
        a = c["analysis"]
        out["circuits"].append({"name": c["name"], "netlist": c["netlist"],
                                "results": a["results"],
                                "voltage_tol": a["tolerances"]["voltage"]})
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dc_controlled_sources.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
