"""Writes tests/golden/basis_h_o.json: the H and O entries ("electron_shells" only) of the basis sets
the end-to-end checks use, read as data from the reference's BSE JSON files (Assets/Basis). Run from
the repository root with the reference tree at /root/reference."""

import json
import os

REF = "/root/reference/Assets/Basis"
SETS = {"cc-pVDZ": "Dunning/cc-pVDZ.json", "STO-3G": "Pople/STO-3G.json", "6-31G": "Pople/6-31G.json"}


def main():
    out = {}
    for name, rel in SETS.items():
        data = json.load(open(os.path.join(REF, rel)))
        out[name] = {z: {"electron_shells": data["elements"][z]["electron_shells"]} for z in ("1", "8")}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "basis_h_o.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
