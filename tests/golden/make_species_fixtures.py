"""Writes tests/golden/species_tutorial.json: the species blocks of the
reference's tutorial configurations (tut_Sexual.xml, tut_EnvironCapAlt.xml)
and the agent rows of tut_Sexual.dat, as data (module names, parameter
strings, priorities, agent columns), read with qhg4_amd.config.

    python tests/golden/make_species_fixtures.py [/root/reference/QHG4]
"""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from qhg4_amd.config import read_species_xml  # noqa: E402


def main():
    qhg = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/QHG4"
    d = os.path.join(qhg, "tutorial_data", "xmldat")
    out = {"species": {}, "agents": {}}
    for name in ("tut_Sexual", "tut_EnvironCapAlt"):
        out["species"][name] = read_species_xml(os.path.join(d, name + ".xml"))
    rows = [ln.rstrip("\n") for ln in open(os.path.join(d, "tut_Sexual.dat")) if ln.strip()]
    out["agents"]["tut_Sexual"] = rows
    path = os.path.join(ROOT, "tests", "golden", "species_tutorial.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(path)


if __name__ == "__main__":
    main()
