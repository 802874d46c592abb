"""Extract the golden geometry vector from the reference's only geometry dump.

Source: obsolete/test_geometry.log of the reference (a QfeLatticeS2(q=5, n=2) member
dump): the 42 site coordinates (6 significant digits) and the neighbour list of every
site.  Writes tests/golden/qfe_n2_geometry.json (data only).  Run once, from a machine
that has the reference checked out:  python tests/golden/make_fixtures.py [REF_ROOT]
"""
import json
import pathlib
import re
import sys

REF = pathlib.Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
OUT = pathlib.Path(__file__).resolve().parent / "qfe_n2_geometry.json"


def main():
    lines = (REF / "obsolete" / "test_geometry.log").read_text().splitlines()
    i = next(k for k, l in enumerate(lines) if l.startswith("std::vector<Vec3> r ="))
    coords = []
    for l in lines[i + 2:]:
        m = re.match(r"^\s*(\d+)\s*:\s*(.*)$", l)
        if not m:
            break
        coords.append([float(v) for v in m.group(2).split()])
    j = next(k for k, l in enumerate(lines) if l.startswith("std::vector<QfeSite> sites ="))
    nbrs = []
    for l in lines[j + 1:]:
        if not l.strip():
            break
        nbrs.append([int(v) for v in l.split()])
    assert len(coords) == 42 and len(nbrs) == 42
    OUT.write_text(json.dumps({"source": "obsolete/test_geometry.log", "n_refine": 2,
                               "sites": coords, "neighbours": nbrs}, indent=0))
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
