"""Copies the stored outputs of the reference notebook
notebooks/Gershgorin_TH3D.ipynb into tests/golden/notebook_outputs.json.

The notebook is the reference's only file with stored expected values: the
SymPy-derived Gershgorin bound (cell 2 output) and the pressure derivative of
the fluid-density polynomial (cell 4 output).  Only the outputs are kept (the
printed Julia expressions); they pin the oracle's GershgorinPoisson and its
d(rho_f)/dp closure.  Run from the repository root:
    python tests/golden/make_notebook_outputs.py [path/to/reference]
"""
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent


def _text(output):
    if "text" in output:
        return "".join(output["text"])
    return "".join(output.get("data", {}).get("text/plain", []))


def main(ref=Path("/root/reference")):
    nb = json.loads((Path(ref) / "notebooks" / "Gershgorin_TH3D.ipynb").read_text())
    code = [c for c in nb["cells"] if c["cell_type"] == "code"]
    outs = {}
    for c in code:
        for o in c.get("outputs", []):
            t = _text(o).strip()
            if t.startswith("abs(kE"):
                outs["gershgorin_bound"] = t
            elif "T_C" in t:
                outs["drhof_dp"] = t.strip("'")
    assert set(outs) == {"gershgorin_bound", "drhof_dp"}, outs.keys()
    outs["source"] = "notebooks/Gershgorin_TH3D.ipynb, stored outputs of the cells that print julia_code(...)"
    (HERE / "notebook_outputs.json").write_text(json.dumps(outs, indent=1) + "\n")
    print(json.dumps(outs, indent=1))


if __name__ == "__main__":
    main(*(sys.argv[1:2]))
