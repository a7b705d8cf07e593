"""Writes tests/golden/optimizer_weights.json: the band combos and the
quadrature weights the reference's wavelength optimizer printed in its
committed reports (reports/*/ranking.csv and report.md of the reference
checkout).  The reports with no "Weight strategy" line predate the option
and used the Voronoi widths (main.rs:44-53 default).

Run in the build container (the reference checkout is not on GPU hosts):
    python tests/golden/make_optimizer_weights.py /root/reference
"""
import csv
import json
import os
import re
import sys


def main(ref_root):
    rows = []
    reports = os.path.join(ref_root, "reports")
    for name in sorted(os.listdir(reports)):
        d = os.path.join(reports, name)
        md = open(os.path.join(d, "report.md")).read()
        m = re.search(r"Weight strategy: `([a-z-]+)`", md)
        strategy = m.group(1) if m else "voronoi"
        with open(os.path.join(d, "ranking.csv")) as f:
            for r in csv.DictReader(f):
                lam = [float(r[k]) for k in sorted(r) if k.startswith("lambda")]
                w = [float(r[k]) for k in sorted(r) if k.startswith("weight")]
                rows.append({"report": name, "strategy": strategy, "lambda_nm": lam, "weights_nm": w})
    out = {"source": "reports/*/ranking.csv of the reference (printed with 6 decimals)",
           "grid_nm": [380.0 + 10.0 * i for i in range(41)], "rows": rows}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "optimizer_weights.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {len(rows)} rows to {path}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
