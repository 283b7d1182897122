"""Extract the published output of the reference example kinFerTron_dns
(KINSOL Newton with and without line search, dense difference-quotient
Jacobian, inequality constraints) into a JSON fixture (data only: per run
the initial guess, the strategy, the printed solution as strings with 6
significant digits, and nni / nfe / nje / nfeD).

Source (read as text from the reference tree):
  examples/kinsol/serial/kinFerTron_dns.out
"""
import json
import os
import re

REF = os.environ.get("SUNDIALS_REFERENCE", "/root/reference")
SRC = os.path.join(REF, "examples", "kinsol", "serial", "kinFerTron_dns.out")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kin_fertron_dns.json")


def main():
    txt = open(SRC).read()
    runs = []
    for gi, block in enumerate(txt.split("Initial guess")[1:]):
        pat = (r"\n(Exact|Modified) Newton( with line search)?\nSolution:\n"
               r"\s*\[x1,x2\] =\s*(\S+)\s+(\S+)\nFinal Statistics:\n"
               r"\s*nni =\s*(\d+)\s+nfe\s*=\s*(\d+)\s*\n\s*nje =\s*(\d+)\s+nfeD =\s*(\d+)")
        for m in re.finditer(pat, block):
            runs.append({"guess": gi + 1, "exact": m.group(1) == "Exact",
                         "linesearch": m.group(2) is not None,
                         "x1": m.group(3), "x2": m.group(4),
                         "nni": int(m.group(5)), "nfe": int(m.group(6)),
                         "nje": int(m.group(7)), "nfeD": int(m.group(8))})
    tol = re.search(r"fnormtol\s*=\s*(\S+)\s*\n\s*scsteptol\s*=\s*(\S+)", txt)
    json.dump({"source": "examples/kinsol/serial/kinFerTron_dns.out",
               "fnormtol": float(tol.group(1)), "scsteptol": float(tol.group(2)),
               "runs": runs}, open(OUT, "w"), indent=1)
    print("wrote %d runs to %s" % (len(runs), OUT))


if __name__ == "__main__":
    main()
