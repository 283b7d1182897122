"""Extract the published outputs of the reference example kinAnalytic_fp
(Anderson-accelerated fixed-point iteration) into a JSON fixture (data only:
the options each run printed, its iteration and function-evaluation counts,
and the printed solution and error values as strings, 6 significant digits).

Sources (read as text from the reference tree):
  examples/kinsol/serial/kinAnalytic_fp*.out
"""
import glob
import json
import os
import re

REF = os.environ.get("SUNDIALS_REFERENCE", "/root/reference")
EX = os.path.join(REF, "examples", "kinsol", "serial")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "kin_analytic_fp.json")


def parse(path):
    txt = open(path).read()
    def num(key):
        return re.search(r"\b%s\s*=\s*(\S+)" % re.escape(key), txt).group(1)
    run = {"file": os.path.basename(path),
           "tol": float(num("tolerance")), "maxiter": int(num("max iters")),
           "m_aa": int(num("m_aa")), "delay_aa": int(num("delay_aa")),
           "damping_aa": float(num("damping_aa")), "damping_fp": float(num("damping_fp")),
           "damping_fn": num("damping_fn") == "ON", "depth_fn": num("depth_fn") == "ON",
           "orth_aa": int(num("orth routine"))}
    run["nni"] = int(re.search(r"Number of nonlinear iterations:\s*(\d+)", txt).group(1))
    run["nfe"] = int(re.search(r"Number of function evaluations:\s*(\d+)", txt).group(1))
    sol = txt.split("Computed solution:")[1]
    run["x"], run["y"], run["z"] = [re.search(r"\b%s = (\S+)" % c, sol).group(1)
                                    for c in ("x", "y", "z")]
    err = txt.split("Solution error:")[1]
    run["ex"], run["ey"], run["ez"] = [re.search(r"\b%s = (\S+)" % c, err).group(1)
                                       for c in ("ex", "ey", "ez")]
    run["pass"] = "PASS" in err
    return run


def main():
    runs = [parse(p) for p in sorted(glob.glob(os.path.join(EX, "kinAnalytic_fp*.out")))]
    json.dump({"source": "examples/kinsol/serial/kinAnalytic_fp*.out", "runs": runs},
              open(OUT, "w"), indent=1)
    print("wrote %d runs to %s" % (len(runs), OUT))


if __name__ == "__main__":
    main()
