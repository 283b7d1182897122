#!/usr/bin/env python3
"""Writes ark_dahlquist_ark324.json: the per-step counts the reference's
ARKStep Dahlquist unit test prints for the IMEX pairs ARK3(2)4L[2]SA (keys
"hermite", "none") and ARK4(3)7L[2]SA (keys "ark437_hermite", "ark437_none")
(test/unit_tests/arkode/CXX_serial/ark_test_dahlquist_ark_0_{0,-1}_0.out:
identity mass matrix, Hermite or no interpolant, trivial predictor; the test
sets a fixed step 0.01, lambda_e = lambda_i = -1, a linearly implicit
autonomous problem).  For each step: cumulative Newton iterations and fe /
fi evaluations ("actual" values).  Reads the reference's output files (run
once; the JSON is committed)."""
import json
import os
import re

REF = "/root/reference/test/unit_tests/arkode/CXX_serial"


def section(path, dirk="ARKODE_ARK324L2SA_DIRK_4_2_3", erk="ARKODE_ARK324L2SA_ERK_4_2_3"):
    lines = open(path).read().split("\n")
    start = next(i for i, l in enumerate(lines)
                 if l.strip() == "DIRK: " + dirk and lines[i + 1].strip() == "ERK: " + erk)
    steps, cur, which = [], None, None
    for l in lines[start + 2:]:
        if l.startswith("====") and cur is not None and steps:
            break
        m = re.match(r"Steps: (\d+)", l)
        if m:
            cur = {"step": int(m.group(1))}
            steps.append(cur)
            continue
        if l.startswith("Dense Output"):
            cur = None
            continue
        if cur is None:
            continue
        m = re.match(r"NLS iters: (\d+)", l)
        if m:
            cur["nni"] = int(m.group(1))
        if l.startswith("Fe RHS"):
            which = "nfe"
        elif l.startswith("Fi RHS"):
            which = "nfi"
        m = re.match(r"\s+actual:\s+(\d+)", l)
        if m and which:
            cur[which] = int(m.group(1))
    return steps


def main():
    out = {"source": "test/unit_tests/arkode/CXX_serial/ark_test_dahlquist_ark_0_{0,-1}_0.out",
           "h": 0.01, "lambda_e": -1.0, "lambda_i": -1.0, "rtol": 1e-4, "atol": 1e-6}
    for tag, fname in (("hermite", "ark_test_dahlquist_ark_0_0_0.out"),
                       ("none", "ark_test_dahlquist_ark_0_-1_0.out")):
        out[tag] = section(os.path.join(REF, fname))
        out["ark437_" + tag] = section(os.path.join(REF, fname), "ARKODE_ARK437L2SA_DIRK_7_3_4",
                                       "ARKODE_ARK437L2SA_ERK_7_3_4")
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "ark_dahlquist_ark324.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
