#!/usr/bin/env python3
"""Writes ark_kpr_logging.json: what the reference's ARKStep logging test
prints for the Kvaerno-Prothero-Robinson problem with the IMEX split, Newton
and SPGMR without preconditioning (test/unit_tests/logging/
test_logging_arkode_arkstep_2_1_0.out, driver test_logging_arkode_arkstep.cpp
with arguments "2 1 0"): the (t, u, v) rows after each of the three
ARK_ONE_STEP calls, to the 16 printed digits, and the statistics table of
ARKodePrintAllStats.  Reads the reference's output file (run once; the JSON
is committed)."""
import json
import os
import re

REF = "/root/reference/test/unit_tests/logging/test_logging_arkode_arkstep_2_1_0.out"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ark_kpr_logging.json")


def parse(path):
    rows, stats = [], {}
    for line in open(path):
        f = line.split()
        if len(f) == 5 and all(re.match(r"^-?\d\.\d{15}e[+-]\d\d$", x) for x in f):
            rows.append([float(x) for x in f[:3]])
            continue
        m = re.match(r"^([A-Za-z][A-Za-z \-/]*?)\s+= (\S+)$", line.rstrip())
        if m:
            v = m.group(2)
            stats[m.group(1).strip()] = float(v) if ("." in v or "e" in v) else int(v)
    return rows, stats


if __name__ == "__main__":
    rows, stats = parse(REF)
    assert len(rows) == 4, rows
    json.dump({"source": "test/unit_tests/logging/test_logging_arkode_arkstep_2_1_0.out",
               "rows_t_u_v": rows, "stats": stats}, open(OUT, "w"), indent=1)
    print("wrote", OUT)
