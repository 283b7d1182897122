"""Extract the published outputs of the reference example cvRoberts_dns into a
JSON fixture (data only: output times, printed solution values and the
per-return integrator counters).

Sources (read as text from the reference tree):
  examples/cvode/serial/cvRoberts_dns.out        printed y at each output time
                                                  and the final statistics
  examples/cvode/serial/cvRoberts_dns_stats.csv  counters at every CVode return
                                                  (16-digit step sizes)
Root returns are kept but flagged; the oracle driver runs without rootfinding,
which leaves the step sequence unchanged, and is compared on normal returns.
"""
import json
import os
import re
import sys

REF = os.environ.get("SUNDIALS_REFERENCE", "/root/reference")
EX = os.path.join(REF, "examples", "cvode", "serial")
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cvRoberts_dns.json")


def parse_out(path):
    returns, final = [], {}
    lines = open(path).read().splitlines()
    for i, ln in enumerate(lines):
        m = re.match(r"\s*At t = (\S+)\s+y =\s+(\S+)\s+(\S+)\s+(\S+)", ln)
        if m:
            is_root = i + 1 < len(lines) and "rootsfound" in lines[i + 1]
            returns.append({"t": float(m.group(1)),
                            "y": [float(m.group(k)) for k in (2, 3, 4)],
                            "root": is_root})
        m = re.match(r"\s*([A-Za-z][A-Za-z .]+?)\s+=\s+(\S+)\s*$", ln)
        if m and returns:
            final[m.group(1).strip()] = float(m.group(2))
    return returns, final


def final_table_lines(path):
    lines = open(path).read().splitlines()
    i = lines.index("Final Statistics:")
    out = []
    for ln in lines[i + 1:]:
        if " = " not in ln:
            break
        out.append(ln)
    return out


def parse_csv(path):
    rows = []
    for ln in open(path).read().splitlines():
        f = ln.split(",")
        if len(f) < 2:
            continue
        rows.append({f[k]: float(f[k + 1]) for k in range(0, len(f) - 1, 2)})
    return rows


def main():
    returns, final = parse_out(os.path.join(EX, "cvRoberts_dns.out"))
    stats = parse_csv(os.path.join(EX, "cvRoberts_dns_stats.csv"))
    if len(stats) != len(returns):
        sys.exit("return count mismatch between .out and _stats.csv")
    for r, s in zip(returns, stats):
        r["stats"] = s
    doc = {
        "source": "examples/cvode/serial/cvRoberts_dns.out, cvRoberts_dns_stats.csv",
        "setup": {"y0": [1.0, 0.0, 0.0], "rtol": 1e-4, "atol": [1e-8, 1e-14, 1e-6],
                  "t1": 0.4, "tmult": 10.0, "nout": 12},
        "returns": returns,
        "final": final,
        # the final statistics table as printed (CVodePrintAllStats, TABLE)
        "final_table_lines": final_table_lines(os.path.join(EX, "cvRoberts_dns.out")),
    }
    json.dump(doc, open(OUT, "w"), indent=1)
    print("wrote", OUT, len(returns), "returns")


if __name__ == "__main__":
    main()
