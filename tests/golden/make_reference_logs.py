"""Extract the reference's logged join results into reference_logs.json.

Reads the reference's run logs (text) and workload config files (text) under
/root/reference and writes one record per distinct (workload, algo, result):
matched_cnt, check_sum, relation sizes and the log file:line it came from.
These are the only golden vectors the reference publishes (it has no tests).

Run in the build container only:  python tests/golden/make_reference_logs.py
"""
import json
import os
import re
import sys

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_logs.json")
ANSI = re.compile(r"\x1b\[[0-9;]*m")


def configs():
    """workload_name -> (r_tuple_num, s_tuple_num, r_tuple_size, theta, cfg path)"""
    out = {}
    for root, _, files in os.walk(os.path.join(REF, "config")):
        for f in files:
            if not f.endswith(".cfg") or f == "cpu.cfg":
                continue
            p = os.path.join(root, f)
            txt = open(p).read()
            kv = dict(re.findall(r"^\s*(\w+)\s*=\s*\"?([^\";]*)\"?;", txt, re.M))
            if "workload_name" not in kv:
                continue
            rel = os.path.relpath(p, REF)
            out.setdefault(kv["workload_name"], []).append(
                {"r_tuple_num": int(kv["r_tuple_num"]), "s_tuple_num": int(kv["s_tuple_num"]),
                 "tuple_size": int(kv.get("r_tuple_size", 16)), "theta": float(kv.get("theta", 0)),
                 "cfg": rel})
    return out


def main():
    if not os.path.isdir(REF):
        sys.exit("reference tree not present")
    cfg = configs()
    recs = {}
    for root, _, files in os.walk(os.path.join(REF, "logs")):
        for f in sorted(files):
            if not f.endswith(".log"):
                continue
            p = os.path.join(root, f)
            lines = [ANSI.sub("", l) for l in open(p, errors="replace").read().splitlines()]
            for i, l in enumerate(lines):
                m = re.search(r"matched_cnt: (\d+)\s+check_sum: (\d+)", l)
                if not m or i + 1 >= len(lines):
                    continue
                w = re.search(r"workload: (\S+)", lines[i + 1])
                a = re.search(r"algo: (\S+)", lines[i + 1])
                if not w or not a:
                    continue
                key = (w.group(1), a.group(1), int(m.group(1)), int(m.group(2)))
                if key in recs:
                    continue
                c = cfg.get(w.group(1), [{}])
                # a workload name can exist in both config trees; keep every candidate
                recs[key] = {"workload": w.group(1), "algo": a.group(1), "matched_cnt": int(m.group(1)),
                             "check_sum": int(m.group(2)), "source": f"{os.path.relpath(p, REF)}:{i + 1}",
                             "configs": c}
    data = sorted(recs.values(), key=lambda r: (r["workload"], r["algo"], r["source"]))
    with open(OUT, "w") as fh:
        json.dump(data, fh, indent=1)
    print(f"wrote {len(data)} records to {OUT}")


if __name__ == "__main__":
    main()
