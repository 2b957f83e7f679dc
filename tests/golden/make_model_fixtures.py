"""Writes the UPPAAL model fixtures and their recorded state counts (run where the reference is).

Copies a selection of the reference's benchmark models (benchmarks/models/xml/*.xml, data files
the reference's benchmark generators wrote) unchanged into tests/golden/models/, and writes
xml_reach_counts.json: for the families whose counts do not depend on the subsumption relation
(a_LU and the G-simulation give the same numbers there), the recorded `states` of each model with
the CSV line it comes from, where every recorded run of it (16, 24 and 32 threads, two
repetitions each) agrees; where they differ (the search order changes the count), the range.

grc, single and hddi are copied too, without counts: their recorded runs used the G-simulation
(`--sim g`, benchmarks/scripts/bench.py:40), which the reference's GPU path does not run
(src/reachability/engine.cpp:1875), and grc's counts differ from run to run. The tests check
those models, and the order-sensitive sizes, against the oracle instead.
"""
import csv
import json
import os
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
CSV = "benchmarks/results/reach/hora/hora_t16.csv"
OTHER = ("benchmarks/results/reach/hora/hora_t24.csv", "benchmarks/results/reach/hora/hora_t32.csv")
MODELS = {
    "arbiter": [4, 8, 12, 16, 18],
    "phil": [4, 6, 8, 10, 12],
    "twolock": [4, 6, 8, 10, 12, 13],
    "csma": [4, 6, 8, 10, 14, 16],
    "fischer": [4, 8, 12],
    "grc": [4, 5],
    "single": [4, 5, 6],
    "hddi": [2, 10],
    "viking": [4, 6, 10, 12, 15, 17, 20],
}
# viking's counts are also verifyta's and tchecker's under a_LU (reach/uppaal_dfs/verifyta_dfs_t1.csv:
# 141-152, reach/tchecker_cov/tchecker_cov_t1.csv:143-150)
COUNTED = ("arbiter", "phil", "twolock", "csma", "fischer", "viking")


def main():
    out_dir = os.path.join(HERE, "models")
    os.makedirs(out_dir, exist_ok=True)
    rows, seen = {}, {}
    for path in (CSV,) + OTHER:
        with open(os.path.join(REF, path)) as f:
            for line_no, r in enumerate(csv.DictReader(f), start=2):
                if r["status"] != "OK":
                    continue
                seen.setdefault(r["model"], set()).add(int(r["states"]))
                if path == CSV and r["model"] not in rows:
                    rows[r["model"]] = (int(r["states"]), r["query"], r["verdict"], line_no)
    counts = {"_source": "States recorded by the reference's reach engine for its benchmark models "
                         f"({CSV}, 16 threads, first OK row; 'range' where the runs in it and in "
                         f"{', '.join(OTHER)} differ); every copied model is SAT, so the count is the "
                         "whole zone graph.",
              "models": {}}
    for fam, sizes in MODELS.items():
        for n in sizes:
            name = f"{fam}{n}"
            shutil.copyfile(os.path.join(REF, "benchmarks/models/xml", name + ".xml"),
                            os.path.join(out_dir, name + ".xml"))
            states, query, verdict, line_no = rows[name]
            entry = {"query": query, "verdict_satisfied": verdict == "SAT"}
            if fam in COUNTED and len(seen[name]) == 1:
                entry.update(states=states, cite=f"{CSV}:{line_no}")
            elif fam in COUNTED:
                entry.update(range=[min(seen[name]), max(seen[name])])
            counts["models"][name] = entry
    with open(os.path.join(HERE, "xml_reach_counts.json"), "w") as f:
        json.dump(counts, f, indent=1)
        f.write("\n")


if __name__ == "__main__":
    main()
