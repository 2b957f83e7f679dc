"""Writes tests/golden/live_recorded_counts.json from the reference's recorded liveness tables (run
where the reference is).

benchmarks/results/live/hora/hora_live_t1.csv and benchmarks/results/live/hora_backward/
hora_live_t1.csv hold one row per (model, leads-to query, repetition): the verdict and the `states`
column. Both tables record the same construction (the backward table's build replaces only the
verdict's final step, `-DHORA_LIVENESS_GFP`), and they agree on every row but hddi's, where the
forward table's numbers are the ones the construction reproduces. The fixture keeps, per model and
query, the verdict and the `states` of the forward table with its line number, and the backward
table's number beside it.

Kept: the families whose networks the tests can build (fischer from the generator, which reads to
the same network as the reference's fischer<N>.xml; the others from tests/golden/models/), and whose
rows are decided on the construction the device runs: every family here but csma and grc, whose
discrete states seed deadlock atoms (their rows need the map D the device does not build).
"""
import csv
import json
import os

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
FWD = "benchmarks/results/live/hora/hora_live_t1.csv"
BWD = "benchmarks/results/live/hora_backward/hora_live_t1.csv"
FAMILIES = ("fischer", "arbiter", "phil", "twolock", "single", "hddi")


def rows(rel):
    out = {}
    with open(os.path.join(REF, rel)) as f:
        for line, r in enumerate(csv.DictReader(f), start=2):
            if r["rep"] != "1" or r["status"] != "OK":
                continue
            out[(r["model"], r["query"])] = (line, int(r["states"]), r["verdict"])
    return out


def main():
    fwd, bwd = rows(FWD), rows(BWD)
    models = sorted(os.listdir(os.path.join(HERE, "models")))
    out = []
    for (model, query), (line, states, verdict) in sorted(fwd.items()):
        fam = model.rstrip("0123456789")
        if fam not in FAMILIES:
            continue
        if fam != "fischer" and model + ".xml" not in models:
            continue
        b = bwd.get((model, query))
        out.append({"model": model, "query": query, "holds": verdict == "SAT", "states": states,
                    "source": f"{FWD}:{line}",
                    "backward_states": b[1] if b else None,
                    "backward_source": f"{BWD}:{b[0]}" if b else None})
    with open(os.path.join(HERE, "live_recorded_counts.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(len(out), "rows")


if __name__ == "__main__":
    main()
