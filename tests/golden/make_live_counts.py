"""Writes tests/golden/live_counts.json: node, edge and closure counts of the oracle's liveness
graph for fischer `P0.req --> P0.wait` (full construction, no early stop) at sizes the device
tests check without re-running the oracle, for the automatic covering route (a_LU here: fischer
seeds no deadlock atom) and for the a_M graph. Run: python tests/golden/make_live_counts.py"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import binding  # noqa: E402
from ptctcl_amd import liveness as lv, models  # noqa: E402


def main():
    out = {"fischer_req_wait": {}, "fischer_req_wait_am": {}}
    for name, cover, sizes in (("fischer_req_wait", lv.COVER_AUTO, (6, 7, 8, 9, 10)),
                               ("fischer_req_wait_am", lv.COVER_AM, (6, 7, 8))):
        for n in sizes:
            net = models.fischer(n)
            q = lv.leads_to(lv.DPred.at(net, "P0", "req"), lv.DPred.at(net, "P0", "wait"), early=False, cover=cover)
            r, st = binding.live(net, q)
            out[name][str(n)] = {"holds": r, "nodes": st.outer_nodes, "edges": st.fwd_edges,
                                 "closure": st.closure, "nonempty": st.nonempty, "route": st.route}
            print(name, n, out[name][str(n)], flush=True)
    with open(os.path.join(HERE, "live_counts.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
