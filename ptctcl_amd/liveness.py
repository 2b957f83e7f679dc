"""Liveness queries over the C ABI: ``p --> q``, ``E[] psi`` and ``A[] not deadlock``.

Mirrors the reference's liveness interface (include/hora/liveness.hpp): ``leads_to_otf(net, p, q,
gs)``, ``exists_invariant_run_otf(net, nq, gs, unrestricted)`` and ``deadlock_reachable(net)``,
with ``GfpStats`` as :class:`LiveStats`. Predicates are the discrete part of ``hora::DiscPred``
(include/hora/predicate.hpp:36-80) in disjunctive normal form; ``E[] psi`` with a clock part runs on
the network restricted by it (the conjunction adjoined to every location invariant), the way the
reference restricts it (include/hora/predicate.hpp:45-58).
"""
from __future__ import annotations

import copy
import ctypes as C
import itertools
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from .network import INF, Network

LEADSTO, EG, DEADLOCK = 0, 1, 2
OP_EQ, OP_NE, OP_LT, OP_LE, OP_GT, OP_GE = range(6)
_NEG = {OP_EQ: OP_NE, OP_NE: OP_EQ, OP_LT: OP_GE, OP_LE: OP_GT, OP_GT: OP_LE, OP_GE: OP_LT}

Atom = Tuple[int, int, int, int]   # (kind 0 location / 1 integer, index, op, value)


@dataclass(frozen=True)
class DPred:
    """A discrete predicate in DNF: true iff some term has all its atoms true."""

    terms: Tuple[Tuple[Atom, ...], ...]

    @staticmethod
    def true() -> "DPred":
        return DPred(((),))

    @staticmethod
    def false() -> "DPred":
        return DPred(())

    @staticmethod
    def at(net: Network, proc: str, loc: str) -> "DPred":
        p, l = net.resolve(proc, loc)
        return DPred((((0, p, OP_EQ, l),),))

    @staticmethod
    def var(net: Network, name: str, op: int, val: int) -> "DPred":
        return DPred((((1, net.intvar_index(name), op, val),),))

    def __and__(self, o: "DPred") -> "DPred":
        return DPred(tuple(a + b for a, b in itertools.product(self.terms, o.terms)))

    def __or__(self, o: "DPred") -> "DPred":
        return DPred(self.terms + o.terms)

    def __invert__(self) -> "DPred":
        out = DPred.true()
        for term in self.terms:
            out = out & DPred(tuple(((k, i, _NEG[op], v),) for k, i, op, v in term))
        return out

    def eval(self, locs: Sequence[int], ints: Sequence[int]) -> bool:
        def holds(a):
            k, i, op, v = a
            x = locs[i] if k == 0 else ints[i]
            return [x == v, x != v, x < v, x <= v, x > v, x >= v][op]
        return any(all(holds(a) for a in t) for t in self.terms)


class hora_dpred(C.Structure):
    _fields_ = [("nterms", C.c_int32), ("term_off", C.c_void_p), ("atoms", C.c_void_p)]


COVER_AUTO, COVER_AM, COVER_ONLINE = 0, 1, 2


class hora_live_query(C.Structure):
    _fields_ = [("kind", C.c_int32), ("early", C.c_int32), ("cover", C.c_int32), ("pad", C.c_int32),
                ("p", hora_dpred), ("q", hora_dpred),
                ("orig_inv_off", C.c_void_p), ("orig_inv_data", C.c_void_p)]


_STAT_I64 = ("outer_nodes", "fwd_edges", "disc_keys", "expanded_nodes", "generated", "covered", "levels",
             "closure", "nonempty", "iters", "pre_calls", "fed_max", "products", "chain_reads",
             "retired", "unions", "enlarged", "fed_hbm_nodes")


class hora_live_stats(C.Structure):
    _fields_ = [(k, C.c_int64) for k in _STAT_I64] + [("early", C.c_int32), ("capped", C.c_int32),
                                                      ("route", C.c_int32), ("seed_keys", C.c_int32)] + \
               [(k, C.c_double) for k in ("seconds", "build_seconds", "gfp_seconds")]


class hora_gfp_graph(C.Structure):
    _fields_ = [("nodes", C.c_int64), ("edges", C.c_int64), ("dk", C.c_void_p), ("zones", C.c_void_p),
                ("row_off", C.c_void_p), ("etgt", C.c_void_p), ("ese", C.c_void_p), ("ere", C.c_void_p),
                ("expanded", C.c_void_p)]


class hora_live_graph(C.Structure):
    _fields_ = [("nodes", C.c_int64), ("edges", C.c_int64), ("d", C.c_int32), ("pad", C.c_int32),
                ("dk", C.c_void_p), ("zones", C.c_void_p), ("row_off", C.c_void_p), ("etgt", C.c_void_p),
                ("ese", C.c_void_p), ("ere", C.c_void_p), ("expanded", C.c_void_p), ("in_s", C.c_void_p),
                ("w_nonempty", C.c_void_p)]


@dataclass
class LiveStats:
    """``GfpStats`` (include/hora/liveness.hpp:48-76) plus construction counters."""

    outer_nodes: int = 0
    fwd_edges: int = 0
    disc_keys: int = 0
    expanded_nodes: int = 0
    generated: int = 0
    covered: int = 0
    levels: int = 0
    closure: int = 0
    nonempty: int = 0
    iters: int = 0
    pre_calls: int = 0
    fed_max: int = 0
    products: int = 0
    chain_reads: int = 0
    retired: int = 0
    unions: int = 0
    enlarged: int = 0
    fed_hbm_nodes: int = 0
    early: int = 0
    capped: int = 0
    route: int = 0
    seed_keys: int = 0
    seconds: float = 0.0
    build_seconds: float = 0.0
    gfp_seconds: float = 0.0

    @classmethod
    def from_c(cls, s: hora_live_stats) -> "LiveStats":
        return cls(**{k: getattr(s, k) for k, _ in s._fields_})


def _pred_arrays(p: DPred):
    off = [0]
    atoms: List[int] = []
    for t in p.terms:
        for a in t:
            atoms.extend(a)
        off.append(len(atoms) // 4)
    return np.asarray(off, np.int32), np.asarray(atoms or [0], np.int32)


@dataclass
class LiveQuery:
    """One liveness query. ``orig`` is the unrestricted network when ``net`` was restricted by a
    clock part of ``psi`` (see :func:`restrict`)."""

    kind: int
    p: DPred = field(default_factory=DPred.true)
    q: DPred = field(default_factory=DPred.false)
    early: bool = True
    orig: Optional[Network] = None
    cover: int = COVER_AUTO

    def encode(self) -> "LiveQueryEnc":
        return LiveQueryEnc(self)


class LiveQueryEnc:
    def __init__(self, q: LiveQuery):
        self.keep = []
        s = hora_live_query()
        s.kind = q.kind
        s.early = 1 if q.early else 0
        s.cover = q.cover
        for name, pred in (("p", q.p), ("q", q.q)):
            off, atoms = _pred_arrays(pred)
            self.keep += [off, atoms]
            dp = getattr(s, name)
            dp.nterms = len(pred.terms)
            dp.term_off = off.ctypes.data
            dp.atoms = atoms.ctypes.data
        if q.orig is not None:
            enc = q.orig.encode()
            self.keep.append(enc)
            s.orig_inv_off = enc.arrays["inv_off"].ctypes.data
            s.orig_inv_data = enc.arrays["inv_data"].ctypes.data
        self.struct = s


def base_route(net: Network) -> int:
    """Where the reference may decide ``p --> q`` by its base construction instead of the online one
    (structurally_empty_d, src/liveness/internal.cpp:59-103): 1 for a time-open network (no
    urgency, no upper invariant; always the base construction), 2 for a network with one possible
    discrete state (the base construction unless that state can seed a deadlock atom), else 0. Same
    test as the library's at open; COVER_ONLINE runs the online construction there instead."""
    time_open = not any(c.urgent for c in net.channels)
    for p in net.procs:
        for loc in p.locations:
            if loc.urgent or loc.committed:
                time_open = False
            if any(i > 0 and j == 0 and b < INF for i, j, b in loc.invariant):
                time_open = False
    if time_open:
        return 1
    one_key = all(len(p.locations) == 1 for p in net.procs) and all(v.lo == v.hi for v in net.intvars)
    return 2 if one_key else 0


def leads_to(p: DPred, q: DPred, early: bool = True, cover: int = COVER_AUTO) -> LiveQuery:
    return LiveQuery(LEADSTO, p, q, early, cover=cover)


def not_deadlock(early: bool = True, cover: int = COVER_AUTO) -> LiveQuery:
    return LiveQuery(DEADLOCK, early=early, cover=cover)


def restrict(net: Network, clocks: Sequence[Tuple[int, int, int]]) -> Network:
    """The network with the conjunction `clocks` adjoined to every location invariant, analyses
    recomputed (include/hora/predicate.hpp:45-58)."""
    r = copy.deepcopy(net)
    for p in r.procs:
        for l in p.locations:
            l.invariant = list(l.invariant) + list(clocks)
    r._analysed = False
    return r.finalize()


def exists_always(net: Network, psi: DPred, clocks: Sequence[Tuple[int, int, int]] = ()):
    """``E[] psi`` with psi = discrete part `psi` /\\ conjunction `clocks`: returns the network to
    open and the query."""
    if not clocks:
        return net, LiveQuery(EG, psi)
    return restrict(net, clocks), LiveQuery(EG, psi, orig=net)


def graph_arrays(nodes: int, edges: int, d: int) -> dict:
    return {
        "dk": np.zeros(max(nodes, 1), np.uint64),
        "zones": np.zeros(max(nodes, 1) * d * d, np.int32),
        "row_off": np.zeros(nodes + 1, np.int64),
        "etgt": np.zeros(max(edges, 1), np.int32),
        "ese": np.zeros(max(edges, 1), np.int32),
        "ere": np.zeros(max(edges, 1), np.int32),
        "expanded": np.zeros(max(nodes, 1), np.int8),
        "in_s": np.zeros(max(nodes, 1), np.int8),
        "w_nonempty": np.zeros(max(nodes, 1), np.int8),
    }


def fill_graph_struct(g: hora_live_graph, arrs: dict) -> None:
    for k, a in arrs.items():
        setattr(g, k, a.ctypes.data)


def trim_graph(arrs: dict, nodes: int, edges: int, d: int) -> dict:
    out = dict(arrs)
    for k in ("dk", "expanded", "in_s", "w_nonempty"):
        out[k] = arrs[k][:nodes]
    out["zones"] = arrs["zones"][: nodes * d * d].reshape(nodes, d * d)
    for k in ("etgt", "ese", "ere"):
        out[k] = arrs[k][:edges]
    return out
