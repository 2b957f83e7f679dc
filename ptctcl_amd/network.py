"""Network IR of a network of timed automata, its static analyses and its flat C encoding.

Mirrors ``hora::Network`` / ``NProc`` / ``NEdge`` / ``Location`` (include/hora/network.hpp:25-140,
include/hora/automaton.hpp:17-30) for the subset the device path covers: clock guards, resets and
invariants as ``x_i - x_j <| c`` triples, ``var OP const`` integer guards, ``var := const``
updates, binary and broadcast channels, urgent and committed locations, urgent channels.

The analyses restate ``Network::compute_bounds``, ``compute_loc_bounds`` and ``compute_active``
(src/model/network.cpp:59-245); ``encode()`` packs everything into :class:`NetEnc`, the
``hora_net_enc`` struct of include/hora_hip.h (the reference's ``netgpu::encode``,
include/hora/gpu/encode.hpp:54-124).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

INF = 1 << 30
ALU_NEG_INF = -(1 << 29)
LE0 = 1

OP_EQ, OP_NE, OP_LT, OP_LE, OP_GT, OP_GE = range(6)
_OPS = {"==": OP_EQ, "!=": OP_NE, "<": OP_LT, "<=": OP_LE, ">": OP_GT, ">=": OP_GE}


def le(v: int) -> int:
    """Raw bound ``<= v`` (HDBM encoding: value << 1 | 1 for a weak bound)."""
    return v * 2 + 1


def lt(v: int) -> int:
    """Raw bound ``< v``."""
    return v * 2


def bound(r: int) -> int:
    return r >> 1


def is_strict(r: int) -> bool:
    return (r & 1) == 0


# Constraint constructors over clock indices (clock 0 is the reference clock), as HDBM's
# c_le / c_lt / c_ge / c_gt build them (include/hora/automaton.hpp:14).
def c_le(i: int, v: int) -> Tuple[int, int, int]:
    return (i, 0, le(v))


def c_lt(i: int, v: int) -> Tuple[int, int, int]:
    return (i, 0, lt(v))


def c_ge(i: int, v: int) -> Tuple[int, int, int]:
    return (0, i, le(-v))


def c_gt(i: int, v: int) -> Tuple[int, int, int]:
    return (0, i, lt(-v))


@dataclass
class Location:
    name: str
    invariant: List[Tuple[int, int, int]] = field(default_factory=list)
    urgent: bool = False
    committed: bool = False


@dataclass
class Edge:
    src: int
    tgt: int
    guard: List[Tuple[int, int, int]] = field(default_factory=list)
    resets: List[int] = field(default_factory=list)
    int_guard: List[Tuple[int, int, int]] = field(default_factory=list)  # (var, op, val)
    int_updates: List[Tuple[int, int]] = field(default_factory=list)  # (var, val)
    chan: int = -1
    dir: int = 0  # +1 send, -1 receive


@dataclass
class Process:
    name: str
    locations: List[Location]
    edges: List[Edge]
    initial: int = 0

    def loc_index(self, name: str) -> int:
        for i, l in enumerate(self.locations):
            if l.name == name:
                return i
        raise KeyError(f"{self.name}.{name}")


@dataclass
class IntVar:
    name: str
    lo: int
    hi: int
    init: int = 0


@dataclass
class Channel:
    name: str
    broadcast: bool = False
    urgent: bool = False


@dataclass
class Target:
    """E<> target: a conjunction (``hora::NetTarget``, include/hora/network.hpp:150-163)."""

    loc_req: List[Tuple[int, int]] = field(default_factory=list)
    int_cond: List[Tuple[int, int, int]] = field(default_factory=list)
    clock_cond: List[Tuple[int, int, int]] = field(default_factory=list)
    never: bool = False


class Network:
    def __init__(self, nclocks: int, intvars: Sequence[IntVar] = (), channels: Sequence[Channel] = (),
                 procs: Sequence[Process] = (), clock_names: Optional[Sequence[str]] = None, name: str = ""):
        self.name = name
        self.nclocks = nclocks
        self.clock_names = list(clock_names) if clock_names else ["0"] + [f"x{i}" for i in range(1, nclocks + 1)]
        self.intvars = list(intvars)
        self.channels = list(channels)
        self.procs = list(procs)
        self._analysed = False

    @property
    def dim(self) -> int:
        return self.nclocks + 1

    def nprocs(self) -> int:
        return len(self.procs)

    def resolve(self, proc: str, loc: str) -> Tuple[int, int]:
        for pi, p in enumerate(self.procs):
            if p.name == proc:
                return pi, p.loc_index(loc)
        raise KeyError(proc)

    def intvar_index(self, name: str) -> int:
        for i, v in enumerate(self.intvars):
            if v.name == name:
                return i
        raise KeyError(name)

    def has_diagonal(self) -> bool:
        """``network_has_diagonal`` (src/reachability/engine.cpp:190-200)."""
        for p in self.procs:
            for l in p.locations:
                if any(i > 0 and j > 0 for i, j, _ in l.invariant):
                    return True
            for e in p.edges:
                if any(i > 0 and j > 0 for i, j, _ in e.guard):
                    return True
        return False

    # ---- analyses (src/model/network.cpp:59-245) ---------------------------------------------
    def _edges_from(self, p: Process, l: int) -> List[Edge]:
        return [e for e in p.edges if e.src == l]

    def compute_bounds(self) -> None:
        d = self.dim
        L = [0] * d
        U = [0] * d

        def scan(cs):
            for i, j, r in cs:
                v = bound(r)
                if i > 0 and j == 0:
                    U[i] = max(U[i], v)
                elif i == 0 and j > 0:
                    L[j] = max(L[j], -v)
                elif i > 0 and j > 0:
                    a = abs(v)
                    U[i] = max(U[i], a)
                    U[j] = max(U[j], a)
                    L[i] = max(L[i], a)
                    L[j] = max(L[j], a)

        for p in self.procs:
            for l in p.locations:
                scan(l.invariant)
            for e in p.edges:
                scan(e.guard)
        L[0] = U[0] = 0
        self.lbound, self.ubound = L, U

    def compute_loc_bounds(self) -> None:
        d = self.dim
        self.lbound_loc, self.ubound_loc = [], []

        def scan(cs, Lc, Uc):
            for i, j, r in cs:
                v = bound(r)
                if i > 0 and j == 0:
                    Uc[i] = max(Uc[i], v)
                elif i == 0 and j > 0:
                    Lc[j] = max(Lc[j], -v)
                elif i > 0 and j > 0:
                    a = abs(v)
                    Uc[i] = max(Uc[i], a)
                    Uc[j] = max(Uc[j], a)
                    Lc[i] = max(Lc[i], a)
                    Lc[j] = max(Lc[j], a)

        for p in self.procs:
            nl = len(p.locations)
            lb = [[ALU_NEG_INF] * d for _ in range(nl)]
            ub = [[ALU_NEG_INF] * d for _ in range(nl)]
            changed = True
            while changed:
                changed = False
                for l in range(nl):
                    Lc = [ALU_NEG_INF] * d
                    Uc = [ALU_NEG_INF] * d
                    scan(p.locations[l].invariant, Lc, Uc)
                    for e in self._edges_from(p, l):
                        scan(e.guard, Lc, Uc)
                        rst = set(e.resets)
                        for c in range(1, d):
                            if c not in rst:
                                Lc[c] = max(Lc[c], lb[e.tgt][c])
                                Uc[c] = max(Uc[c], ub[e.tgt][c])
                    for c in range(1, d):
                        if Lc[c] > lb[l][c]:
                            lb[l][c] = Lc[c]
                            changed = True
                        if Uc[c] > ub[l][c]:
                            ub[l][c] = Uc[c]
                            changed = True
            self.lbound_loc.append(lb)
            self.ubound_loc.append(ub)

    def compute_active(self) -> None:
        d = self.dim
        self.active_mask = []
        for p in self.procs:
            nl = len(p.locations)
            act = [set() for _ in range(nl)]
            changed = True
            while changed:
                changed = False
                for l in range(nl):
                    A = set()
                    for i, j, _ in p.locations[l].invariant:
                        if i > 0:
                            A.add(i)
                        if j > 0:
                            A.add(j)
                    for e in self._edges_from(p, l):
                        for i, j, _ in e.guard:
                            if i > 0:
                                A.add(i)
                            if j > 0:
                                A.add(j)
                        rst = set(e.resets)
                        A |= {c for c in act[e.tgt] if c not in rst}
                    if not A <= act[l]:
                        act[l] |= A
                        changed = True
            masks = []
            for l in range(nl):
                m = 0
                for c in act[l]:
                    if 1 <= c < d:
                        m |= 1 << c
                masks.append(m)
            self.active_mask.append(masks)

    def finalize(self) -> "Network":
        self.compute_bounds()
        self.compute_loc_bounds()
        self.compute_active()
        self._analysed = True
        return self

    # ---- flat encoding --------------------------------------------------------------------
    def encode(self) -> "NetEnc":
        if not self._analysed:
            self.finalize()
        return NetEnc(self)

    def key_bits(self) -> int:
        def fb(n):
            b = 0
            while (1 << b) < n:
                b += 1
            return b
        return sum(fb(len(p.locations)) for p in self.procs) + sum(fb(v.hi - v.lo + 1) for v in self.intvars)


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(-1))


class hora_net_enc(C.Structure):
    _fields_ = [
        ("nclocks", C.c_int32), ("nprocs", C.c_int32), ("nintvars", C.c_int32), ("nchans", C.c_int32),
        ("nedges", C.c_int32), ("total_locs", C.c_int32),
        ("int_lo", C.c_void_p), ("int_hi", C.c_void_p), ("int_init", C.c_void_p),
        ("chan_bcast", C.c_void_p), ("chan_urgent", C.c_void_p),
        ("locbase", C.c_void_p), ("loc_init", C.c_void_p), ("loc_flags", C.c_void_p),
        ("inv_off", C.c_void_p), ("inv_data", C.c_void_p),
        ("edge_proc", C.c_void_p), ("edge_src", C.c_void_p), ("edge_tgt", C.c_void_p),
        ("edge_chan", C.c_void_p), ("edge_dir", C.c_void_p),
        ("cg_off", C.c_void_p), ("cg_data", C.c_void_p),
        ("rs_off", C.c_void_p), ("rs_clk", C.c_void_p),
        ("ig_off", C.c_void_p), ("ig_data", C.c_void_p),
        ("iu_off", C.c_void_p), ("iu_data", C.c_void_p),
        ("lbound", C.c_void_p), ("ubound", C.c_void_p),
        ("lbound_loc", C.c_void_p), ("ubound_loc", C.c_void_p),
        ("active_mask", C.c_void_p),
    ]


class hora_target(C.Structure):
    _fields_ = [
        ("never", C.c_int32), ("nloc_req", C.c_int32), ("loc_req", C.c_void_p),
        ("nint_cond", C.c_int32), ("int_cond", C.c_void_p),
        ("nclock_cond", C.c_int32), ("clock_cond", C.c_void_p),
    ]


class hora_reach_stats(C.Structure):
    _fields_ = [
        ("explored", C.c_uint64), ("generated", C.c_uint64), ("expanded", C.c_uint64),
        ("stored_live", C.c_uint64), ("keys", C.c_uint64), ("inserted", C.c_uint64),
        ("subsumed", C.c_uint64), ("killed", C.c_uint64), ("unions", C.c_uint64),
        ("chain_reads", C.c_uint64),
        ("levels", C.c_int32), ("found", C.c_int32), ("emode_switch_level", C.c_int32),
        ("zone_entry_bytes", C.c_int32),
        ("seconds", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class NetEnc:
    """Owns the numpy arrays behind one ``hora_net_enc`` (keeps them alive)."""

    def __init__(self, net: Network):
        d = net.dim
        arrs = {}
        locbase = [0]
        for p in net.procs:
            locbase.append(locbase[-1] + len(p.locations))
        tl = locbase[-1]
        inv_off, inv_data, flags = [0], [], []
        for p in net.procs:
            for l in p.locations:
                for c in l.invariant:
                    inv_data.extend(c)
                inv_off.append(len(inv_data) // 3)
                flags.append((1 if l.committed else 0) | (2 if l.urgent else 0))
        eproc, esrc, etgt, echan, edir = [], [], [], [], []
        cg_off, cg, rs_off, rs, ig_off, ig, iu_off, iu = [0], [], [0], [], [0], [], [0], []
        for pi, p in enumerate(net.procs):
            for e in p.edges:
                eproc.append(pi); esrc.append(e.src); etgt.append(e.tgt); echan.append(e.chan); edir.append(e.dir)
                for c in e.guard:
                    cg.extend(c)
                cg_off.append(len(cg) // 3)
                rs.extend(e.resets)
                rs_off.append(len(rs))
                for a in e.int_guard:
                    ig.extend(a)
                ig_off.append(len(ig) // 3)
                for a in e.int_updates:
                    iu.extend(a)
                iu_off.append(len(iu) // 2)
        lloc = np.zeros((tl, d), dtype=np.int32)
        uloc = np.zeros((tl, d), dtype=np.int32)
        amask = np.zeros(tl, dtype=np.uint64)
        for pi, p in enumerate(net.procs):
            for l in range(len(p.locations)):
                g = locbase[pi] + l
                lloc[g] = net.lbound_loc[pi][l]
                uloc[g] = net.ubound_loc[pi][l]
                amask[g] = net.active_mask[pi][l]
        arrs["int_lo"] = _i32([v.lo for v in net.intvars] or [0])
        arrs["int_hi"] = _i32([v.hi for v in net.intvars] or [0])
        arrs["int_init"] = _i32([v.init for v in net.intvars] or [0])
        arrs["chan_bcast"] = np.asarray([1 if c.broadcast else 0 for c in net.channels] or [0], dtype=np.uint8)
        arrs["chan_urgent"] = np.asarray([1 if c.urgent else 0 for c in net.channels] or [0], dtype=np.uint8)
        arrs["locbase"] = _i32(locbase)
        arrs["loc_init"] = _i32([p.initial for p in net.procs])
        arrs["loc_flags"] = _i32(flags)
        arrs["inv_off"] = _i32(inv_off)
        arrs["inv_data"] = _i32(inv_data or [0])
        for k, v in (("edge_proc", eproc), ("edge_src", esrc), ("edge_tgt", etgt), ("edge_chan", echan),
                     ("edge_dir", edir)):
            arrs[k] = _i32(v or [0])
        arrs["cg_off"] = _i32(cg_off); arrs["cg_data"] = _i32(cg or [0])
        arrs["rs_off"] = _i32(rs_off); arrs["rs_clk"] = _i32(rs or [0])
        arrs["ig_off"] = _i32(ig_off); arrs["ig_data"] = _i32(ig or [0])
        arrs["iu_off"] = _i32(iu_off); arrs["iu_data"] = _i32(iu or [0])
        arrs["lbound"] = _i32(net.lbound); arrs["ubound"] = _i32(net.ubound)
        arrs["lbound_loc"] = np.ascontiguousarray(lloc.reshape(-1))
        arrs["ubound_loc"] = np.ascontiguousarray(uloc.reshape(-1))
        arrs["active_mask"] = np.ascontiguousarray(amask)
        self.arrays = arrs
        s = hora_net_enc()
        s.nclocks = net.nclocks
        s.nprocs = len(net.procs)
        s.nintvars = len(net.intvars)
        s.nchans = len(net.channels)
        s.nedges = len(eproc)
        s.total_locs = tl
        for k, a in arrs.items():
            setattr(s, k, a.ctypes.data)
        self.struct = s
        self.net = net

    def ptr(self):
        return C.byref(self.struct)


class TargetEnc:
    def __init__(self, t: Optional[Target]):
        t = t or Target(never=True)
        self.lr = _i32([x for pr in t.loc_req for x in pr] or [0])
        self.ic = _i32([x for a in t.int_cond for x in a] or [0])
        self.cc = _i32([x for c in t.clock_cond for x in c] or [0])
        s = hora_target()
        s.never = 1 if t.never else 0
        s.nloc_req = len(t.loc_req)
        s.loc_req = self.lr.ctypes.data
        s.nint_cond = len(t.int_cond)
        s.int_cond = self.ic.ctypes.data
        s.nclock_cond = len(t.clock_cond)
        s.clock_cond = self.cc.ctypes.data
        self.struct = s

    def ptr(self):
        return C.byref(self.struct)


def parse_op(op: str) -> int:
    return _OPS[op]
