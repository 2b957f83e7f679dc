"""UPPAAL XML models and safety queries, the subset the back end's networks represent.

Mirrors the reference's ``load_network`` (src/frontend/model_reader.cpp:987-1261) for flat models:

* declarations (global, and per template with the instance name as prefix): ``clock``, ``chan``
  (``broadcast``, ``urgent``), ``const int``, ``int[lo,hi]`` and ``bool`` variables with constant
  initialisers; comments ``//`` and ``/* */``;
* templates without parameters, instantiated in the order of the ``system`` line (one process per
  template in document order without one), directly or through argument-free instantiations
  (``P2 = P1();`` before the ``system`` line); clocks, variables and channels are numbered in
  declaration order, globals first, then each instance's locals in process order;
* locations with a name (else the id), a conjunction of clock upper bounds as invariant, and the
  ``committed`` / ``urgent`` flags; the ``init`` reference;
* transitions: a guard that is a conjunction (``&&``, ``and``) of clock bounds, clock differences
  and integer comparisons against constant expressions; a ``ch!`` / ``ch?`` synchronisation; an
  assignment list of clock resets to 0 and integer variables set to integer expressions (``=`` or
  ``:=``). An expression that reads integer variables (the reference's general update,
  NEdge::gen_update) is compiled into constant updates: the edge becomes one edge per valuation of
  what it reads (``_Builder.split_expression_edges``).

Anything outside the subset (template parameters, ``select``, arrays, functions, disjunctive
guards) raises :class:`ModelError` naming it, as the reference refuses what it cannot represent
rather than verify a different model.

:func:`safety_target` turns the reachability queries the reference's benchmarks record
(``A[] not (P.l && Q.m)``, ``A[] true``, ``A[] x <= c``, ``E<> ...``) into a :class:`Target`
(src/frontend/query_reader.cpp:31-66).
"""
from __future__ import annotations

import itertools
import os
import re
import xml.etree.ElementTree as ET
from typing import Dict, List, Optional, Tuple

from .network import (Channel, Edge, IntVar, Location, Network, OP_EQ, OP_GE, OP_GT, OP_LE, OP_LT, OP_NE,
                      Process, Target, le, lt)


class ModelError(ValueError):
    """The model (or query) lies outside the subset the networks represent."""


_OPS = {"==": OP_EQ, "!=": OP_NE, "<": OP_LT, "<=": OP_LE, ">": OP_GT, ">=": OP_GE}
_FLIP = {"==": "==", "!=": "!=", "<": ">", "<=": ">=", ">": "<", ">=": "<="}
_NEG = {"==": "!=", "!=": "==", "<": ">=", "<=": ">", ">": "<=", ">=": "<"}
_TOKEN = re.compile(r"\s*(?:(\d+)|([A-Za-z_][A-Za-z_0-9]*(?:\.[A-Za-z_][A-Za-z_0-9]*)?)|(:=|==|!=|<=|>=|&&|\|\||[-+*/%()<>=!,\[\]]))")


def _strip_comments(s: str) -> str:
    s = re.sub(r"/\*.*?\*/", " ", s, flags=re.S)
    return re.sub(r"//[^\n]*", " ", s)


def _tokens(s: str) -> List[str]:
    out, pos, s = [], 0, s.strip()
    while pos < len(s):
        m = _TOKEN.match(s, pos)
        if not m or m.end() == pos:
            if s[pos:].strip() == "":
                break
            raise ModelError(f"cannot read {s[pos:]!r} in {s!r}")
        out.append(m.group(m.lastindex))
        pos = m.end()
    return out


class _Scope:
    """Names visible in one template instance: constants, clocks, integer variables, channels."""

    def __init__(self, parent: Optional["_Scope"] = None):
        self.parent = parent
        self.consts: Dict[str, int] = {}
        self.clocks: Dict[str, int] = {}
        self.ints: Dict[str, int] = {}
        self.chans: Dict[str, int] = {}

    def find(self, table: str, name: str):
        s = self
        while s is not None:
            t = getattr(s, table)
            if name in t:
                return t[name]
            s = s.parent
        return None


class _Expr:
    """Integer expressions: + - * / % unary minus, parentheses, named constants, and (given a
    valuation `env` of the integer variables by index) integer variables."""

    def __init__(self, toks: List[str], scope: _Scope, env: Optional[Dict[int, int]] = None):
        self.t, self.i, self.sc, self.env = toks, 0, scope, env

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else None

    def take(self):
        tok = self.peek()
        self.i += 1
        return tok

    def expr(self) -> int:
        v = self.term()
        while self.peek() in ("+", "-"):
            v = v + self.term() if self.take() == "+" else v - self.term()
        return v

    def term(self) -> int:
        v = self.unary()
        while self.peek() in ("*", "/", "%"):
            op, r = self.take(), self.unary()
            if op == "*":
                v *= r
            elif r == 0:
                raise ModelError("division by zero in a constant expression")
            else:   # C semantics: truncation toward zero
                q = abs(v) // abs(r)
                q = q if (v >= 0) == (r >= 0) else -q
                v = q if op == "/" else v - q * r
        return v

    def unary(self) -> int:
        if self.peek() == "-":
            self.take()
            return -self.unary()
        if self.peek() == "+":
            self.take()
            return self.unary()
        tok = self.take()
        if tok == "(":
            v = self.expr()
            if self.take() != ")":
                raise ModelError("unbalanced parentheses")
            return v
        if tok is None:
            raise ModelError("expression ends early")
        if tok.isdigit():
            return int(tok)
        c = self.sc.find("consts", tok)
        if c is not None:
            return c
        v = self.sc.find("ints", tok) if self.env is not None else None
        if v is None:
            raise ModelError(f"not a constant: {tok}")
        return self.env[v]


def _const(toks: List[str], scope: _Scope, env: Optional[Dict[int, int]] = None) -> int:
    e = _Expr(toks, scope, env)
    v = e.expr()
    if e.peek() is not None:
        raise ModelError(f"unexpected {e.peek()!r} in {' '.join(toks)!r}")
    return v


def _reads(toks: List[str], scope: _Scope) -> List[int]:
    """The integer variables an expression names, in first-use order."""
    out: List[int] = []
    for t in toks:
        if scope.find("consts", t) is None:
            v = scope.find("ints", t)
            if v is not None and v not in out:
                out.append(v)
    return out


def _split_top(toks: List[str], sep: str) -> List[List[str]]:
    parts, cur, depth = [], [], 0
    for t in toks:
        depth += t in ("(", "[")
        depth -= t in (")", "]")
        if t == sep and depth == 0:
            parts.append(cur)
            cur = []
        else:
            cur.append(t)
    parts.append(cur)
    return parts


class _Builder:
    def __init__(self):
        self.clock_names = ["0"]
        self.intvars: List[IntVar] = []
        self.channels: List[Channel] = []
        # edges whose updates read integer variables: id(edge) -> (edge, [(var, tokens, scope) or
        # (var, None, constant)] in assignment order, from the first such update on)
        self.pending: Dict[int, Tuple[Edge, list]] = {}

    # ---- declarations (model_reader.cpp parse_decl_block)
    def declare(self, text: str, scope: _Scope, prefix: str) -> None:
        for stmt in _strip_comments(text).split(";"):
            toks = _tokens(stmt)
            if not toks:
                continue
            if toks[0] == "clock":
                for part in _split_top(toks[1:], ","):
                    if len(part) != 1:
                        raise ModelError(f"unsupported clock declaration: {stmt.strip()}")
                    scope.clocks[part[0]] = len(self.clock_names)
                    self.clock_names.append(prefix + part[0])
            elif "chan" in toks:
                k = toks.index("chan")
                quals = set(toks[:k])
                if not quals <= {"broadcast", "urgent"}:
                    raise ModelError(f"unsupported channel declaration: {stmt.strip()}")
                for part in _split_top(toks[k + 1:], ","):
                    if len(part) != 1:
                        raise ModelError(f"channel arrays are not supported: {stmt.strip()}")
                    scope.chans[part[0]] = len(self.channels)
                    self.channels.append(Channel(prefix + part[0], "broadcast" in quals, "urgent" in quals))
            elif toks[0] == "const":
                if toks[1:2] != ["int"]:
                    raise ModelError(f"unsupported constant: {stmt.strip()}")
                for part in _split_top(toks[2:], ","):
                    if len(part) < 3 or part[1] != "=":
                        raise ModelError(f"unsupported constant: {stmt.strip()}")
                    scope.consts[part[0]] = _const(part[2:], scope)
            elif toks[0] in ("int", "bool"):
                rest = toks[1:]
                if toks[0] == "bool":
                    lo, hi = 0, 1
                elif rest[:1] == ["["]:
                    close = rest.index("]")
                    lo_t, hi_t = _split_top(rest[1:close], ",")
                    lo, hi = _const(lo_t, scope), _const(hi_t, scope)
                    rest = rest[close + 1:]
                else:
                    lo, hi = -32768, 32767   # UPPAAL's default int range
                for part in _split_top(rest, ","):
                    if not part or "[" in part:
                        raise ModelError(f"integer arrays are not supported: {stmt.strip()}")
                    init = 0
                    if len(part) > 1:
                        if part[1] != "=":
                            raise ModelError(f"unsupported declaration: {stmt.strip()}")
                        v = part[2:]
                        init = 1 if v == ["true"] else 0 if v == ["false"] else _const(v, scope)
                    scope.ints[part[0]] = len(self.intvars)
                    self.intvars.append(IntVar(prefix + part[0], lo, hi, init))
            else:
                raise ModelError(f"unsupported declaration: {stmt.strip()}")

    # ---- guards, invariants, assignments (parse_net_guard / parse_net_assign)
    def atom(self, toks: List[str], scope: _Scope, clock_out: list, int_out: Optional[list]) -> None:
        if toks == ["true"]:
            return
        ops = [k for k, t in enumerate(toks) if t in _OPS]
        if len(ops) != 1:
            raise ModelError(f"unsupported constraint: {' '.join(toks)}")
        k = ops[0]
        lhs, op, rhs = toks[:k], toks[k], toks[k + 1:]
        if self._is_var(rhs, scope) and not self._is_var(lhs, scope):
            lhs, rhs, op = rhs, lhs, _FLIP[op]
        c = _const(rhs, scope)
        if len(lhs) == 1 and scope.find("clocks", lhs[0]) is not None:
            self._clock(scope.find("clocks", lhs[0]), 0, op, c, clock_out)
        elif len(lhs) == 3 and lhs[1] == "-" and scope.find("clocks", lhs[0]) is not None \
                and scope.find("clocks", lhs[2]) is not None:
            self._clock(scope.find("clocks", lhs[0]), scope.find("clocks", lhs[2]), op, c, clock_out)
        elif len(lhs) == 1 and scope.find("ints", lhs[0]) is not None and int_out is not None:
            int_out.append((scope.find("ints", lhs[0]), _OPS[op], c))
        else:
            raise ModelError(f"unsupported constraint: {' '.join(toks)}")

    @staticmethod
    def _is_var(toks, scope) -> bool:
        return bool(toks) and (scope.find("clocks", toks[0]) is not None or scope.find("ints", toks[0]) is not None)

    @staticmethod
    def _clock(i: int, j: int, op: str, c: int, out: list) -> None:
        """x_i - x_j op c as DBM constraints (i, j, raw): x_i - x_j <= / < c, x_j - x_i <= / < -c."""
        if op in ("<=", "=="):
            out.append((i, j, le(c)))
        if op == "<":
            out.append((i, j, lt(c)))
        if op in (">=", "=="):
            out.append((j, i, le(-c)))
        if op == ">":
            out.append((j, i, lt(-c)))
        if op == "!=":
            raise ModelError("clock disequalities are not supported")

    def conj(self, text: str, scope: _Scope, clock_out: list, int_out: Optional[list]) -> None:
        toks = ["&&" if t == "and" else t for t in _tokens(text)]
        if "||" in toks or "or" in toks or "not" in toks or "!" in toks:
            raise ModelError(f"only conjunctions are supported: {text.strip()}")
        for part in _split_top(_unwrap(toks), "&&"):
            part = _unwrap(part)
            if part:
                self.atom(part, scope, clock_out, int_out)

    def assign(self, text: str, scope: _Scope, edge: Edge) -> None:
        for part in _split_top(_tokens(text), ","):
            if not part:
                continue
            if len(part) < 3 or part[1] not in ("=", ":="):
                raise ModelError(f"unsupported assignment: {text.strip()}")
            name, rhs = part[0], part[2:]
            if scope.find("clocks", name) is not None:
                if _const(rhs, scope) != 0:
                    raise ModelError(f"clock assignments other than reset to 0: {text.strip()}")
                edge.resets.append(scope.find("clocks", name))
            elif scope.find("ints", name) is not None:
                var = scope.find("ints", name)
                if _reads(rhs, scope):
                    # an expression over integer variables: compiled away by split_expression_edges
                    _Expr(rhs, scope, {v: 0 for v in _reads(rhs, scope)})   # names resolve
                    self.pending.setdefault(id(edge), (edge, []))[1].append((var, rhs, scope))
                    continue
                v = _const(rhs, scope)
                if id(edge) in self.pending:   # after an expression update: keeps its place in order
                    self.pending[id(edge)][1].append((var, None, v))
                else:
                    edge.int_updates.append((var, v))
            else:
                raise ModelError(f"assignment to an unknown name: {text.strip()}")

    # ---- expression updates (model_reader.cpp's general updates, NEdge::gen_update)
    def split_expression_edges(self, procs: List["Process"]) -> None:
        """Compile every update that reads integer variables into constant updates: the edge is
        replaced, in place, by one copy per valuation of the variables its updates read (each copy
        guarded by that valuation, its updates evaluated in order with C's truncating division).
        In any state exactly one copy is enabled, with the successor the original edge gives, and
        the copies stand where the edge stood, so successors keep their order. A valuation whose
        update leaves a variable's declared range disables its copy (a faulting update kills the
        edge). The valuations range over the values each variable can take: its initial value and
        every value an update can give it, closed under the updates (at most 4,096 per variable)."""
        if not self.pending:
            return
        vals = {k: {v.init} for k, v in enumerate(self.intvars)}
        for p in procs:
            for e in p.edges:
                for var, val in e.int_updates:
                    vals[var].add(val)

        def reads(seq):
            r: List[int] = []
            for _, toks, sc in seq:
                for v in (_reads(toks, sc) if toks is not None else []):
                    if v not in r:
                        r.append(v)
            return r

        def valuations(seq):
            r = reads(seq)
            n = 1
            for v in r:
                n *= len(vals[v])
            if n > 100000:
                raise ModelError("an expression update over more than 100,000 valuations")
            for vs in itertools.product(*[sorted(vals[v]) for v in r]):
                yield r, dict(zip(r, vs))

        def evaluate(e, seq, env):
            env = dict(env)
            for var, val in e.int_updates:      # constant updates written before the first expression
                env[var] = val
            out = []
            for var, toks, arg in seq:
                x = arg if toks is None else _const(toks, arg, env)
                iv = self.intvars[var]
                if not iv.lo <= x <= iv.hi:
                    return None
                env[var] = x
                out.append((var, x))
            return out

        # the values' closure under the updates, semi-naively: each pass evaluates only the
        # valuations that use a value found in the previous pass
        old = {k: set() for k in vals}
        new = {k: set(v) for k, v in vals.items()}
        first = True
        while first or any(new.values()):
            added = {k: set() for k in vals}
            for e, seq in self.pending.values():
                r = reads(seq)
                pools_list = [[]] if not r and first else []
                for k in range(len(r)):
                    pools_list.append([(old[r[i]] | new[r[i]]) if i < k else new[r[i]] if i == k else old[r[i]]
                                       for i in range(len(r))])
                for pools in pools_list:
                    for vs in itertools.product(*[sorted(p) for p in pools]):
                        for var, x in evaluate(e, seq, dict(zip(r, vs))) or []:
                            if x not in old[var] and x not in new[var]:
                                added[var].add(x)
            for k in vals:
                old[k] |= new[k]
                new[k] = added[k]
                if len(old[k]) + len(new[k]) > 4096:
                    raise ModelError(f"updates give {self.intvars[k].name} more than 4,096 values")
            first = False
        vals = old
        for p in procs:
            new_edges = []
            for e in p.edges:
                if id(e) not in self.pending:
                    new_edges.append(e)
                    continue
                seq = self.pending[id(e)][1]
                for r, env in valuations(seq):
                    if not all(_atom_holds(env[v], op, c) for v, op, c in e.int_guard if v in env):
                        continue   # the edge's own guard disables it at this valuation
                    out = evaluate(e, seq, env)
                    if out is None:
                        continue
                    new_edges.append(Edge(e.src, e.tgt, list(e.guard), list(e.resets),
                                          list(e.int_guard) + [(v, OP_EQ, env[v]) for v in r],
                                          list(e.int_updates) + out, e.chan, e.dir))
            p.edges = new_edges
        self.pending = {}


def _atom_holds(x: int, op: int, c: int) -> bool:
    return {OP_EQ: x == c, OP_NE: x != c, OP_LT: x < c, OP_LE: x <= c, OP_GT: x > c, OP_GE: x >= c}[op]


def _text(el) -> str:
    return (el.text or "") if el is not None else ""


def _system_order(text: str) -> Tuple[Optional[List[str]], Dict[str, str]]:
    """The process names of the `system` line, and the instantiations before it (`P = T();`:
    instance name -> template name; instances of templates without parameters only)."""
    body = _strip_comments(text)
    m = re.search(r"\bsystem\b([^;]*);", body)
    if not m:
        return None, {}
    inst: Dict[str, str] = {}
    for stmt in body[:m.start()].split(";"):
        if not stmt.strip():
            continue
        mi = re.fullmatch(r"\s*([A-Za-z_]\w*)\s*=\s*([A-Za-z_]\w*)\s*\(\s*\)\s*", stmt)
        if not mi:
            raise ModelError(f"process instantiations with arguments are not supported: {stmt.strip()}")
        inst[mi.group(1)] = mi.group(2)
    names = [n.strip() for n in m.group(1).replace("<", ",").split(",")]
    return [n for n in names if n], inst


def parse_xml(text: str, name: str = "") -> Network:
    """The network of a flat UPPAAL XML model (the subset in the module docstring)."""
    root = ET.fromstring(text)
    if root.tag != "nta":
        raise ModelError("not an <nta> model")
    b = _Builder()
    glob = _Scope()
    b.declare(_text(root.find("declaration")), glob, "")
    templates = {}
    order_doc = []
    for t in root.findall("template"):
        tn = _text(t.find("name")).strip() or "P"
        templates[tn] = t
        order_doc.append(tn)
    sysel = root.find("system")
    order, inst = _system_order(_text(sysel)) if sysel is not None else (None, {})
    order = order if order else order_doc
    procs = []
    for pname in order:
        t = templates.get(inst.get(pname, pname))
        if t is None:
            raise ModelError(f"system declaration names unknown process or template: {pname}")
        if _text(t.find("parameter")).strip():
            raise ModelError(f"template parameters are not supported: {pname}")
        sc = _Scope(glob)
        b.declare(_text(t.find("declaration")), sc, pname + ".")
        locs, lid = [], {}
        for el in t.findall("location"):
            nm = el.find("name")
            loc = Location(_text(nm).strip() if nm is not None else el.get("id"),
                           committed=el.find("committed") is not None, urgent=el.find("urgent") is not None)
            for lb in el.findall("label"):
                if lb.get("kind") == "invariant":
                    b.conj(_text(lb), sc, loc.invariant, None)
            lid[el.get("id")] = len(locs)
            locs.append(loc)
        init = t.find("init")
        edges = []
        for tr in t.findall("transition"):
            e = Edge(lid[tr.find("source").get("ref")], lid[tr.find("target").get("ref")])
            for lb in tr.findall("label"):
                kind, body = lb.get("kind"), _text(lb)
                if kind == "guard":
                    b.conj(body, sc, e.guard, e.int_guard)
                elif kind == "assignment":
                    b.assign(body, sc, e)
                elif kind == "synchronisation":
                    s = body.strip()
                    if not s:
                        continue
                    if s[-1] not in "!?":
                        raise ModelError(f"malformed synchronisation: {s}")
                    ch = s[:-1].strip()
                    if "[" in ch or sc.find("chans", ch) is None:
                        raise ModelError(f"unsupported channel: {ch}")
                    e.chan, e.dir = sc.find("chans", ch), 1 if s[-1] == "!" else -1
                elif kind in ("select",):
                    raise ModelError("select bindings are not supported")
                elif kind in ("comments", "probability", "testcode"):
                    continue
                elif body.strip():
                    raise ModelError(f"unsupported label kind: {kind}")
            edges.append(e)
        procs.append(Process(pname, locs, edges, lid.get(init.get("ref"), 0) if init is not None else 0))
    b.split_expression_edges(procs)
    net = Network(len(b.clock_names) - 1, b.intvars, b.channels, procs, clock_names=b.clock_names, name=name)
    return net.finalize()


def load(path: str) -> Network:
    """The network of the model file at `path`, named after the file."""
    with open(path, encoding="utf-8") as f:
        text = f.read()
    return parse_xml(text, os.path.splitext(os.path.basename(path))[0])


def _unwrap(toks: List[str]) -> List[str]:
    """Drop parentheses that enclose the whole token list."""
    while toks[:1] == ["("] and toks[-1:] == [")"]:
        depth = 0
        for k, t in enumerate(toks):
            depth += t == "("
            depth -= t == ")"
            if depth == 0 and k < len(toks) - 1:
                return toks
        toks = toks[1:-1]
    return toks


def _target_atoms(net: Network, text: str, tgt: Target, negate: bool) -> None:
    toks = _unwrap(["&&" if t == "and" else t for t in _tokens(text)])
    if "||" in toks or "or" in toks:
        raise ModelError(f"only conjunctions are supported in queries: {text}")
    parts = _split_top(toks, "&&")
    if negate and len(parts) != 1:
        raise ModelError(f"A[] of a conjunction is not a single target: {text}")
    glob = _Scope()
    glob.clocks = {n: i for i, n in enumerate(net.clock_names) if i > 0}
    glob.ints = {v.name: i for i, v in enumerate(net.intvars)}
    for part in parts:
        part = _unwrap(part)
        if len(part) == 1 and "." in part[0]:
            if negate:
                raise ModelError(f"A[] of a location is not a single target: {text}")
            p, l = part[0].split(".", 1)
            tgt.loc_req.append(net.resolve(p, l))
            continue
        if negate:
            k = next(k for k, t in enumerate(part) if t in _OPS)
            part = part[:k] + [_NEG[part[k]]] + part[k + 1:]
        _Builder().atom(part, glob, tgt.clock_cond, tgt.int_cond)


def safety_target(net: Network, query: str) -> Tuple[bool, Target]:
    """(is_invariance, target) of ``A[] not (conj)``, ``A[] true``, ``A[] atom`` or ``E<> conj``:
    the invariance holds iff the target is unreachable; ``E<>`` holds iff it is reachable."""
    q = query.strip()
    if q.startswith("E<>"):
        t = Target()
        _target_atoms(net, q[3:], t, negate=False)
        return False, t
    if not q.startswith("A[]"):
        raise ModelError(f"not a safety query: {q}")
    body = q[3:].strip()
    if body == "true":
        return True, Target(never=True)
    t = Target()
    if body.startswith("not"):
        _target_atoms(net, body[3:], t, negate=False)
    else:
        _target_atoms(net, body, t, negate=True)
    return True, t
