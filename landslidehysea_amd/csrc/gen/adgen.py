"""Source-transformation reverse-mode AD for straight-line device code.

A program is written with Python operator overloading into an SSA graph of
scalar nodes (FP32-or-FP64 "F" values and FP64 "D" values) and boolean
conditions.  ``emit`` prints one C++ device function that recomputes the
forward values and then runs the reverse sweep: every adjoint rule is the
derivative of one primitive, and every branch (comparison, min/max, select,
|x|) is frozen at the choice the forward values make, which is the discrete
adjoint of the branchy step with its branches held fixed.

Primitive rules (z = op(...), gz the adjoint of z):
  a+b, a-b, a*b, a/b, -a, sqrt, rsq (a^-1/2), rcp, exp, a^c (constant c),
  |a|, min, max, select(c, a, b), casts.
"""
from __future__ import annotations

import itertools


class Cond:
    """A boolean: kind "cmp" (args: operator, node, node), "and", "or" (two
    conditions) or "not" (one)."""

    def __init__(self, prog, kind, args):
        self.prog, self.kind, self.args = prog, kind, args
        self.id = prog._new_cond(self)

    def __and__(self, o):
        return Cond(self.prog, "and", [self, o])

    def __or__(self, o):
        return Cond(self.prog, "or", [self, o])

    def __invert__(self):
        return Cond(self.prog, "not", [self])

    def ref(self):
        return f"c{self.id}"


class Node:
    def __init__(self, prog, op, args, dt, active, code=None):
        self.prog, self.op, self.args, self.dt, self.active = prog, op, args, dt, active
        self.code = code
        self.id = prog._new_node(self)

    def ref(self):
        return f"v{self.id}"

    # arithmetic
    def _w(self, o):
        return o if isinstance(o, Node) else self.prog.const(o, self.dt)

    def __add__(self, o): return self.prog.op("add", self, self._w(o))
    def __radd__(self, o): return self.prog.op("add", self._w(o), self)
    def __sub__(self, o): return self.prog.op("sub", self, self._w(o))
    def __rsub__(self, o): return self.prog.op("sub", self._w(o), self)
    def __mul__(self, o): return self.prog.op("mul", self, self._w(o))
    def __rmul__(self, o): return self.prog.op("mul", self._w(o), self)
    def __truediv__(self, o): return self.prog.op("div", self, self._w(o))
    def __rtruediv__(self, o): return self.prog.op("div", self._w(o), self)
    def __neg__(self): return self.prog.op("neg", self)

    def _cmp(self, o, s):
        o = self._w(o)
        return Cond(self.prog, "cmp", [s, self, o])

    def __lt__(self, o): return self._cmp(o, "<")
    def __le__(self, o): return self._cmp(o, "<=")
    def __gt__(self, o): return self._cmp(o, ">")
    def __ge__(self, o): return self._cmp(o, ">=")


class Prog:
    def __init__(self, name):
        self.name = name
        self.nodes, self.conds, self.seq = [], [], []
        self.inputs, self.params, self.outputs = [], [], []
        self._consts = {}

    def _new_node(self, n):
        self.nodes.append(n)
        self.seq.append(n)
        return len(self.nodes) - 1

    def _new_cond(self, c):
        self.conds.append(c)
        self.seq.append(c)
        return len(self.conds) - 1

    def inp(self, name, dt="F"):
        n = Node(self, "inp", [], dt, True, code=name)
        self.inputs.append(n)
        return n

    def param(self, name, dt="F"):
        n = Node(self, "param", [], dt, False, code=name)
        self.params.append(n)
        return n

    def const(self, v, dt="F"):
        key = (repr(float(v)), dt)
        if key not in self._consts:
            self._consts[key] = Node(self, "const", [], dt, False, code=repr(float(v)))
        return self._consts[key]

    def op(self, op, *args):
        dt = "D" if any(a.dt == "D" for a in args if isinstance(a, Node)) and op not in ("tof",) else args[0].dt
        if op == "tof":
            dt = "F"
        if op == "tod":
            dt = "D"
        act = any(a.active for a in args if isinstance(a, Node))
        return Node(self, op, list(args), dt, act)

    def out(self, name, node):
        self.outputs.append((name, node))


# functions usable in programs
def sqrt(a): return a.prog.op("sqrt", a)
def rsq(a): return a.prog.op("rsq", a)
def rcp(a): return a.prog.op("rcp", a)
def exp(a): return a.prog.op("exp", a)
def fabs(a): return a.prog.op("fabs", a)
def fmin(a, b): return a.prog.op("fmin", a, a._w(b))
def fmax(a, b): return a.prog.op("fmax", a, a._w(b))
def powc(a, c): return a.prog.op("powc", a, a.prog.const(c, a.dt))
def tof(a): return a.prog.op("tof", a)
def tod(a): return a.prog.op("tod", a)


def select(c, a, b):
    p = c.prog
    a = a if isinstance(a, Node) else p.const(a, b.dt if isinstance(b, Node) else "F")
    b = b if isinstance(b, Node) else p.const(b, a.dt)
    n = p.op("select", a, b)
    n.cond = c
    return n


def _ty(dt):
    return "F" if dt == "F" else "double"


class Knowledge:
    """What the emitter may assume about forward values.

    zeros: input names whose value is exactly 0; positive / nonneg: parameter
    or input names known > 0 / >= 0.  From these, every node gets a sign
    class ("zero", "pos", "neg", "nonneg", "nonpos" or "unk"), may be an alias
    of another node (x + 0, select with a known condition, min/max with a
    known order), and conditions may be decided.  Values are assumed finite,
    so x * 0 = 0.  With no assumptions nothing is folded."""

    def __init__(self, prog, zeros=(), positive=(), nonneg=()):
        self.on = bool(zeros or positive or nonneg)
        self.sign, self.val, self.rep, self.truth = {}, {}, {}, {}
        self.zeros, self.positive, self.nonneg = set(zeros), set(positive), set(nonneg)
        if not self.on:
            return
        for x in prog.seq:
            if isinstance(x, Cond):
                self.truth[x.id] = self._cond(x)
            else:
                self._node(x)

    # ---- queries
    def Z(self, n):
        return self.on and self.sign.get(n.id) == "zero"

    def S(self, n):
        return self.sign.get(n.id, "unk") if self.on else "unk"

    def T(self, c):
        return self.truth.get(c.id) if self.on else None

    def R(self, n):
        """The node whose forward value n has."""
        while self.on and n.id in self.rep:
            n = self.rep[n.id]
        return n

    def le(self, a, b):
        """a <= b known: True / False / None."""
        sa, sb = self.S(a), self.S(b)
        va, vb = self.val.get(a.id), self.val.get(b.id)
        if va is not None and vb is not None:
            return va <= vb
        if self.R(a) is self.R(b):
            return True
        if sa in ("neg", "nonpos", "zero") and sb in ("pos", "nonneg", "zero"):
            return True
        if sa == "pos" and sb in ("neg", "nonpos", "zero"):
            return False
        if sa in ("pos", "nonneg", "zero") and sb == "neg":
            return False
        return None

    def lt(self, a, b):
        sa, sb = self.S(a), self.S(b)
        va, vb = self.val.get(a.id), self.val.get(b.id)
        if va is not None and vb is not None:
            return va < vb
        if sa == "neg" and sb in ("pos", "nonneg", "zero"):
            return True
        if sa in ("nonpos", "zero") and sb == "pos":
            return True
        r = self.le(b, a)
        if r is True:
            return False
        return None

    # ---- propagation
    def _cond(self, c):
        if c.kind == "cmp":
            op, a, b = c.args
            return {"<": lambda: self.lt(a, b), "<=": lambda: self.le(a, b),
                    ">": lambda: self.lt(b, a), ">=": lambda: self.le(b, a)}[op]()
        t = [self.truth.get(x.id) for x in c.args]
        if c.kind == "not":
            return None if t[0] is None else (not t[0])
        if c.kind == "and":
            if False in t:
                return False
            return True if t == [True, True] else None
        if True in t:
            return True
        return False if t == [False, False] else None

    def _set(self, n, sign, val=None, rep=None):
        self.sign[n.id] = sign
        if val is not None:
            self.val[n.id] = val
        if rep is not None and rep is not n:
            self.rep[n.id] = rep

    @staticmethod
    def _mulsign(a, b):
        if "zero" in (a, b):
            return "zero"
        if "unk" in (a, b):
            return "unk"
        pos = {"pos": 1, "nonneg": 1, "neg": -1, "nonpos": -1}
        strict = a in ("pos", "neg") and b in ("pos", "neg")
        if pos[a] * pos[b] > 0:
            return "pos" if strict else "nonneg"
        return "neg" if strict else "nonpos"

    @staticmethod
    def _negsign(a):
        return {"pos": "neg", "neg": "pos", "nonneg": "nonpos", "nonpos": "nonneg"}.get(a, a)

    def _node(self, n):
        A = [y for y in n.args if isinstance(y, Node)]
        sg = [self.S(y) for y in A]
        if n.op == "inp":
            self._set(n, "zero" if n.code in self.zeros else
                      "pos" if n.code in self.positive else "nonneg" if n.code in self.nonneg else "unk",
                      0.0 if n.code in self.zeros else None)
        elif n.op == "param":
            self._set(n, "pos" if n.code in self.positive else "nonneg" if n.code in self.nonneg else "unk")
        elif n.op == "const":
            v = float(n.code)
            self._set(n, "zero" if v == 0 else "pos" if v > 0 else "neg", v)
        elif n.op == "add":
            if sg[0] == "zero":
                self._set(n, sg[1], self.val.get(A[1].id), A[1])
            elif sg[1] == "zero":
                self._set(n, sg[0], self.val.get(A[0].id), A[0])
            elif sg[0] in ("pos", "nonneg") and sg[1] in ("pos", "nonneg"):
                self._set(n, "pos" if "pos" in sg else "nonneg")
            elif sg[0] in ("neg", "nonpos") and sg[1] in ("neg", "nonpos"):
                self._set(n, "neg" if "neg" in sg else "nonpos")
            else:
                self._set(n, "unk")
        elif n.op == "sub":
            if sg[1] == "zero":
                self._set(n, sg[0], self.val.get(A[0].id), A[0])
            elif sg[0] == "zero":
                self._set(n, self._negsign(sg[1]))
            elif sg[0] in ("pos", "nonneg") and sg[1] in ("neg", "nonpos"):
                self._set(n, "pos" if "pos" in (sg[0], self._negsign(sg[1])) else "nonneg")
            else:
                self._set(n, "unk")
        elif n.op == "mul":
            self._set(n, self._mulsign(sg[0], sg[1]), 0.0 if "zero" in sg else None)
        elif n.op == "div":
            s = "zero" if sg[0] == "zero" else self._mulsign(sg[0], sg[1])
            self._set(n, s, 0.0 if s == "zero" else None)
        elif n.op == "neg":
            self._set(n, self._negsign(sg[0]), 0.0 if sg[0] == "zero" else None)
        elif n.op == "sqrt":
            self._set(n, {"zero": "zero", "pos": "pos"}.get(sg[0], "nonneg"), 0.0 if sg[0] == "zero" else None)
        elif n.op in ("rsq", "rcp"):
            self._set(n, sg[0] if sg[0] in ("pos",) or (n.op == "rcp" and sg[0] == "neg") else "unk")
        elif n.op == "exp":
            if sg[0] == "zero":
                self._set(n, "pos", 1.0)
            else:
                self._set(n, "nonneg")
        elif n.op == "powc":
            c = float(A[1].code)
            if sg[0] == "zero" and c > 0:
                self._set(n, "zero", 0.0)
            else:
                self._set(n, "pos" if sg[0] == "pos" else "unk")
        elif n.op == "fabs":
            self._set(n, "zero" if sg[0] == "zero" else "pos" if sg[0] in ("pos", "neg") else "nonneg",
                      0.0 if sg[0] == "zero" else None)
        elif n.op in ("fmin", "fmax"):
            first = self.le(A[0], A[1]) if n.op == "fmin" else self.le(A[1], A[0])
            if first is True:
                self._set(n, sg[0], self.val.get(A[0].id), A[0])
            elif first is False:
                self._set(n, sg[1], self.val.get(A[1].id), A[1])
            else:
                self._set(n, "unk")
        elif n.op == "select":
            t = self.truth.get(n.cond.id)
            if t is True:
                self._set(n, sg[0], self.val.get(A[0].id), A[0])
            elif t is False:
                self._set(n, sg[1], self.val.get(A[1].id), A[1])
            elif sg[0] == "zero" and sg[1] == "zero":
                self._set(n, "zero", 0.0)
            else:
                self._set(n, "unk")
        elif n.op in ("tof", "tod"):
            s = sg[0]
            if n.op == "tof" and s == "pos":
                s = "nonneg"          # may underflow to 0 in FP32
            self._set(n, s, 0.0 if s == "zero" else None, A[0] if (n.op == "tod" and s != "zero") else None)
        else:
            raise ValueError(n.op)


def emit(prog: Prog, forward_only: bool = False, name: str = None, zeros=(), positive=(), nonneg=()) -> str:
    """C++ device function: void NAME(params..., inputs..., out values..., out-adjoints in, input-adjoints out).

    forward_only: NAME_fwd(params..., inputs..., out values...) with the same arithmetic.
    zeros / positive / nonneg: assumptions on inputs and parameters (see
    Knowledge); the function then drops every forward term and reverse
    contribution that a known-zero value removes, decides known conditions,
    and still returns the adjoints of every input (zero-valued ones too)."""
    K = Knowledge(prog, zeros, positive, nonneg)
    L = []
    nodes = prog.nodes

    def ref(y):
        """Forward value of node y in the emitted code."""
        if not isinstance(y, Node):
            return str(y)
        if K.Z(y):
            return f"{_ty(y.dt)}(0)"
        if K.on and y.id in K.val and y.op != "const":
            return f"{_ty(y.dt)}({K.val[y.id]!r})"
        r = K.R(y)
        return r.ref()

    def cref(c):
        t = K.T(c)
        return c.ref() if t is None else ("true" if t else "false")

    def contributions(n):
        """(arg, expression template) pairs of n's reverse rule under K; "{g}" is n's adjoint."""
        A = [y for y in n.args if isinstance(y, Node)]
        z = ref(n)
        T = _ty(n.dt)
        out = []
        if n.op == "add":
            out = [(A[0], "{g}"), (A[1], "{g}")]
        elif n.op == "sub":
            out = [(A[0], "{g}"), (A[1], "-{g}")]
        elif n.op == "mul":
            if not K.Z(A[1]):
                out.append((A[0], "{g} * " + ref(A[1])))
            if not K.Z(A[0]):
                out.append((A[1], "{g} * " + ref(A[0])))
        elif n.op == "div":
            out.append((A[0], "ad_div({g}, " + ref(A[1]) + ")"))
            if not K.Z(n):
                out.append((A[1], "-ad_div({g} * " + z + ", " + ref(A[1]) + ")"))
        elif n.op == "neg":
            out = [(A[0], "-{g}")]
        elif n.op == "sqrt":
            if K.S(n) == "pos":
                out = [(A[0], "ad_div({g} * " + T + "(0.5), " + z + ")")]
            elif not K.Z(n):
                out = [(A[0], "(" + z + " > 0 ? ad_div({g} * " + T + "(0.5), " + z + ") : " + T + "(0))")]
        elif n.op == "rsq":
            out = [(A[0], T + "(-0.5) * {g} * " + z + " * " + z + " * " + z)]
        elif n.op == "rcp":
            out = [(A[0], "-{g} * " + z + " * " + z)]
        elif n.op == "exp":
            out = [(A[0], "{g} * " + z)]
        elif n.op == "powc":
            if not K.Z(A[0]):
                a0 = ref(A[0])
                out = [(A[0], "(" + a0 + " != 0 ? ad_div({g} * " + ref(A[1]) + " * " + z + ", " + a0 + ") : " + T +
                        "(0))")]
        elif n.op == "fabs":
            s0 = K.S(A[0])
            if s0 in ("zero", "pos", "nonneg"):
                out = [(A[0], "{g}")]
            elif s0 == "neg":
                out = [(A[0], "-{g}")]
            else:
                out = [(A[0], "(" + ref(A[0]) + " >= 0 ? {g} : -{g})")]
        elif n.op in ("fmin", "fmax"):
            first = K.le(A[0], A[1]) if n.op == "fmin" else K.le(A[1], A[0])
            if first is True:
                out = [(A[0], "{g}")]
            elif first is False:
                out = [(A[1], "{g}")]
            else:
                c = (ref(A[0]) + (" <= " if n.op == "fmin" else " >= ") + ref(A[1]))
                out = [(A[0], "(" + c + " ? {g} : " + T + "(0))"), (A[1], "(" + c + " ? " + T + "(0) : {g})")]
        elif n.op == "select":
            t = K.T(n.cond)
            if t is True:
                out = [(A[0], "{g}")]
            elif t is False:
                out = [(A[1], "{g}")]
            else:
                c = n.cond.ref()
                out = [(A[0], "(" + c + " ? {g} : " + T + "(0))"), (A[1], "(" + c + " ? " + T + "(0) : {g})")]
        elif n.op in ("tof", "tod"):
            out = [(A[0], "{g}")]
        else:
            raise ValueError(n.op)
        return out

    # which active nodes can reach an output through the reverse rules
    need = set()
    for _, o in prog.outputs:
        if o.active:
            need.add(o.id)
    for n in reversed(nodes):
        if n.id in need and n.active and n.op not in ("inp", "param", "const"):
            for y, _ in contributions(n):
                if y.active:
                    need.add(y.id)
    args = []
    for n in prog.params:
        args.append(f"{_ty(n.dt)} {n.code}")
    for n in prog.inputs:
        args.append(f"{_ty(n.dt)} {n.code}")
    for nm, o in prog.outputs:
        args.append(f"{_ty(o.dt)}& {nm}")
    if not forward_only:
        for nm, o in prog.outputs:
            args.append(f"{_ty(o.dt)} g_{nm}")
        for n in prog.inputs:
            args.append(f"{_ty(n.dt)}& g_{n.code}")
    L.append("template <typename F>")
    fname = (name or prog.name) + ("_fwd" if forward_only else "")
    L.append(f"__device__ __forceinline__ void {fname}(" + ", ".join(args) + ")\n{")
    # forward
    for x in prog.seq:
        if isinstance(x, Cond):
            if K.T(x) is not None:
                continue
            if x.kind == "cmp":
                code = f"({ref(x.args[1])} {x.args[0]} {ref(x.args[2])})"
            elif x.kind == "not":
                code = f"(!{cref(x.args[0])})"
            else:
                code = f"({cref(x.args[0])} {'&&' if x.kind == 'and' else '||'} {cref(x.args[1])})"
            L.append(f"    const bool {x.ref()} = {code};")
            continue
        n = x
        if K.on and n.op != "const" and (K.Z(n) or n.id in K.rep or n.id in K.val):
            continue                     # a literal or another node's value
        t = _ty(n.dt)
        a = [ref(y) for y in n.args]
        if n.op in ("inp", "param"):
            e = n.code
        elif n.op == "const":
            e = f"{t}({n.code})"
        elif n.op == "add": e = f"{a[0]} + {a[1]}"
        elif n.op == "sub": e = f"-{a[1]}" if K.Z(n.args[0]) else f"{a[0]} - {a[1]}"
        elif n.op == "mul": e = f"{a[0]} * {a[1]}"
        elif n.op == "div": e = f"ad_div({a[0]}, {a[1]})"
        elif n.op == "neg": e = f"-{a[0]}"
        elif n.op == "sqrt": e = f"ad_sqrt({a[0]})"
        elif n.op == "rsq": e = f"ad_rsq({a[0]})"
        elif n.op == "rcp": e = f"ad_rcp({a[0]})"
        elif n.op == "exp": e = f"ad_exp({a[0]})"
        elif n.op == "powc": e = f"ad_pow({a[0]}, {a[1]})"
        elif n.op == "fabs": e = f"fabs({a[0]})"
        elif n.op == "fmin": e = f"({a[0]} <= {a[1]} ? {a[0]} : {a[1]})"
        elif n.op == "fmax": e = f"({a[0]} >= {a[1]} ? {a[0]} : {a[1]})"
        elif n.op == "select": e = f"({n.cond.ref()} ? {a[0]} : {a[1]})"
        elif n.op == "tof": e = f"(F)({a[0]})"
        elif n.op == "tod": e = f"(double)({a[0]})"
        else:
            raise ValueError(n.op)
        L.append(f"    const {t} {n.ref()} = {e};")
    for nm, o in prog.outputs:
        L.append(f"    {nm} = {ref(o)};")
    if forward_only:
        L.append("}")
        return "\n".join(L) + "\n"
    # reverse
    L.append("    // reverse sweep")
    for n in nodes:
        if n.id in need:
            L.append(f"    {_ty(n.dt)} g{n.id} = 0;")
    for nm, o in prog.outputs:
        if o.id in need:
            L.append(f"    g{o.id} += g_{nm};")
    for n in reversed(nodes):
        if n.id not in need or n.op in ("inp", "param", "const"):
            continue
        g = f"g{n.id}"
        # rules with unbounded partials are skipped when the adjoint is zero, so an
        # infinite value in a branch that the forward did not take cannot turn
        # a zero adjoint into NaN
        guard = n.op in ("div", "rcp", "rsq", "sqrt", "powc")
        for y, tmpl in contributions(n):
            if not (y.active and y.id in need):
                continue
            expr = tmpl.replace("{g}", g)
            if guard:
                expr = f"({g} != 0 ? ({expr}) : {_ty(n.dt)}(0))"
            if y.dt != n.dt:
                expr = f"({_ty(y.dt)})({expr})"
            L.append(f"    g{y.id} += {expr};")
    for n in prog.inputs:
        L.append(f"    g_{n.code} = {('g%d' % n.id) if n.id in need else _ty(n.dt) + '(0)'};")
    L.append("}")
    return "\n".join(L) + "\n"
