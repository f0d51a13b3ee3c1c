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
    def __init__(self, prog, code):
        self.prog, self.code = prog, code
        self.id = prog._new_cond(self)

    def __and__(self, o):
        return Cond(self.prog, f"({self.ref()} && {o.ref()})")

    def __or__(self, o):
        return Cond(self.prog, f"({self.ref()} || {o.ref()})")

    def __invert__(self):
        return Cond(self.prog, f"(!{self.ref()})")

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
        return Cond(self.prog, f"({self.ref()} {s} {o.ref()})")

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


def emit(prog: Prog, forward_only: bool = False) -> str:
    """C++ device function: void NAME(params..., inputs..., out values..., out-adjoints in, input-adjoints out).

    forward_only: NAME_fwd(params..., inputs..., out values...) with the same arithmetic."""
    L = []
    nodes = prog.nodes
    # which active nodes can reach an output (prune the reverse sweep)
    need = set()
    for _, o in prog.outputs:
        need.add(o.id)
    for n in reversed(nodes):
        if n.id in need and n.active:
            for a in n.args:
                if isinstance(a, Node) and a.active:
                    need.add(a.id)
    args = []
    for n in prog.params:
        args.append(f"{_ty(n.dt)} {n.code}")
    for n in prog.inputs:
        args.append(f"{_ty(n.dt)} {n.code}")
    for name, o in prog.outputs:
        args.append(f"{_ty(o.dt)}& {name}")
    if not forward_only:
        for name, o in prog.outputs:
            args.append(f"{_ty(o.dt)} g_{name}")
        for n in prog.inputs:
            args.append(f"{_ty(n.dt)}& g_{n.code}")
    L.append("template <typename F>")
    fname = prog.name + ("_fwd" if forward_only else "")
    L.append(f"__device__ __forceinline__ void {fname}(" + ", ".join(args) + ")\n{")
    # forward
    for x in prog.seq:
        if isinstance(x, Cond):
            L.append(f"    const bool {x.ref()} = {x.code};")
            continue
        n = x
        t = _ty(n.dt)
        a = [y.ref() if isinstance(y, Node) else str(y) for y in n.args]
        if n.op in ("inp", "param"):
            e = n.code
        elif n.op == "const":
            e = f"{t}({n.code})"
        elif n.op == "add": e = f"{a[0]} + {a[1]}"
        elif n.op == "sub": e = f"{a[0]} - {a[1]}"
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
    for name, o in prog.outputs:
        L.append(f"    {name} = {o.ref()};")
    if forward_only:
        L.append("}")
        return "\n".join(L) + "\n"
    # reverse
    L.append("    // reverse sweep")
    for n in nodes:
        if n.id in need:
            L.append(f"    {_ty(n.dt)} g{n.id} = 0;")
    for name, o in prog.outputs:
        if o.id in need:
            L.append(f"    g{o.id} += g_{name};")
    for n in reversed(nodes):
        if n.id not in need or n.op in ("inp", "param", "const"):
            continue
        g = f"g{n.id}"
        A = [y for y in n.args if isinstance(y, Node)]

        # rules with unbounded partials are skipped when the adjoint is zero, so an
        # infinite value in a branch that the forward did not take cannot turn
        # a zero adjoint into NaN
        guard = n.op in ("div", "rcp", "rsq", "sqrt", "powc")

        def acc(y, expr):
            if guard:
                expr = f"({g} != 0 ? ({expr}) : {_ty(n.dt)}(0))"
            if y.active and y.id in need:
                if y.dt != n.dt:
                    expr = f"({_ty(y.dt)})({expr})"
                L.append(f"    g{y.id} += {expr};")

        z = n.ref()
        if n.op == "add":
            acc(A[0], g); acc(A[1], g)
        elif n.op == "sub":
            acc(A[0], g); acc(A[1], f"-{g}")
        elif n.op == "mul":
            acc(A[0], f"{g} * {A[1].ref()}"); acc(A[1], f"{g} * {A[0].ref()}")
        elif n.op == "div":
            acc(A[0], f"ad_div({g}, {A[1].ref()})"); acc(A[1], f"-ad_div({g} * {z}, {A[1].ref()})")
        elif n.op == "neg":
            acc(A[0], f"-{g}")
        elif n.op == "sqrt":
            acc(A[0], f"({z} > 0 ? ad_div({g} * {_ty(n.dt)}(0.5), {z}) : {_ty(n.dt)}(0))")
        elif n.op == "rsq":
            acc(A[0], f"{_ty(n.dt)}(-0.5) * {g} * {z} * {z} * {z}")
        elif n.op == "rcp":
            acc(A[0], f"-{g} * {z} * {z}")
        elif n.op == "exp":
            acc(A[0], f"{g} * {z}")
        elif n.op == "powc":
            acc(A[0], f"({A[0].ref()} != 0 ? ad_div({g} * {A[1].ref()} * {z}, {A[0].ref()}) : {_ty(n.dt)}(0))")
        elif n.op == "fabs":
            acc(A[0], f"({A[0].ref()} >= 0 ? {g} : -{g})")
        elif n.op == "fmin":
            acc(A[0], f"({A[0].ref()} <= {A[1].ref()} ? {g} : {_ty(n.dt)}(0))")
            acc(A[1], f"({A[0].ref()} <= {A[1].ref()} ? {_ty(n.dt)}(0) : {g})")
        elif n.op == "fmax":
            acc(A[0], f"({A[0].ref()} >= {A[1].ref()} ? {g} : {_ty(n.dt)}(0))")
            acc(A[1], f"({A[0].ref()} >= {A[1].ref()} ? {_ty(n.dt)}(0) : {g})")
        elif n.op == "select":
            acc(A[0], f"({n.cond.ref()} ? {g} : {_ty(n.dt)}(0))")
            acc(A[1], f"({n.cond.ref()} ? {_ty(n.dt)}(0) : {g})")
        elif n.op in ("tof", "tod"):
            acc(A[0], g)
        else:
            raise ValueError(n.op)
    for n in prog.inputs:
        L.append(f"    g_{n.code} = {('g%d' % n.id) if n.id in need else _ty(n.dt) + '(0)'};")
    L.append("}")
    return "\n".join(L) + "\n"
