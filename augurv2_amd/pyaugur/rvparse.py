"""Parser of the AugurV2 modelling language and kernel schedules.

Grammar and tokens follow compiler/augur/src/Rv/ParseRv.y:87-215 and
compiler/augur/src/Rv/LexRv.x:60-113 (models: ``(hypers) => { decls }``;
decls ``param|data x[i, ..] ~|= exp for i <- e until e, ...;``; schedules
``K1 (*) K2 (*) ...`` with HMC / NUTS / RSlice settings lists and MWG
proposals).  The result is a small AST of tuples and dataclasses.

Expressions:
    ("lit", value) | ("var", name) | ("proj", exp, [exps])
    ("prim", name, [exps])      name in {"sigmoid", "dotprod"}
    ("dist", name, [exps])      a distribution
    ("bin", op, lhs, rhs)       op in "+-*/"
Types: ("Unit",) ("Int",) ("Real",) ("Vec", t) ("Mat", t)
"""
import re
from dataclasses import dataclass, field
from typing import List, Optional, Tuple

DISTS = {"Bernoulli", "Categorical", "Geometric", "Poisson", "Beta", "Exponential", "Gamma",
         "InvGamma", "Normal", "Uniform", "Dirichlet", "MvNormal", "InvWishart"}
PRIMS = {"sigmoid", "dotprod"}
KEYWORDS = {"data", "param", "let", "in", "for", "until", "Unit", "Int", "Real", "Vec", "Mat",
            "MWG", "HMC", "NUTS", "DiscGibbs", "ConjGibbs", "ESlice", "RSlice"}

_TOKEN = re.compile(r"""
    (?P<ws>\s+|--[^\n]*)
  | (?P<flt>[0-9]+\.[0-9]+([eE][-+]?[0-9]+)?|[0-9]+[eE][-+]?[0-9]+)
  | (?P<int>[0-9]+)
  | (?P<op>\(\*\)|<-|=>|[()\[\]{};:,.|~=+\-*/])
  | (?P<id>[A-Za-z_][A-Za-z0-9_']*)
""", re.VERBOSE)


class RvParseError(ValueError):
    pass


def tokenize(src):
    pos, out = 0, []
    while pos < len(src):
        m = _TOKEN.match(src, pos)
        if not m:
            raise RvParseError(f"lexical error at offset {pos}: {src[pos:pos + 20]!r}")
        pos = m.end()
        kind = m.lastgroup
        text = m.group(kind)
        if kind == "ws":
            continue
        if kind == "id":
            if text in KEYWORDS:
                kind = "kw"
            elif text in DISTS:
                kind = "dist"
            elif text in PRIMS:
                kind = "prim"
        out.append((kind, text))
    out.append(("eof", ""))
    return out


@dataclass
class Decl:
    kind: str                      # "param" | "data"
    name: str
    idxs: List[str]
    ty: tuple
    stochastic: bool               # '~' (True) or '=' deterministic (False)
    exp: tuple
    grids: List[Tuple[str, tuple, tuple]] = field(default_factory=list)  # (var, lo, hi)


@dataclass
class Model:
    hypers: List[Tuple[str, tuple]]
    decls: List[Decl]

    def decl(self, name):
        for d in self.decls:
            if d.name == name:
                return d
        return None

    @property
    def params(self):
        return [d for d in self.decls if d.kind == "param"]

    @property
    def data(self):
        return [d for d in self.decls if d.kind == "data"]


@dataclass
class Kern:
    kind: str                      # HMC | NUTS | RSlice | MWG | DiscGibbs | ConjGibbs | ESlice
    vars: List[str]
    settings: Optional[tuple] = None
    proposal: Optional[tuple] = None


class _P:
    def __init__(self, src):
        self.t = tokenize(src)
        self.i = 0

    def peek(self, k=0):
        return self.t[self.i + k]

    def next(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def expect(self, text):
        kind, val = self.next()
        if val != text:
            raise RvParseError(f"expected {text!r}, found {val!r} (token {self.i})")
        return val

    def accept(self, text):
        if self.peek()[1] == text and self.peek()[0] != "eof":
            self.i += 1
            return True
        return False

    def ident(self):
        kind, val = self.next()
        if kind != "id":
            raise RvParseError(f"expected an identifier, found {val!r}")
        return val

    # ---- types ----
    def typ(self):
        kind, val = self.next()
        if val in ("Unit", "Int", "Real"):
            return (val,)
        if val in ("Vec", "Mat"):
            return (val, self.typ())
        if val == "(":
            t = self.typ()
            self.expect(")")
            return t
        raise RvParseError(f"bad type at {val!r}")

    # ---- expressions (precedence climbing: + - < * /) ----
    def exp(self):
        lhs = self.term()
        while self.peek()[1] in ("+", "-"):
            op = self.next()[1]
            lhs = ("bin", op, lhs, self.term())
        return lhs

    def term(self):
        lhs = self.cexp()
        while self.peek()[1] in ("*", "/"):
            op = self.next()[1]
            lhs = ("bin", op, lhs, self.cexp())
        return lhs

    def exps(self, close):
        out = []
        if self.peek()[1] == close:
            return out
        out.append(self.exp())
        while self.accept(","):
            out.append(self.exp())
        return out

    def cexp(self):
        kind, val = self.peek()
        if kind in ("dist", "prim"):
            self.next()
            self.expect("(")
            args = self.exps(")")
            self.expect(")")
            return (kind, val, args)
        return self.aexp()

    def aexp(self):
        kind, val = self.next()
        if kind == "int":
            return ("lit", int(val))
        if kind == "flt":
            return ("lit", float(val))
        if val == "-" and self.peek()[0] in ("int", "flt"):
            k2, v2 = self.next()
            return ("lit", -(int(v2) if k2 == "int" else float(v2)))
        if kind == "id":
            if self.accept("["):
                idx = self.exps("]")
                self.expect("]")
                return ("proj", ("var", val), idx)
            return ("var", val)
        if val == "(":
            e = self.exp()
            self.expect(")")
            return e
        raise RvParseError(f"unexpected {val!r} in expression")

    # ---- models ----
    def model(self):
        self.expect("(")
        hypers = []
        if self.peek()[1] != ")":
            while True:
                name = self.ident()
                self.expect(":")
                hypers.append((name, self.typ()))
                if not self.accept(","):
                    break
        self.expect(")")
        self.expect("=>")
        self.expect("{")
        decls = []
        while self.peek()[1] in ("param", "data"):
            decls.append(self.decl())
        self.expect("}")
        if self.peek()[0] != "eof":
            raise RvParseError(f"trailing input at {self.peek()[1]!r}")
        return Model(hypers, decls)

    def grids(self):
        out = []
        while True:
            var = self.ident()
            self.expect("<-")
            lo = self.exp()
            self.expect("until")
            hi = self.exp()
            out.append((var, lo, hi))
            if not self.accept(","):
                return out

    def decl(self):
        kind = self.next()[1]
        name = self.ident()
        idxs = []
        ty = ("Unit",)
        if self.accept("["):
            idxs.append(self.ident())
            while self.accept(","):
                idxs.append(self.ident())
            self.expect("]")
            if self.accept(":"):
                ty = self.typ()
        elif self.accept(":"):
            ty = self.typ()
        ass = self.next()[1]
        if ass not in ("~", "="):
            raise RvParseError(f"expected '~' or '=' in declaration of {name}")
        e = self.exp()
        grids = []
        if idxs or self.peek()[1] == "for":
            self.expect("for")
            grids = self.grids()
        self.expect(";")
        return Decl(kind, name, idxs, ty, ass == "~", e, grids)

    # ---- schedules ----
    def floats(self):
        self.expect("[")
        vals = [self._num()]
        while self.accept(","):
            vals.append(self._num())
        self.expect("]")
        return tuple(vals)

    def _num(self):
        kind, val = self.next()
        neg = False
        if val == "-":
            neg = True
            kind, val = self.next()
        if kind not in ("int", "flt"):
            raise RvParseError(f"expected a number, found {val!r}")
        v = float(val)
        return -v if neg else v

    def varlist(self):
        self.expect("[")
        vs = [self.ident()]
        while self.accept(","):
            vs.append(self.ident())
        self.expect("]")
        return vs

    def kern(self):
        kind, val = self.next()
        if val in ("HMC", "RSlice", "NUTS"):
            vs = self.varlist()
            settings = self.floats() if self.peek()[1] == "[" else None
            want = 1 if val == "NUTS" else 2
            if settings is not None and len(settings) != want:
                raise RvParseError(f"{val} takes {want} setting(s)")
            return Kern(val, vs, settings)
        if val in ("DiscGibbs", "ConjGibbs", "ESlice"):
            vs = self.varlist()
            if len(vs) != 1:
                raise RvParseError(f"{val} updates one variable")
            return Kern(val, vs)
        if val == "MWG":
            vs = self.varlist()
            self.expect("~")
            prop = self.exp()
            grids = self.grids() if self.accept("for") else []
            return Kern(val, vs, None, (prop, grids))
        raise RvParseError(f"unknown kernel {val!r}")

    def schedule(self):
        ks = [self.kern()]
        while self.accept("(*)"):
            ks.append(self.kern())
        if self.peek()[0] != "eof":
            raise RvParseError(f"trailing input at {self.peek()[1]!r}")
        return ks


def parse_model(src):
    return _P(src).model()


def parse_schedule(src):
    return _P(src).schedule()
