"""Static FP64 operation count of the reference's tregn96 per refined layer
(test infrastructure: it reads the reference's Fortran source and prints
numbers; tests/test_flop_counts.py checks bench.py's constant against it,
nothing else imports it).

Counts, per layer of one (column, period) solve, the arithmetic of the
routines tregn96 runs for every layer (code/src_inv_iso_joint/tregn96.f):
  svfunc (:1559-1832), the interface contraction, which runs
    up (:1834-1990) with gettiegn (:3166-3361), varsv (:3363-3475),
    dnka_tregn (:1992-2987) and cnormc (:1475-1511), and
    down (:3561-3709) with gettiegn, varsv, hska (:3477-3559) and rnormc
    (:1513-1557);
  energy (:3777-3994) with getmat (:3996-4071), its six intijr calls
  (:4073-4245, with ffunc/gfunc/h1func/h2func :1385-1473) and getdcdh
  (:4247-4381).
The per-layer body of each routine's layer loop is counted once; the
statements outside the layer loops (run once per solve) are counted with
it, a small overestimate of the per-layer work.

Cost model (real operations): real +,-,*,/ = 1; complex+complex = 2,
complex+real = 1, complex*complex = 6, complex*real = 2, complex/complex =
11, complex/real = 2, real/complex = 7; integer arithmetic and comparisons
are free; multiplication by the literal 1 (the reference's (+1)*, (-1)*
factors) is free.  Transcendentals (exp, sqrt, log, sin, cos) are counted
apart, one each, complex exp as exp + cos + sin + 2 multiplications,
complex sqrt as 1 sqrt + 8, complex abs as 3 + 1 sqrt.  Each if / else if /
else chain is counted at its costliest branch (the general path; the
special branches are small-argument or degenerate cases), constant-trip
loops by their trip counts.

Usage: python oracle/count_tregn96_flops.py [path/to/tregn96.f]
"""
import re
import sys

SRC = "/root/reference/code/src_inv_iso_joint/tregn96.f"

# per-layer routine bodies: (routine, start line, end line)
ROUTINES = {
    "ffunc": (1385, 1408), "gfunc": (1409, 1423), "h1func": (1424, 1444),
    "h2func": (1445, 1473), "cnormc": (1475, 1511), "rnormc": (1513, 1557),
    "svfunc": (1559, 1832), "up": (1834, 1990), "dnka_tregn": (1992, 2987),
    "gettiegn": (3166, 3361), "varsv": (3363, 3475), "hska": (3477, 3559),
    "down": (3561, 3709), "energy": (3777, 3994), "getmat": (3996, 4071),
    "intijr": (4073, 4245), "getdcdh": (4247, 4381),
}
INTRINSIC_FREE = {"dcmplx", "cmplx", "dble", "dreal", "real", "dimag", "aimag", "sngl", "float",
                  "int", "nint", "dfloat", "conjg", "dconjg", "max", "min", "dmax1", "dmin1",
                  "sign", "dsign", "mod", "iabs", "abs_r"}
TRANS_REAL = {"exp", "dexp", "sqrt", "dsqrt", "log", "dlog", "sin", "cos", "dsin", "dcos",
              "atan", "datan", "atan2", "datan2", "log10", "alog"}


def read_statements(lines, a, b):
    """Joined fixed-form statements of lines a..b (1-based), comments removed."""
    out = []
    for ln in lines[a - 1:b]:
        if not ln.strip() or ln[0] in "cC*!":
            continue
        body = ln.rstrip("\n")
        cont = len(body) > 5 and body[5] not in " 0"
        text = body[6:] if len(body) > 6 else ""
        text = re.sub(r"!.*$", "", text)
        if cont and out:
            out[-1] += " " + text.strip()
        else:
            lab = body[:5].strip()
            out.append((lab + " " if lab else "") + text.strip())
    return out


class Types:
    def __init__(self):
        self.t = {}
        self.arr = set()

    def declare(self, stmt):
        s = stmt.lower()
        m = re.match(r"(complex\*16|double complex|complex|real\*8|double precision|real\*4|"
                     r"real|integer|logical)\s+(.*)$", s)
        if not m or s.startswith("real function") or "function" in s.split()[:3]:
            return False
        kind = {"complex*16": "c", "double complex": "c", "complex": "c", "integer": "i",
                "logical": "l"}.get(m.group(1), "r")
        for name, dims in self._names(m.group(2)):
            self.t[name] = kind
            if dims:
                self.arr.add(name)
        return True

    @staticmethod
    def _names(s):
        depth, cur, parts = 0, "", []
        for ch in s:
            if ch == "(":
                depth += 1
            elif ch == ")":
                depth -= 1
            if ch == "," and depth == 0:
                parts.append(cur)
                cur = ""
            else:
                cur += ch
        parts.append(cur)
        for p in parts:
            p = p.strip()
            if not p:
                continue
            m = re.match(r"([a-z_][a-z0-9_]*)\s*(\(.*\))?", p)
            if m:
                yield m.group(1), m.group(2)

    def of(self, name):
        if name in self.t:
            return self.t[name]
        return "i" if name[0] in "ijklmn" else "r"


TOK = re.compile(r"\s*(?:(\d+\.\d*(?:[de][+-]?\d+)?|\.\d+(?:[de][+-]?\d+)?|\d+[de][+-]?\d+|\d+)|"
                 r"([a-z_][a-z0-9_]*)|(\*\*|[-+*/(),]))")


class Cost:
    def __init__(self, flops=0.0, trans=0.0):
        self.flops, self.trans = flops, trans

    def __iadd__(self, o):
        self.flops += o.flops
        self.trans += o.trans
        return self

    def scaled(self, k):
        return Cost(self.flops * k, self.trans * k)

    def total(self):
        return self.flops + self.trans


class Expr:
    """Pratt parser over a Fortran expression: returns (type, literal, Cost)."""

    def __init__(self, text, types, calls):
        self.toks = []
        pos = 0
        text = text.lower()
        while pos < len(text):
            m = TOK.match(text, pos)
            if not m or m.end() == pos:
                raise ValueError("cannot tokenize: " + text[pos:pos + 30])
            pos = m.end()
            if m.group(1):
                self.toks.append(("num", m.group(1)))
            elif m.group(2):
                self.toks.append(("id", m.group(2)))
            elif m.group(3):
                self.toks.append(("op", m.group(3)))
        self.i = 0
        self.types = types
        self.calls = calls
        self.cost = Cost()

    def peek(self):
        return self.toks[self.i] if self.i < len(self.toks) else (None, None)

    def next(self):
        t = self.peek()
        self.i += 1
        return t

    def parse(self):
        t = self.expr(0)
        return t

    BP = {"+": 10, "-": 10, "*": 20, "/": 20, "**": 30}

    def expr(self, rbp):
        left = self.unary()
        while True:
            k, v = self.peek()
            if k != "op" or v not in self.BP or self.BP[v] <= rbp:
                break
            self.next()
            bp = self.BP[v] - (1 if v == "**" else 0)
            right = self.expr(bp)
            left = self.binop(v, left, right)
        return left

    def unary(self):
        k, v = self.peek()
        if k == "op" and v in "+-":
            self.next()
            t = self.expr(25)
            return (t[0], t[1], t[2])  # negation is free
        return self.atom()

    def atom(self):
        k, v = self.next()
        if k == "num":
            return ("i" if re.fullmatch(r"\d+", v) else "r", v, None)
        if k == "op" and v == "(":
            t = self.expr(0)
            self.expect(")")
            return (t[0], None, None)
        if k == "id":
            if self.peek() == ("op", "("):
                self.next()
                args = []
                if self.peek() != ("op", ")"):
                    while True:
                        args.append(self.expr(0))
                        if self.peek() == ("op", ","):
                            self.next()
                            continue
                        break
                self.expect(")")
                return self.call(v, args)
            return (self.types.of(v), None, None)
        raise ValueError("unexpected token %r" % (v,))

    def expect(self, s):
        k, v = self.next()
        if v != s:
            raise ValueError("expected %s, got %r" % (s, v))

    def call(self, name, args):
        if name in self.types.arr:
            return (self.types.of(name), None, None)
        if name in self.calls:
            self.calls[name] += 1
            return (self.types.of(name), None, None)
        at = args[0][0] if args else "r"
        if name in ("dcmplx", "cmplx"):
            return ("c", None, None)
        if name in INTRINSIC_FREE:
            return ("i" if name in ("int", "nint", "mod", "iabs") else "r", None, None)
        if name in ("abs", "dabs", "cdabs", "zabs"):
            if at == "c" or name in ("cdabs", "zabs"):
                self.cost += Cost(3, 1)
            return ("r", None, None)
        if name in ("cdsqrt", "zsqrt") or (name == "sqrt" and at == "c"):
            self.cost += Cost(8, 1)
            return ("c", None, None)
        if name in ("cdexp", "zexp") or (name == "exp" and at == "c"):
            self.cost += Cost(2, 3)
            return ("c", None, None)
        if name in TRANS_REAL:
            self.cost += Cost(0, 1)
            return ("r", None, None)
        return (self.types.of(name), None, None)

    def binop(self, op, a, b):
        ta, la, _ = a
        tb, lb, _ = b
        if ta == "i" and tb == "i":
            return ("i", None, None)
        if op == "**":
            p = lb if lb and re.fullmatch(r"\d+", lb) else None
            n = int(p) - 1 if p else 1
            if not p:
                self.cost += Cost(0, 2)  # exp(y log x)
            else:
                self.cost += Cost(n * (6 if ta == "c" else 1), 0)
            return (ta if ta != "i" else "r", None, None)
        ta = "r" if ta == "i" else ta
        tb = "r" if tb == "i" else tb
        if op == "*" and ((la == "1") or (lb == "1")):
            return (tb if la == "1" else ta, None, None)
        c = {("+", "r", "r"): 1, ("+", "c", "c"): 2, ("+", "c", "r"): 1, ("+", "r", "c"): 1,
             ("-", "r", "r"): 1, ("-", "c", "c"): 2, ("-", "c", "r"): 1, ("-", "r", "c"): 1,
             ("*", "r", "r"): 1, ("*", "c", "c"): 6, ("*", "c", "r"): 2, ("*", "r", "c"): 2,
             ("/", "r", "r"): 1, ("/", "c", "c"): 11, ("/", "c", "r"): 2, ("/", "r", "c"): 7}
        self.cost += Cost(c[(op, ta, tb)], 0)
        return ("c" if "c" in (ta, tb) else "r", None, None)


def split_assign(s):
    depth = 0
    for i, ch in enumerate(s):
        if ch == "(":
            depth += 1
        elif ch == ")":
            depth -= 1
        elif ch == "=" and depth == 0 and s[i - 1:i + 1] != "==" and s[i:i + 2] != "==":
            return s[:i], s[i + 1:]
    return None


class Routine:
    def __init__(self, name, stmts):
        self.name = name
        self.types = Types()
        self.exec = []
        for s in stmts:
            body = re.sub(r"^\d+\s+", "", s)
            low = body.lower()
            if re.match(r"(implicit|common|parameter|external|save|data|equivalence|"
                        r"character|dimension)\b", low):
                continue
            if self.types.declare(body):
                continue
            if re.match(r"(subroutine|(complex\*16|real\*8|real\*4|double precision|real|integer)?"
                        r"\s*function)\b", low):
                continue
            self.exec.append(s)


def count(routines, name, memo):
    """Cost of one call of routine `name`, with per-layer loops counted once."""
    if name in memo:
        return memo[name]
    r = routines[name]
    # block structure: a stack of frames; an if-chain keeps its branches' costs
    root = {"kind": "seq", "cost": Cost(), "mult": 1}
    stack = [root]
    do_labels = []  # (label, frame)

    def add(c):
        stack[-1]["cost"] += c

    def close_if():
        fr = stack.pop()
        branches = fr["branches"] + [fr["cost"]]
        best = max(branches, key=lambda c: c.total())
        add(best)

    def close_do():
        fr = stack.pop()
        add(fr["cost"].scaled(fr["mult"]))

    for s in r.exec:
        lab_m = re.match(r"^(\d+)\s+(.*)$", s)
        lab = lab_m.group(1) if lab_m else None
        body = lab_m.group(2) if lab_m else s
        low = body.lower().strip()
        calls = {k: 0 for k in routines}
        cost = Cost()

        def expr_cost(text):
            e = Expr(text, r.types, calls)
            e.parse()
            return e.cost

        if re.match(r"if\s*\(.*\)\s*then$", low):
            stack.append({"kind": "if", "cost": Cost(), "branches": []})
        elif re.match(r"else\s*if\s*\(.*\)\s*then$", low) or low == "else":
            fr = stack[-1]
            fr["branches"].append(fr["cost"])
            fr["cost"] = Cost()
        elif low in ("endif", "end if"):
            close_if()
        elif low.startswith("do "):
            m = re.match(r"do\s+(\d+)?\s*,?\s*([a-z0-9_]+)\s*=\s*(.*)$", low)
            lo_hi = [x.strip() for x in Types._names(m.group(3))] if False else None
            parts = m.group(3).split(",")
            try:
                trip = int(parts[1]) - int(parts[0]) + 1 if len(parts) == 2 else None
            except ValueError:
                trip = None
            if trip is None:
                # a loop over layers (or another data-sized loop): one trip per layer
                trip = 1
            stack.append({"kind": "do", "cost": Cost(), "mult": max(trip, 0)})
            if m.group(1):
                do_labels.append(m.group(1))
            else:
                do_labels.append(None)
        elif low in ("enddo", "end do"):
            do_labels.pop()
            close_do()
        elif low.startswith("call "):
            m = re.match(r"call\s+([a-z0-9_]+)\s*(\((.*)\))?$", low)
            callee = m.group(1)
            if m.group(3):
                for a in Types._names(m.group(3)):
                    pass
            if callee in routines:
                add(count(routines, callee, memo))
        elif re.match(r"if\s*\(", low):
            # logical if with a statement: count the statement (condition free)
            depth, i = 0, low.index("(")
            for j in range(i, len(low)):
                depth += low[j] == "("
                depth -= low[j] == ")"
                if depth == 0:
                    break
            rest = body[j + 1:].strip()
            sa = split_assign(rest)
            if sa:
                cost += expr_cost(sa[1])
            add(cost)
        elif re.match(r"(return|continue|go\s*to|goto|end$|stop)", low):
            pass
        else:
            sa = split_assign(body)
            if sa:
                cost += expr_cost(sa[1])
            add(cost)
        for k, n in calls.items():
            if n:
                add(count(routines, k, memo).scaled(n))
        # labelled loop ends
        while lab and do_labels and do_labels[-1] == lab:
            do_labels.pop()
            close_do()
    while len(stack) > 1:
        fr = stack[-1]
        if fr["kind"] == "if":
            close_if()
        else:
            close_do()
    memo[name] = root["cost"]
    return root["cost"]


def per_layer(path=SRC):
    """(total Cost per layer, {routine: Cost of one call}) of tregn96 at path."""
    lines = open(path).read().split("\n")
    routines = {n: Routine(n, read_statements(lines, a, b)) for n, (a, b) in ROUTINES.items()}
    for fn in ("ffunc", "gfunc", "h1func", "h2func", "intijr"):
        for r in routines.values():
            r.types.t.setdefault(fn, "c" if fn != "intijr" else "r")
    memo = {}
    tot = Cost()
    for n in ("svfunc", "energy"):
        tot += count(routines, n, memo)
    for n in ROUTINES:
        count(routines, n, memo)
    return tot, memo


def main():
    tot, memo = per_layer(sys.argv[1] if len(sys.argv) > 1 else SRC)
    for n in ("svfunc", "energy", "up", "down", "gettiegn", "varsv", "dnka_tregn", "hska",
              "intijr", "getdcdh"):
        c = memo[n]
        print("%-10s flops %6.0f  transcendentals %4.0f (one call)" % (n, c.flops, c.trans))
    print("per layer (svfunc + energy): flops %.0f + transcendentals %.0f = %.0f operations"
          % (tot.flops, tot.trans, tot.total()))


if __name__ == "__main__":
    main()
