"""A small interpreter for the subset of Lua 5.4 that Aether's scripts use.

Aether runs its simulation scripts, structure scripts and material library
files (``mat_lib/*.lua``) through liblua 5.4 (Lua 5.4.4 in its build; not vendored
in the reference and not installed here).  This module restates the language
from the Lua 5.4 reference manual for what those scripts need: literals
(decimal and hexadecimal numbers with the integer/float split, short and long
strings, comments), locals and globals, multiple assignment and returns,
``if``/``while``/``repeat``/numeric and generic ``for``/``do``/``break``,
functions and closures (``function f(...)``, ``local function``,
``function t.a:b()``, varargs), tables with constructors, field and method
access (``obj:method(...)``), and the operators with the manual's precedence
(``or`` < ``and`` < comparison < ``|`` < ``~`` < ``&`` < shifts < ``..``
(right) < ``+ -`` < ``* / // %`` < unary < ``^`` (right)).  Libraries:
``math``, ``string.format/rep/sub/len/upper/lower``, ``table.insert/unpack``,
``print``, ``tostring``, ``tonumber``, ``type``, ``pairs``, ``ipairs``,
``select``, ``error``, ``assert``, ``pcall``.  Not provided: coroutines,
metatables set from scripts, ``goto``, the io/os libraries.

Host objects (modes, materials, structures, sensors) are Python objects with
a ``lua_methods`` dict; ``obj:name(args)`` calls ``lua_methods[name](obj,
*args)``.
"""
from __future__ import annotations

import math
import re

__all__ = ["LuaError", "LuaTable", "Interpreter", "tostring", "to_integer", "to_number"]


class LuaError(RuntimeError):
    pass


class LuaTable:
    """A Lua table: a dict keyed by Lua values (floats with integral values
    are normalised to ints, as Lua does for table keys)."""

    def __init__(self, items=None):
        self.d = {}
        if items:
            for k, v in items.items():
                self[k] = v

    @staticmethod
    def _key(k):
        if isinstance(k, float) and k.is_integer():
            return int(k)
        return k

    def __getitem__(self, k):
        return self.d.get(self._key(k))

    def __setitem__(self, k, v):
        if k is None:
            raise LuaError("table index is nil")
        k = self._key(k)
        if v is None:
            self.d.pop(k, None)
        else:
            self.d[k] = v

    def length(self):
        """A border of the table (#t): the largest n with t[1..n] non-nil."""
        n = 0
        while (n + 1) in self.d:
            n += 1
        return n

    def array(self):
        return [self.d[i] for i in range(1, self.length() + 1)]


# ---------------------------------------------------------------- values

def lua_type(v):
    if v is None:
        return "nil"
    if isinstance(v, bool):
        return "boolean"
    if isinstance(v, (int, float)):
        return "number"
    if isinstance(v, str):
        return "string"
    if isinstance(v, LuaTable):
        return "table"
    if isinstance(v, LuaFunction) or callable(v):
        return "function"
    return "userdata"


def truthy(v):
    return v is not None and v is not False


def tostring(v):
    """Lua 5.4's tostring for numbers: integers as such, floats with "%.14g"
    and a ".0" suffix when the result would read as an integer."""
    if v is None:
        return "nil"
    if v is True:
        return "true"
    if v is False:
        return "false"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        if math.isinf(v):
            return "inf" if v > 0 else "-inf"
        if math.isnan(v):
            return "nan" if math.copysign(1.0, v) < 0 else "-nan"
        s = "%.14g" % v
        if re.fullmatch(r"-?\d+", s):
            s += ".0"
        return s
    if isinstance(v, str):
        return v
    if isinstance(v, LuaTable):
        return "table: 0x%08x" % (id(v) & 0xFFFFFFFF)
    return "%s: 0x%08x" % (lua_type(v), id(v) & 0xFFFFFFFF)


_NUM_RE = re.compile(r"\s*([+-]?)(0[xX][0-9a-fA-F]*(\.[0-9a-fA-F]*)?([pP][+-]?\d+)?|"
                     r"(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)\s*$")


def str_to_number(s):
    m = _NUM_RE.match(s)
    if not m:
        return None
    sign, body = m.group(1), m.group(2)
    neg = sign == "-"
    if body[:2] in ("0x", "0X"):
        if "." in body or "p" in body or "P" in body:
            v = float.fromhex(body)
        else:
            v = int(body, 16)
            v = (v + 2 ** 63) % 2 ** 64 - 2 ** 63
    elif re.fullmatch(r"\d+", body):
        v = int(body)
        if v >= 2 ** 63:
            v = float(body)
    else:
        v = float(body)
    return -v if neg else v


def to_number(v):
    if isinstance(v, bool):
        return None
    if isinstance(v, (int, float)):
        return v
    if isinstance(v, str):
        return str_to_number(v)
    return None


def to_integer(v):
    """lua_tointeger: a number (or numeric string) with an integral value, else None."""
    n = to_number(v)
    if n is None:
        return None
    if isinstance(n, float):
        return int(n) if n.is_integer() else None
    return n


def arith_num(v, op):
    n = to_number(v)
    if n is None:
        raise LuaError(f"attempt to perform arithmetic ({op}) on a {lua_type(v)} value")
    return n


def fdiv(a, b):
    a, b = float(a), float(b)
    if b == 0.0:
        if a == 0.0 or math.isnan(a):
            return math.nan
        return math.copysign(math.inf, a) * math.copysign(1.0, b)
    return a / b


def lua_mod(a, b):
    if isinstance(a, int) and isinstance(b, int):
        if b == 0:
            raise LuaError("attempt to perform 'n%%0'")
        return a % b
    a, b = float(a), float(b)
    if b == 0.0:
        return math.nan
    r = math.fmod(a, b)
    if r != 0.0 and (r < 0) != (b < 0):
        r += b
    return r


def lua_idiv(a, b):
    if isinstance(a, int) and isinstance(b, int):
        if b == 0:
            raise LuaError("attempt to perform 'n//0'")
        return a // b
    return math.floor(fdiv(a, b)) * 1.0 if b != 0 else fdiv(a, b)


def arith(op, a, b):
    if op == "..":
        for v in (a, b):
            if not isinstance(v, (str, int, float)) or isinstance(v, bool):
                raise LuaError(f"attempt to concatenate a {lua_type(v)} value")
        return tostring(a) + tostring(b)
    a = arith_num(a, op)
    b = arith_num(b, op)
    both_int = isinstance(a, int) and isinstance(b, int)
    if op == "+":
        return a + b if both_int else float(a) + float(b)
    if op == "-":
        return a - b if both_int else float(a) - float(b)
    if op == "*":
        return a * b if both_int else float(a) * float(b)
    if op == "/":
        return fdiv(a, b)
    if op == "%":
        return lua_mod(a, b)
    if op == "//":
        return lua_idiv(a, b)
    if op == "^":
        try:
            return math.pow(float(a), float(b))
        except (OverflowError, ValueError):
            return math.inf if float(a) > 0 else math.nan
    ia, ib = to_integer(a), to_integer(b)
    if ia is None or ib is None:
        raise LuaError("number has no integer representation")
    if op == "&":
        return ia & ib
    if op == "|":
        return ia | ib
    if op == "~":
        return ia ^ ib
    if op == "<<":
        return (ia << ib) if ib >= 0 else (ia >> -ib)
    if op == ">>":
        return (ia >> ib) if ib >= 0 else (ia << -ib)
    raise LuaError(f"unknown operator {op}")


def compare(op, a, b):
    if op == "==":
        return raw_equal(a, b)
    if op == "~=":
        return not raw_equal(a, b)
    num = (isinstance(a, (int, float)) and not isinstance(a, bool)
           and isinstance(b, (int, float)) and not isinstance(b, bool))
    if not (num or (isinstance(a, str) and isinstance(b, str))):
        raise LuaError(f"attempt to compare {lua_type(a)} with {lua_type(b)}")
    return {"<": a < b, "<=": a <= b, ">": a > b, ">=": a >= b}[op]


def raw_equal(a, b):
    if isinstance(a, bool) or isinstance(b, bool):
        return a is b
    if isinstance(a, (int, float)) and isinstance(b, (int, float)):
        return a == b
    if isinstance(a, str) and isinstance(b, str):
        return a == b
    return a is b


# ---------------------------------------------------------------- lexer

KEYWORDS = {"and", "break", "do", "else", "elseif", "end", "false", "for", "function", "goto", "if", "in",
            "local", "nil", "not", "or", "repeat", "return", "then", "true", "until", "while"}
SYMBOLS = ["...", "..", "==", "~=", "<=", ">=", "//", "::", "<<", ">>",
           "+", "-", "*", "/", "%", "^", "#", "&", "~", "|", "<", ">", "=", "(", ")", "{", "}", "[", "]",
           ";", ":", ",", "."]
ESCAPES = {"n": "\n", "t": "\t", "r": "\r", "a": "\a", "b": "\b", "f": "\f", "v": "\v", "\\": "\\",
           '"': '"', "'": "'", "\n": "\n"}


def tokenize(src, chunk="script"):
    toks = []
    i, n, line = 0, len(src), 1
    if src.startswith("#!"):
        while i < n and src[i] != "\n":
            i += 1

    def err(msg):
        raise LuaError(f"{chunk}:{line}: {msg}")

    def long_bracket(j):
        m = re.compile(r"\[(=*)\[").match(src, j)
        if not m:
            return None
        close = "]" + m.group(1) + "]"
        end = src.find(close, m.end())
        if end < 0:
            err("unfinished long string/comment")
        body = src[m.end():end]
        if body.startswith("\r\n"):
            body = body[2:]
        elif body.startswith("\n"):
            body = body[1:]
        return body, end + len(close)

    while i < n:
        c = src[i]
        if c == "\n":
            line += 1
            i += 1
            continue
        if c in " \t\r\f\v":
            i += 1
            continue
        if src.startswith("--", i):
            lb = long_bracket(i + 2) if src.startswith("[", i + 2) else None
            if lb:
                line += src.count("\n", i, lb[1])
                i = lb[1]
            else:
                while i < n and src[i] != "\n":
                    i += 1
            continue
        if c.isalpha() or c == "_":
            j = i
            while j < n and (src[j].isalnum() or src[j] == "_"):
                j += 1
            w = src[i:j]
            toks.append(("kw" if w in KEYWORDS else "name", w, line))
            i = j
            continue
        if c.isdigit() or (c == "." and i + 1 < n and src[i + 1].isdigit()):
            if src.startswith(("0x", "0X"), i):
                m = re.compile(r"0[xX][0-9a-fA-F]*(\.[0-9a-fA-F]*)?([pP][+-]?\d+)?").match(src, i)
            else:
                m = re.compile(r"(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?").match(src, i)
            v = str_to_number(m.group(0))
            if v is None:
                err(f"malformed number near '{m.group(0)}'")
            toks.append(("num", v, line))
            i = m.end()
            continue
        if c in "\"'":
            q = c
            j = i + 1
            out = []
            while True:
                if j >= n or src[j] == "\n":
                    err("unfinished string")
                ch = src[j]
                if ch == q:
                    j += 1
                    break
                if ch == "\\":
                    e = src[j + 1] if j + 1 < n else ""
                    if e in ESCAPES:
                        out.append(ESCAPES[e])
                        if e == "\n":
                            line += 1
                        j += 2
                    elif e.isdigit():
                        m = re.compile(r"\d{1,3}").match(src, j + 1)
                        out.append(chr(int(m.group(0))))
                        j = m.end()
                    elif e == "x":
                        out.append(chr(int(src[j + 2:j + 4], 16)))
                        j += 4
                    elif e == "z":
                        j += 2
                        while j < n and src[j] in " \t\r\n\f\v":
                            line += src[j] == "\n"
                            j += 1
                    else:
                        err(f"invalid escape sequence '\\{e}'")
                    continue
                out.append(ch)
                j += 1
            toks.append(("str", "".join(out), line))
            i = j
            continue
        if c == "[" and src.startswith(("[[", "[="), i):
            lb = long_bracket(i)
            if lb:
                toks.append(("str", lb[0], line))
                line += src.count("\n", i, lb[1])
                i = lb[1]
                continue
        for s in SYMBOLS:
            if src.startswith(s, i):
                toks.append(("op", s, line))
                i += len(s)
                break
        else:
            err(f"unexpected symbol near '{c}'")
    toks.append(("eof", None, line))
    return toks


# ---------------------------------------------------------------- parser (to nested tuples)

BINPRI = {"or": (1, 1), "and": (2, 2), "<": (3, 3), ">": (3, 3), "<=": (3, 3), ">=": (3, 3), "~=": (3, 3),
          "==": (3, 3), "|": (4, 4), "~": (5, 5), "&": (6, 6), "<<": (7, 7), ">>": (7, 7), "..": (9, 8),
          "+": (10, 10), "-": (10, 10), "*": (11, 11), "/": (11, 11), "//": (11, 11), "%": (11, 11),
          "^": (14, 13)}
UNARY_PRI = 12


class Parser:
    def __init__(self, toks, chunk):
        self.t, self.p, self.chunk = toks, 0, chunk

    def peek(self, k=0):
        return self.t[self.p + k]

    def next(self):
        tok = self.t[self.p]
        self.p += 1
        return tok

    def check(self, kind, val=None):
        tok = self.peek()
        return tok[0] == kind and (val is None or tok[1] == val)

    def accept(self, kind, val=None):
        if self.check(kind, val):
            return self.next()
        return None

    def expect(self, kind, val=None):
        if not self.check(kind, val):
            tok = self.peek()
            raise LuaError(f"{self.chunk}:{tok[2]}: '{val or kind}' expected near '{tok[1]}'")
        return self.next()

    def name(self):
        return self.expect("name")[1]

    # statements
    def block(self):
        stats = []
        while True:
            tok = self.peek()
            if tok[0] == "eof" or (tok[0] == "kw" and tok[1] in ("end", "else", "elseif", "until")):
                break
            if tok[0] == "kw" and tok[1] == "return":
                self.next()
                exprs = []
                if not (self.check("eof") or self.check("op", ";") or
                        (self.peek()[0] == "kw" and self.peek()[1] in ("end", "else", "elseif", "until"))):
                    exprs = self.exprlist()
                self.accept("op", ";")
                stats.append(("return", exprs, tok[2]))
                break
            s = self.statement()
            if s is not None:
                stats.append(s)
        return stats

    def statement(self):
        tok = self.peek()
        line = tok[2]
        if self.accept("op", ";"):
            return None
        if tok[0] == "kw":
            k = tok[1]
            if k == "if":
                self.next()
                arms = []
                cond = self.expr()
                self.expect("kw", "then")
                arms.append((cond, self.block()))
                other = None
                while True:
                    if self.accept("kw", "elseif"):
                        c = self.expr()
                        self.expect("kw", "then")
                        arms.append((c, self.block()))
                    elif self.accept("kw", "else"):
                        other = self.block()
                        self.expect("kw", "end")
                        break
                    else:
                        self.expect("kw", "end")
                        break
                return ("if", arms, other, line)
            if k == "while":
                self.next()
                c = self.expr()
                self.expect("kw", "do")
                b = self.block()
                self.expect("kw", "end")
                return ("while", c, b, line)
            if k == "do":
                self.next()
                b = self.block()
                self.expect("kw", "end")
                return ("do", b, line)
            if k == "repeat":
                self.next()
                b = self.block()
                self.expect("kw", "until")
                return ("repeat", b, self.expr(), line)
            if k == "for":
                self.next()
                n1 = self.name()
                if self.accept("op", "="):
                    a = self.expr()
                    self.expect("op", ",")
                    b = self.expr()
                    c = self.expr() if self.accept("op", ",") else None
                    self.expect("kw", "do")
                    body = self.block()
                    self.expect("kw", "end")
                    return ("fornum", n1, a, b, c, body, line)
                names = [n1]
                while self.accept("op", ","):
                    names.append(self.name())
                self.expect("kw", "in")
                exprs = self.exprlist()
                self.expect("kw", "do")
                body = self.block()
                self.expect("kw", "end")
                return ("forin", names, exprs, body, line)
            if k == "function":
                self.next()
                target = ("name", self.name(), line)
                method = False
                while self.check("op", ".") or self.check("op", ":"):
                    sep = self.next()[1]
                    key = ("const", self.name())
                    target = ("index", target, key, line)
                    if sep == ":":
                        method = True
                        break
                f = self.funcbody(line, method)
                return ("assign", [target], [f], line)
            if k == "local":
                self.next()
                if self.accept("kw", "function"):
                    n = self.name()
                    return ("localfunc", n, self.funcbody(line, False), line)
                names = [self.name()]
                self.attrib()
                while self.accept("op", ","):
                    names.append(self.name())
                    self.attrib()
                exprs = self.exprlist() if self.accept("op", "=") else []
                return ("local", names, exprs, line)
            if k == "break":
                self.next()
                return ("break", line)
            if k == "goto":
                raise LuaError(f"{self.chunk}:{line}: goto is not supported")
        if self.accept("op", "::"):
            raise LuaError(f"{self.chunk}:{line}: labels are not supported")
        e = self.suffixedexp()
        if self.check("op", "=") or self.check("op", ","):
            targets = [e]
            while self.accept("op", ","):
                targets.append(self.suffixedexp())
            self.expect("op", "=")
            for t in targets:
                if t[0] not in ("name", "index"):
                    raise LuaError(f"{self.chunk}:{line}: syntax error near '='")
            return ("assign", targets, self.exprlist(), line)
        if e[0] not in ("call", "method"):
            raise LuaError(f"{self.chunk}:{line}: syntax error (statement expected)")
        return ("exprstat", e, line)

    def attrib(self):
        if self.accept("op", "<"):
            self.name()
            self.expect("op", ">")

    def funcbody(self, line, method):
        self.expect("op", "(")
        params, vararg = (["self"] if method else []), False
        if not self.check("op", ")"):
            while True:
                if self.accept("op", "..."):
                    vararg = True
                    break
                params.append(self.name())
                if not self.accept("op", ","):
                    break
        self.expect("op", ")")
        body = self.block()
        self.expect("kw", "end")
        return ("function", params, vararg, body, line)

    def exprlist(self):
        out = [self.expr()]
        while self.accept("op", ","):
            out.append(self.expr())
        return out

    def primaryexp(self):
        tok = self.peek()
        if tok[0] == "name":
            self.next()
            return ("name", tok[1], tok[2])
        if self.accept("op", "("):
            e = self.expr()
            self.expect("op", ")")
            return ("paren", e)
        raise LuaError(f"{self.chunk}:{tok[2]}: unexpected symbol near '{tok[1]}'")

    def suffixedexp(self):
        e = self.primaryexp()
        while True:
            tok = self.peek()
            line = tok[2]
            if self.accept("op", "."):
                e = ("index", e, ("const", self.name()), line)
            elif self.accept("op", "["):
                k = self.expr()
                self.expect("op", "]")
                e = ("index", e, k, line)
            elif self.accept("op", ":"):
                n = self.name()
                e = ("method", e, n, self.callargs(), line)
            elif tok[0] == "str" or self.check("op", "(") or self.check("op", "{"):
                e = ("call", e, self.callargs(), line)
            else:
                return e

    def callargs(self):
        tok = self.peek()
        if tok[0] == "str":
            self.next()
            return [("const", tok[1])]
        if self.check("op", "{"):
            return [self.table()]
        self.expect("op", "(")
        if self.accept("op", ")"):
            return []
        args = self.exprlist()
        self.expect("op", ")")
        return args

    def table(self):
        line = self.expect("op", "{")[2]
        items = []
        while not self.check("op", "}"):
            if self.check("op", "["):
                self.next()
                k = self.expr()
                self.expect("op", "]")
                self.expect("op", "=")
                items.append(("kv", k, self.expr()))
            elif self.peek()[0] == "name" and self.peek(1)[0] == "op" and self.peek(1)[1] == "=":
                k = self.name()
                self.next()
                items.append(("kv", ("const", k), self.expr()))
            else:
                items.append(("pos", self.expr()))
            if not (self.accept("op", ",") or self.accept("op", ";")):
                break
        self.expect("op", "}")
        return ("table", items, line)

    def simpleexp(self):
        tok = self.peek()
        if tok[0] == "num":
            self.next()
            return ("const", tok[1])
        if tok[0] == "str":
            self.next()
            return ("const", tok[1])
        if tok[0] == "kw" and tok[1] in ("nil", "true", "false"):
            self.next()
            return ("const", {"nil": None, "true": True, "false": False}[tok[1]])
        if self.accept("op", "..."):
            return ("vararg", tok[2])
        if self.check("op", "{"):
            return self.table()
        if tok[0] == "kw" and tok[1] == "function":
            self.next()
            return self.funcbody(tok[2], False)
        return self.suffixedexp()

    def expr(self, limit=0):
        tok = self.peek()
        if (tok[0] == "kw" and tok[1] == "not") or (tok[0] == "op" and tok[1] in ("-", "#", "~")):
            self.next()
            e = ("unop", tok[1], self.expr(UNARY_PRI), tok[2])
        else:
            e = self.simpleexp()
        while True:
            tok = self.peek()
            op = tok[1] if tok[0] in ("op", "kw") else None
            if op not in BINPRI or BINPRI[op][0] <= limit:
                return e
            self.next()
            rhs = self.expr(BINPRI[op][1])
            e = ("binop", op, e, rhs, tok[2])


# ---------------------------------------------------------------- evaluator

class _Break(Exception):
    pass


class _Return(Exception):
    def __init__(self, vals):
        self.vals = vals


class Scope:
    __slots__ = ("vars", "parent")

    def __init__(self, parent=None):
        self.vars = {}
        self.parent = parent

    def find(self, name):
        s = self
        while s is not None:
            if name in s.vars:
                return s
            s = s.parent
        return None


class LuaFunction:
    def __init__(self, interp, node, scope):
        self.interp, self.node, self.scope = interp, node, scope

    def __call__(self, *args):
        return self.interp.call_lua(self, list(args))


def _callee(e):
    if e[0] == "name":
        return f"global '{e[1]}'"
    if e[0] == "index" and e[2][0] == "const":
        return f"field '{e[2][1]}'"
    return "expression"


def _ret(v):
    """Normalises a Python host function's result to a list of values."""
    if v is None:
        return []
    if isinstance(v, tuple):
        return list(v)
    return [v]


class Interpreter:
    """One Lua state: a global table plus the standard functions."""

    def __init__(self, chunk="script", out=None):
        self.chunk = chunk
        self.out = out if out is not None else []
        self.G = {}
        self._stdlib()

    # ---- running
    def execute(self, source, chunk=None):
        chunk = chunk or self.chunk
        tree = Parser(tokenize(source, chunk), chunk).block()
        fn = LuaFunction(self, ("function", [], True, tree, 0), None)
        return self.call_lua(fn, [])

    def call(self, f, args):
        if isinstance(f, LuaFunction):
            return self.call_lua(f, args)
        if callable(f) and not isinstance(f, (LuaTable, str)):
            return _ret(f(*args))
        raise LuaError(f"attempt to call a {lua_type(f)} value")

    def call_lua(self, f, args):
        _, params, vararg, body, _line = f.node
        sc = Scope(f.scope)
        for i, p in enumerate(params):
            sc.vars[p] = args[i] if i < len(args) else None
        sc.vars["..."] = args[len(params):] if vararg else []
        try:
            self.run_block(body, sc)
        except _Return as r:
            return r.vals
        return []

    def run_block(self, stats, scope):
        for s in stats:
            self.stat(s, scope)

    # ---- variables
    def get_var(self, name, scope):
        s = scope.find(name) if scope is not None else None
        if s is not None:
            return s.vars[name]
        return self.G.get(name)

    def set_var(self, name, value, scope):
        s = scope.find(name) if scope is not None else None
        if s is not None:
            s.vars[name] = value
        elif value is None:
            self.G.pop(name, None)
        else:
            self.G[name] = value

    def index(self, obj, key, line=0):
        if isinstance(obj, LuaTable):
            return obj[key]
        if isinstance(obj, str):
            return self.G["string"][key]
        methods = getattr(obj, "lua_methods", None)
        if methods is not None:
            if key in methods:
                return methods[key]
            raise LuaError(f"{self.chunk}:{line}: method '{key}' is not provided")
        raise LuaError(f"{self.chunk}:{line}: attempt to index a {lua_type(obj)} value")

    def setindex(self, obj, key, value, line=0):
        if isinstance(obj, LuaTable):
            obj[key] = value
            return
        raise LuaError(f"{self.chunk}:{line}: attempt to index a {lua_type(obj)} value")

    # ---- statements
    def stat(self, s, sc):
        k = s[0]
        if k == "exprstat":
            self.multi(s[1], sc)
        elif k == "local":
            vals = self.explist(s[2], sc)
            for i, n in enumerate(s[1]):
                sc.vars[n] = vals[i] if i < len(vals) else None
        elif k == "assign":
            vals = self.explist(s[2], sc)
            targets = []
            for t in s[1]:
                if t[0] == "index":
                    targets.append((t, self.eval(t[1], sc), self.eval(t[2], sc)))
                else:
                    targets.append((t, None, None))
            for i, (t, obj, key) in enumerate(targets):
                v = vals[i] if i < len(vals) else None
                if t[0] == "name":
                    self.set_var(t[1], v, sc)
                else:
                    self.setindex(obj, key, v, t[3])
        elif k == "localfunc":
            sc.vars[s[1]] = None
            sc.vars[s[1]] = LuaFunction(self, s[2], sc)
        elif k == "if":
            for cond, body in s[1]:
                if truthy(self.eval(cond, sc)):
                    self.run_block(body, Scope(sc))
                    return
            if s[2] is not None:
                self.run_block(s[2], Scope(sc))
        elif k == "while":
            try:
                while truthy(self.eval(s[1], sc)):
                    self.run_block(s[2], Scope(sc))
            except _Break:
                pass
        elif k == "repeat":
            try:
                while True:
                    inner = Scope(sc)
                    self.run_block(s[1], inner)
                    if truthy(self.eval(s[2], inner)):
                        break
            except _Break:
                pass
        elif k == "do":
            self.run_block(s[1], Scope(sc))
        elif k == "fornum":
            self.fornum(s, sc)
        elif k == "forin":
            self.forin(s, sc)
        elif k == "return":
            raise _Return(self.explist(s[1], sc))
        elif k == "break":
            raise _Break()
        else:
            raise LuaError(f"unknown statement {k}")

    def fornum(self, s, sc):
        _, name, a, b, c, body, line = s
        start = arith_num(self.eval(a, sc), "for")
        stop = arith_num(self.eval(b, sc), "for")
        step = arith_num(self.eval(c, sc), "for") if c is not None else 1
        if step == 0:
            raise LuaError(f"{self.chunk}:{line}: 'for' step is zero")
        if not (isinstance(start, int) and isinstance(step, int)):
            start, stop, step = float(start), float(stop), float(step)
        elif isinstance(stop, float):
            stop = math.floor(stop) if step > 0 else math.ceil(stop)
        v = start
        try:
            while (v <= stop) if step > 0 else (v >= stop):
                inner = Scope(sc)
                inner.vars[name] = v
                self.run_block(body, inner)
                v = v + step
        except _Break:
            pass

    def forin(self, s, sc):
        _, names, exprs, body, _line = s
        vals = self.explist(exprs, sc) + [None, None, None]
        f, st, ctl = vals[0], vals[1], vals[2]
        try:
            while True:
                r = self.call(f, [st, ctl]) + [None] * len(names)
                if r[0] is None:
                    break
                ctl = r[0]
                inner = Scope(sc)
                for i, n in enumerate(names):
                    inner.vars[n] = r[i]
                self.run_block(body, inner)
        except _Break:
            pass

    # ---- expressions
    def explist(self, exprs, sc):
        out = []
        for i, e in enumerate(exprs):
            if i == len(exprs) - 1:
                out.extend(self.multi(e, sc))
            else:
                out.append(self.eval(e, sc))
        return out

    def multi(self, e, sc):
        """All values of an expression (calls and ... give several)."""
        k = e[0]
        if k == "call":
            f = self.eval(e[1], sc)
            if f is None:
                raise LuaError(f"{self.chunk}:{e[3]}: attempt to call a nil value ({_callee(e[1])})")
            return self.call(f, self.explist(e[2], sc))
        if k == "method":
            obj = self.eval(e[1], sc)
            f = self.index(obj, e[2], e[4])
            args = [obj] + self.explist(e[3], sc)
            return self.call(f, args)
        if k == "vararg":
            return list(sc.find("...").vars["..."]) if sc.find("...") else []
        return [self.eval(e, sc)]

    def eval(self, e, sc):
        k = e[0]
        if k == "const":
            return e[1]
        if k == "name":
            return self.get_var(e[1], sc)
        if k == "index":
            return self.index(self.eval(e[1], sc), self.eval(e[2], sc), e[3])
        if k in ("call", "method", "vararg"):
            r = self.multi(e, sc)
            return r[0] if r else None
        if k == "paren":
            return self.eval(e[1], sc)
        if k == "binop":
            op = e[1]
            if op == "and":
                a = self.eval(e[2], sc)
                return self.eval(e[3], sc) if truthy(a) else a
            if op == "or":
                a = self.eval(e[2], sc)
                return a if truthy(a) else self.eval(e[3], sc)
            a, b = self.eval(e[2], sc), self.eval(e[3], sc)
            try:
                if op in ("==", "~=", "<", "<=", ">", ">="):
                    return compare(op, a, b)
                return arith(op, a, b)
            except LuaError as err:
                raise LuaError(f"{self.chunk}:{e[4]}: {err}") from None
        if k == "unop":
            v = self.eval(e[2], sc)
            op = e[1]
            if op == "not":
                return not truthy(v)
            if op == "-":
                n = arith_num(v, "unm")
                return -n
            if op == "#":
                if isinstance(v, str):
                    return len(v)
                if isinstance(v, LuaTable):
                    return v.length()
                raise LuaError(f"{self.chunk}:{e[3]}: attempt to get length of a {lua_type(v)} value")
            if op == "~":
                return ~to_integer(v)
        if k == "table":
            t = LuaTable()
            n = 1
            items = e[1]
            for i, it in enumerate(items):
                if it[0] == "kv":
                    t[self.eval(it[1], sc)] = self.eval(it[2], sc)
                elif i == len(items) - 1:
                    for v in self.multi(it[1], sc):
                        t[n] = v
                        n += 1
                else:
                    t[n] = self.eval(it[1], sc)
                    n += 1
            return t
        if k == "function":
            return LuaFunction(self, e, sc)
        raise LuaError(f"unknown expression {k}")

    # ---- standard library
    def _stdlib(self):
        G = self.G
        out = self.out

        def lua_print(*args):
            out.append("\t".join(tostring(a) for a in args))

        def lua_tonumber(v, base=None):
            if base is None:
                return to_number(v)
            try:
                return int(str(v).strip().lower(), int(base))
            except ValueError:
                return None

        def lua_ipairs(t):
            def it(tt, i):
                i = i + 1
                v = tt[i]
                return None if v is None else (i, v)
            return it, t, 0

        def lua_next(t, k=None):
            keys = list(t.d.keys())
            if k is None:
                idx = 0
            else:
                k = LuaTable._key(k)
                idx = keys.index(k) + 1
            if idx >= len(keys):
                return None
            return keys[idx], t.d[keys[idx]]

        def lua_pairs(t):
            return lua_next, t, None

        def lua_select(n, *args):
            if n == "#":
                return len(args)
            n = to_integer(n)
            return tuple(args[n - 1:]) if n > 0 else tuple(args[n:])

        def lua_error(msg=None, level=1):
            raise LuaError(tostring(msg))

        def lua_assert(v=None, msg="assertion failed!", *rest):
            if not truthy(v):
                raise LuaError(tostring(msg))
            return (v, msg) + rest

        def lua_pcall(f, *args):
            try:
                return tuple([True] + self.call(f, list(args)))
            except LuaError as err:
                return False, str(err)

        def lua_type_(v):
            return lua_type(v)

        G.update(print=lua_print, tostring=tostring, tonumber=lua_tonumber, ipairs=lua_ipairs,
                 pairs=lua_pairs, next=lua_next, select=lua_select, error=lua_error,
                 pcall=lua_pcall, type=lua_type_, rawequal=raw_equal)
        G["assert"] = lua_assert

        def floor(x):
            x = arith_num(x, "floor")
            return x if isinstance(x, int) else int(math.floor(x))

        def ceil(x):
            x = arith_num(x, "ceil")
            return x if isinstance(x, int) else int(math.ceil(x))

        def lmax(*a):
            return max(a, key=lambda v: arith_num(v, "max"))

        def lmin(*a):
            return min(a, key=lambda v: arith_num(v, "min"))

        def log(x, base=None):
            x = float(x)
            if x == 0:
                return -math.inf
            if x < 0:
                return math.nan
            if base is None:
                return math.log(x)
            if base == 2:
                return math.log2(x)
            if base == 10:
                return math.log10(x)
            return math.log(x) / math.log(base)

        def modf(x):
            x = float(x)
            f = math.fmod(x, 1.0)
            return float(x - f), f

        def tointeger(x):
            return to_integer(x) if isinstance(x, (int, float)) else None

        def fabs(x):
            x = arith_num(x, "abs")
            return abs(x)

        def sqrt(x):
            x = float(x)
            return math.sqrt(x) if x >= 0 else math.nan

        fmath = {n: (lambda fn: (lambda x: fn(float(x))))(getattr(math, n))
                 for n in ("sin", "cos", "tan", "asin", "acos", "exp")}

        def atan(y, x=1.0):
            return math.atan2(float(y), float(x))

        G["math"] = LuaTable(dict(fmath, sqrt=sqrt, atan=atan, log=log, floor=floor, ceil=ceil, abs=fabs,
                                  max=lmax, min=lmin, pi=math.pi, huge=math.inf, fmod=lambda a, b: math.fmod(a, b),
                                  modf=modf, tointeger=tointeger, maxinteger=2 ** 63 - 1,
                                  mininteger=-2 ** 63, pow=lambda a, b: math.pow(a, b)))

        def fmt(f, *args):
            out_s, ai = [], 0
            for m in re.finditer(r"%([-+ #0]*\d*(?:\.\d+)?)([diouxXeEfgGqsc%aA])|[^%]+", f):
                if m.group(2) is None:
                    out_s.append(m.group(0))
                    continue
                spec, conv = m.group(1), m.group(2)
                if conv == "%":
                    out_s.append("%")
                    continue
                a = args[ai]
                ai += 1
                if conv in "di":
                    out_s.append(("%" + spec + "d") % to_integer(a))
                elif conv in "ouxXc":
                    out_s.append(("%" + spec + conv) % (to_integer(a) if conv != "c" else chr(to_integer(a))))
                elif conv in "eEfgGaA":
                    out_s.append(("%" + spec + (conv if conv not in "aA" else "e")) % float(a))
                elif conv == "q":
                    out_s.append('"' + tostring(a).replace("\\", "\\\\").replace('"', '\\"') + '"')
                else:
                    out_s.append(("%" + spec + "s") % tostring(a))
            return "".join(out_s)

        def sub(s, i=1, j=-1):
            s = tostring(s)
            n = len(s)
            i, j = to_integer(i), to_integer(j)
            if i < 0:
                i = max(n + i + 1, 1)
            elif i == 0:
                i = 1
            if j < 0:
                j = n + j + 1
            elif j > n:
                j = n
            return s[i - 1:j] if i <= j else ""

        G["string"] = LuaTable(dict(format=fmt, rep=lambda s, n, sep="": sep.join([tostring(s)] * to_integer(n)),
                                    sub=sub, len=lambda s: len(tostring(s)),
                                    upper=lambda s: tostring(s).upper(), lower=lambda s: tostring(s).lower()))

        def tinsert(t, *args):
            if len(args) == 1:
                t[t.length() + 1] = args[0]
            else:
                pos, v = to_integer(args[0]), args[1]
                n = t.length()
                for i in range(n, pos - 1, -1):
                    t[i + 1] = t[i]
                t[pos] = v

        def unpack(t, i=1, j=None):
            j = t.length() if j is None else to_integer(j)
            return tuple(t[k] for k in range(to_integer(i), j + 1))

        G["table"] = LuaTable(dict(insert=tinsert, unpack=unpack))
        G["unpack"] = unpack
