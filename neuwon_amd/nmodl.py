"""NMODL front end for linear kinetic schemes: a .mod file with a KINETIC block becomes an
``channels.LTIScheme`` (one input: the membrane voltage) and a density mechanism for the model.

The reference reads .mod files with BlueBrain's ``nmodl`` package and sympy
(neuwon/rxd/nmodl/__init__.py:243-326, lti_sim/lti_model.py:5-78); neither is available here, so
this module parses the subset of the language that the kinetic library uses
(nmodl_library/kinetic/*.mod, lti_sim/tests/*.mod):

* blocks NEURON (SUFFIX | POINT_PROCESS, USEION ... READ ... WRITE ..., NONSPECIFIC_CURRENT,
  POINTER, RANGE, GLOBAL), PARAMETER, CONSTANT, ASSIGNED, STATE, top-level LOCAL, INITIAL,
  BREAKPOINT, KINETIC, FUNCTION, PROCEDURE; UNITS, INDEPENDENT, TITLE and comments are skipped;
* statements: assignment, procedure call, ``SOLVE ... METHOD ...``, ``SOLVE ... STEADYSTATE ...``,
  reactions ``~ A + 2 B <-> C (kf, kb)``, ``CONSERVE a + b = total``, LOCAL, UNITSOFF/UNITSON;
* expressions with + - * / ^, unary minus, parentheses, calls of exp/log/sqrt/fabs/pow and of the
  file's FUNCTIONs; unit annotations after numbers and in declarations are ignored.

Evaluation follows the reference's order: PARAMETER defaults (temperature into ``celsius``), the
INITIAL assignments, then per derivative evaluation the KINETIC block (procedure calls,
assignments, reactions) at the given inputs and states.  Each reaction contributes the flux
kf * prod(reactants) - kb * prod(products); linearity in the states is checked as in
lti_model.py:40-49, so a nonlinear scheme raises.
"""
import math
import os
import re

import numpy as np

from . import channels

__all__ = ["NmodlError", "KineticModel", "load_kinetic", "kinetic_mechanism"]


class NmodlError(ValueError):
    """The file uses NMODL outside the supported subset, or is not a linear kinetic scheme."""


# ------------------------------------------------------------------------------------------------
# Tokens
# ------------------------------------------------------------------------------------------------
_TOKEN = re.compile(r"""
    (?P<num>(\d+\.\d*|\.\d+|\d+)([eE][-+]?\d+)?)
  | (?P<name>[A-Za-z_][A-Za-z_0-9]*)
  | (?P<op><->|<=|>=|==|!=|&&|\|\||[-+*/^(){},=~<>!])
  | (?P<ws>\s+)
""", re.VERBOSE)


def _strip_comments(text):
    text = re.sub(r"(?ms)^\s*COMMENT\b.*?^\s*ENDCOMMENT\b", "", text)
    text = re.sub(r"(?m)^\s*TITLE\b.*$", "", text)
    text = re.sub(r":[^\n]*", "", text)        # ':' comments run to the end of the line
    text = re.sub(r"\?[^\n]*", "", text)       # '?' comments likewise
    return text


def _tokenize(text):
    toks, pos = [], 0
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise NmodlError(f"unexpected character {text[pos]!r} at offset {pos}")
        pos = m.end()
        if m.lastgroup == "ws":
            continue
        toks.append((m.lastgroup, m.group()))
    return toks


class _Tokens:
    def __init__(self, toks):
        self.toks, self.i = toks, 0

    def peek(self, k=0):
        j = self.i + k
        return self.toks[j] if j < len(self.toks) else (None, None)

    def next(self):
        t = self.peek()
        if t[0] is None:
            raise NmodlError("unexpected end of file")
        self.i += 1
        return t

    def expect(self, value):
        kind, val = self.next()
        if val != value:
            raise NmodlError(f"expected {value!r}, found {val!r}")
        return val

    def accept(self, value):
        if self.peek()[1] == value:
            self.i += 1
            return True
        return False

    def skip_unit(self):
        """A parenthesized unit annotation, e.g. (mV), (/ms), (joule/degC), ()."""
        if self.peek()[1] != "(":
            return
        depth = 0
        while True:
            _, val = self.next()
            if val == "(":
                depth += 1
            elif val == ")":
                depth -= 1
                if depth == 0:
                    return

    def block(self):
        """Tokens of a { ... } block (the braces consumed)."""
        self.expect("{")
        start, depth = self.i, 1
        while depth:
            _, val = self.next()
            depth += {"{": 1, "}": -1}.get(val, 0)
        return _Tokens(self.toks[start:self.i - 1])

    def done(self):
        return self.i >= len(self.toks)


# ------------------------------------------------------------------------------------------------
# Expressions -> Python source
# ------------------------------------------------------------------------------------------------
_BUILTINS = {"exp": "_m.exp", "log": "_m.log", "log10": "_m.log10", "sqrt": "_m.sqrt",
             "fabs": "abs", "pow": "_pow", "sin": "_m.sin", "cos": "_m.cos", "tanh": "_m.tanh"}


def _expr(ts, ref, prec=0):
    """Precedence climbing; ``ref(name)`` renders a variable.  Returns Python source."""
    left = _unary(ts, ref)
    binops = {"||": (1, "or"), "&&": (2, "and"), "==": (3, "=="), "!=": (3, "!="), "<": (3, "<"),
              ">": (3, ">"), "<=": (3, "<="), ">=": (3, ">="), "+": (4, "+"), "-": (4, "-"),
              "*": (5, "*"), "/": (5, "/"), "^": (7, "**")}
    while True:
        kind, op = ts.peek()
        if kind != "op" or op not in binops or binops[op][0] < prec:
            return left
        p, py = binops[op]
        ts.next()
        right = _expr(ts, ref, p if op == "^" else p + 1)   # '^' is right-associative
        left = f"({left} {py} {right})"


def _unary(ts, ref):
    kind, val = ts.peek()
    if val == "-":
        ts.next()
        return f"(-{_expr(ts, ref, 6)})"
    if val == "+":
        ts.next()
        return _expr(ts, ref, 6)
    if val == "!":
        ts.next()
        return f"(not {_expr(ts, ref, 6)})"
    return _atom(ts, ref)


def _atom(ts, ref):
    kind, val = ts.next()
    if kind == "num":
        ts.skip_unit()                      # 20(degC), 96.485(joule/mV)
        return repr(float(val))
    if val == "(":
        inner = _expr(ts, ref)
        ts.expect(")")
        return f"({inner})"
    if kind == "name":
        if ts.peek()[1] == "(":             # a call
            ts.next()
            args = []
            if not ts.accept(")"):
                while True:
                    args.append(_expr(ts, ref))
                    if ts.accept(")"):
                        break
                    ts.expect(",")
            fn = _BUILTINS.get(val, f"_f_{val}")
            return f"{fn}({', '.join(args)})"
        return ref(val)
    raise NmodlError(f"unexpected {val!r} in an expression")


# ------------------------------------------------------------------------------------------------
# The file
# ------------------------------------------------------------------------------------------------
class KineticModel:
    """A parsed .mod file with a linear KINETIC block.

    name: SUFFIX (or POINT_PROCESS) name; states: sorted STATE names (the reference's order,
    lti_sim/tests/test_lti_sim.py); parameters: PARAMETER/CONSTANT defaults; inputs: sorted
    names the scheme depends on from outside (v, and POINTER / READ concentrations);
    conserve_sum: CONSERVE total (None if absent); ions: USEION read/write lists; current: the
    variable BREAKPOINT writes (e.g. ina), with its ion.
    """

    def __init__(self, path, celsius=37.0, inputs=None):
        self.path = str(path)
        with open(self.path, "rt") as f:
            text = f.read()
        self.celsius = float(celsius)
        self.name = None
        self.point_process = False
        self.ions = {}              # ion -> {"read": [...], "write": [...]}
        self.nonspecific = []
        self.pointers = []
        self.parameters = {}
        self.assigned = []
        self.state_decl = []
        self.locals = []
        self.initial = None
        self.breakpoint = None
        self.kinetic = {}           # name -> _Tokens
        self.functions = {}         # name -> (args, _Tokens)
        self.procedures = {}
        self.derivative_blocks = []
        self._parse(_Tokens(_tokenize(_strip_comments(text))))
        if self.name is None:
            raise NmodlError(f"{self.path}: no SUFFIX or POINT_PROCESS")
        if self.derivative_blocks and not self.kinetic:
            raise NmodlError(f"{self.name}: DERIVATIVE blocks are not a kinetic scheme")
        if not self.kinetic:
            raise NmodlError(f"{self.name}: no KINETIC block")
        self.states = sorted(self.state_decl)
        self._compile()
        used = self._free_names()
        if inputs is None:
            inputs = [n for n in ("v",) if n in used] + [p for p in self.pointers if p in used]
            for ion in self.ions.values():
                inputs += [r for r in ion["read"] if r in used and not r.startswith("e")
                           and r not in inputs]
        self.inputs = sorted(inputs)
        self.num_inputs = len(self.inputs)
        self.num_states = len(self.states)
        self._check_linear()

    def _check_linear(self):
        """lti_model.py:40-49: d(2 s) = 2 d(s) at random inputs (voltages in [-100, 100] mV,
        other inputs in [0, 1000]) and random states."""
        rng = np.random.default_rng(12345)
        for _ in range(3):
            ins = [rng.uniform(-100.0, 100.0) if n == "v" else rng.uniform(0.0, 1000.0)
                   for n in self.inputs]
            s1 = rng.uniform(0.0, 1.0, self.num_states)
            d1 = np.asarray(self.derivative(*ins, *s1), dtype=np.float64)
            d2 = np.asarray(self.derivative(*ins, *(2.0 * s1)), dtype=np.float64)
            if not (np.all(np.isfinite(d1)) and np.all(np.abs(d1 - d2 / 2.0) <= 1e-12 *
                                                         np.maximum(1.0, np.abs(d1)))):
                raise NmodlError(f"{self.name}: non-linear system detected")

    # -------------------------------------------------------------------- parsing
    def _parse(self, ts):
        while not ts.done():
            kind, word = ts.next()
            if word == "NEURON":
                self._neuron(ts.block())
            elif word in ("PARAMETER", "CONSTANT"):
                self._parameters(ts.block())
            elif word == "ASSIGNED":
                self.assigned += self._declared_names(ts.block())
            elif word == "STATE":
                self.state_decl += self._declared_names(ts.block())
            elif word == "LOCAL":
                self.locals += self._name_list(ts)
            elif word == "INITIAL":
                self.initial = ts.block()
            elif word == "BREAKPOINT":
                self.breakpoint = ts.block()
            elif word == "KINETIC":
                name = ts.next()[1]
                self.kinetic[name] = ts.block()
            elif word == "DERIVATIVE":
                self.derivative_blocks.append(ts.next()[1])
                ts.block()
            elif word in ("FUNCTION", "PROCEDURE"):
                name = ts.next()[1]
                args = self._arg_list(ts)
                ts.skip_unit()                       # FUNCTION f(x) (units) { ... }
                body = ts.block()
                (self.functions if word == "FUNCTION" else self.procedures)[name] = (args, body)
            elif word in ("UNITS", "INDEPENDENT", "UNITSOFF", "UNITSON"):
                if ts.peek()[1] == "{":
                    ts.block()
            elif word in ("NET_RECEIVE", "VERBATIM", "FUNCTION_TABLE", "LINEAR", "NONLINEAR",
                          "DISCRETE", "PARTIAL", "BEFORE", "AFTER", "CONSTRUCTOR", "DESTRUCTOR"):
                raise NmodlError(f"{word} blocks are not supported")
            else:
                raise NmodlError(f"unexpected {word!r} at the top level")

    def _neuron(self, ts):
        while not ts.done():
            _, word = ts.next()
            if word in ("SUFFIX", "POINT_PROCESS", "ARTIFICIAL_CELL"):
                self.name = ts.next()[1]
                self.point_process = word != "SUFFIX"
            elif word == "USEION":
                ion = ts.next()[1]
                entry = self.ions.setdefault(ion, {"read": [], "write": []})
                while ts.peek()[1] in ("READ", "WRITE", "VALENCE"):
                    mode = ts.next()[1]
                    if mode == "VALENCE":
                        ts.accept("-")
                        ts.next()
                        continue
                    entry[mode.lower()] += self._name_list(ts)
            elif word == "NONSPECIFIC_CURRENT":
                self.nonspecific += self._name_list(ts)
            elif word == "POINTER":
                self.pointers += self._name_list(ts)
            elif word in ("RANGE", "GLOBAL", "BBCOREPOINTER", "ELECTRODE_CURRENT", "THREADSAFE"):
                if word != "THREADSAFE":
                    self._name_list(ts)
            else:
                raise NmodlError(f"unexpected {word!r} in the NEURON block")

    def _name_list(self, ts):
        names = [ts.next()[1]]
        while ts.accept(","):
            names.append(ts.next()[1])
        return names

    def _arg_list(self, ts):
        ts.expect("(")
        args = []
        if ts.accept(")"):
            return args
        while True:
            args.append(ts.next()[1])
            ts.skip_unit()
            if ts.accept(")"):
                return args
            ts.expect(",")

    def _declared_names(self, ts):
        names = []
        while not ts.done():
            kind, name = ts.next()
            if kind != "name":
                raise NmodlError(f"unexpected {name!r} in a declaration block")
            names.append(name)
            ts.skip_unit()
            if ts.accept("FROM"):                # STATE { x FROM 0 TO 1 }
                _expr(ts, lambda n: "0")
                ts.expect("TO")
                _expr(ts, lambda n: "0")
        return names

    def _parameters(self, ts):
        while not ts.done():
            kind, name = ts.next()
            if kind != "name":
                raise NmodlError(f"unexpected {name!r} in PARAMETER")
            value = None
            if ts.accept("="):
                sign = -1.0 if ts.accept("-") else 1.0
                k2, num = ts.next()
                if k2 != "num":
                    raise NmodlError(f"parameter {name}: expected a number, found {num!r}")
                value = sign * float(num)
            ts.skip_unit()
            if ts.accept("<"):                   # range limits <lo, hi>
                while not ts.accept(">"):
                    ts.next()
            self.parameters[name] = value

    # -------------------------------------------------------------------- code generation
    def _statements(self, ts, ref, locals_, out, indent, kinetic=False):
        """Python source lines of a block's statements."""
        pad = " " * indent
        while not ts.done():
            kind, word = ts.peek()
            if word in ("UNITSOFF", "UNITSON"):
                ts.next()
            elif word == "LOCAL":
                ts.next()
                locals_.update(self._name_list(ts))
            elif word == "SOLVE":
                ts.next()
                ts.next()
                while ts.peek()[1] in ("METHOD", "STEADYSTATE"):
                    ts.next()
                    ts.next()
            elif word == "~":
                if not kinetic:
                    raise NmodlError("reaction outside a KINETIC block")
                ts.next()
                lhs = self._species(ts)
                ts.expect("<->")
                rhs = self._species(ts)
                ts.expect("(")
                kf = _expr(ts, ref)
                ts.expect(",")
                kb = _expr(ts, ref)
                ts.expect(")")
                fwd = " * ".join([f"({kf})"] + [f"_s[{self._state_index(n)}] ** {c}" for n, c in lhs])
                bwd = " * ".join([f"({kb})"] + [f"_s[{self._state_index(n)}] ** {c}" for n, c in rhs])
                out.append(f"{pad}_flux = {fwd} - {bwd}")
                for n, c in lhs:
                    out.append(f"{pad}_d[{self._state_index(n)}] -= {c} * _flux")
                for n, c in rhs:
                    out.append(f"{pad}_d[{self._state_index(n)}] += {c} * _flux")
            elif word == "CONSERVE":
                ts.next()
                while not ts.accept("="):
                    ts.next()
                total = _expr(ts, lambda n: repr(self.parameters.get(n) or 0.0))
                self._conserve = float(eval(total, {"_m": math}))
            elif word == "if":
                raise NmodlError("if statements are not supported")
            elif kind == "name" and ts.peek(1)[1] == "(":
                ts.next()
                ts.next()
                args = []
                if not ts.accept(")"):
                    while True:
                        args.append(_expr(ts, ref))
                        if ts.accept(")"):
                            break
                        ts.expect(",")
                if word not in self.procedures:
                    raise NmodlError(f"call of unknown procedure {word}")
                out.append(f"{pad}_p_{word}({', '.join(args)})")
            elif kind == "name" and ts.peek(1)[1] == "=":
                ts.next()
                ts.next()
                value = _expr(ts, ref)
                out.append(f"{pad}{ref(word)} = {value}")
            else:
                raise NmodlError(f"unsupported statement starting with {word!r}")

    def _species(self, ts):
        terms = []
        while True:
            coef = 1
            if ts.peek()[0] == "num":
                coef = int(float(ts.next()[1]))
            name = ts.next()[1]
            if name not in self.state_decl:
                raise NmodlError(f"{name} in a reaction is not a STATE")
            terms.append((name, coef))
            if not ts.accept("+"):
                return terms

    def _state_index(self, name):
        return self.states.index(name)

    def _compile(self):
        """Python functions for every FUNCTION/PROCEDURE, the INITIAL assignments, the KINETIC
        derivative and the BREAKPOINT current."""
        ns = {"_m": math, "_pow": math.pow}
        lines = []
        self._conserve = None

        def global_ref(name):
            return f"_g[{name!r}]"

        for kind, table in (("f", self.functions), ("p", self.procedures)):
            for name, (args, body) in table.items():
                locals_ = set(args)
                if kind == "f":
                    locals_.add(name)
                ref = (lambda locs: (lambda n: f"_l_{n}" if n in locs else f"_g[{n!r}]"))(locals_)
                body_lines = []
                self._statements(_Tokens(list(body.toks)), ref, locals_, body_lines, 4)
                head = f"def _{kind}_{name}({', '.join('_l_' + a for a in args)}):"
                lines.append(head)
                lines += [f"    _l_{n} = 0.0" for n in sorted(locals_ - set(args))]
                lines += body_lines or ["    pass"]
                lines.append(f"    return _l_{name}" if kind == "f" else "    return None")
        # INITIAL assignments (the SOLVE ... STEADYSTATE is the scheme's steady state)
        lines.append("def _initial():")
        body_lines = []
        if self.initial is not None:
            self._statements(_Tokens(list(self.initial.toks)), global_ref, set(), body_lines, 4)
        lines += body_lines or ["    pass"]
        # KINETIC: derivative of the states (first KINETIC block solved by BREAKPOINT)
        kin = next(iter(self.kinetic.values()))
        lines.append("def _kinetic(_s):")
        lines.append(f"    _d = [0.0] * {len(self.state_decl)}")
        body_lines = []
        self._statements(_Tokens(list(kin.toks)), global_ref, set(), body_lines, 4, kinetic=True)
        lines += body_lines
        lines.append("    return _d")
        lines.append("def _breakpoint():")
        body_lines = []
        if self.breakpoint is not None:
            self._statements(_Tokens(list(self.breakpoint.toks)), global_ref, set(), body_lines, 4)
        lines += body_lines or ["    pass"]
        self.source = "\n".join(lines)
        self._ns = ns
        try:
            exec(compile(self.source, self.path, "exec"), ns)
        except SyntaxError as e:   # the parser produced something Python rejects
            raise NmodlError(f"{self.name}: generated code does not compile ({e})")
        self.conserve_sum = self._conserve

    def _free_names(self):
        """Globals read by the KINETIC block and the FUNCTIONs and PROCEDUREs (not BREAKPOINT)."""
        src = self.source.split("def _initial():")[0] + \
            self.source.split("def _kinetic(_s):")[1].split("def _breakpoint():")[0]
        return set(re.findall(r"_g\['([A-Za-z_0-9]+)'\]", src))

    # -------------------------------------------------------------------- evaluation
    def _globals(self, inputs):
        g = {k: (0.0 if v is None else float(v)) for k, v in self.parameters.items()}
        g.update({k: 0.0 for k in self.assigned if k not in g})
        g.update({k: 0.0 for k in self.locals})
        g["celsius"] = self.celsius
        for name in self.state_decl:
            g[name] = 0.0
        g.update({k: float(v) for k, v in inputs.items()})
        return g

    def _run(self, fn, inputs, *args):
        g = self._globals(inputs)
        self._ns["_g"] = g
        self._ns["_initial"]()
        g.update({k: float(v) for k, v in inputs.items()})   # inputs win over INITIAL
        return g, self._ns[fn](*args)

    def rate_matrices(self, *arrays):
        """Rate matrices [..., S, S] at arrays of input values (sorted input order, broadcast
        together): the KINETIC block run once per state on whole arrays, numpy functions in place
        of math.  Blocks whose control flow depends on an input fall back to one evaluation per
        point."""
        arrays = np.broadcast_arrays(*[np.asarray(a, dtype=np.float64) for a in arrays])
        S = self.num_states
        try:
            ns = self._vector_ns
        except AttributeError:
            ns = self._vector_ns = {"_m": np, "_pow": np.power}
            exec(compile(self.source, self.path, "exec"), ns)
        try:
            with np.errstate(all="ignore"):
                g = self._globals({})
                g.update(dict(zip(self.inputs, arrays)))
                ns["_g"] = g
                ns["_initial"]()
                g.update(dict(zip(self.inputs, arrays)))
                A = np.zeros(arrays[0].shape + (S, S))
                for j in range(S):
                    d = ns["_kinetic"]([float(k == j) for k in range(S)])
                    for i in range(S):
                        A[..., i, j] = d[i]
            return A
        except (TypeError, ValueError):
            flat = [a.ravel() for a in arrays]
            A = np.empty((flat[0].size, S, S))
            for k in range(flat[0].size):
                x = [f[k] for f in flat]
                for j in range(S):
                    A[k, :, j] = self.derivative(*x, *[float(q == j) for q in range(S)])
            return A.reshape(arrays[0].shape + (S, S))

    def derivative(self, *values):
        """d(states)/dt (1/ms) at inputs (sorted order) and states (sorted order)."""
        inputs = dict(zip(self.inputs, values[:self.num_inputs]))
        states = [float(x) for x in values[self.num_inputs:]]
        _, d = self._run("_kinetic", inputs, states)
        return d

    def conductance_weights(self, reversal_name=None, **inputs):
        """Conductance of each state per unit occupancy from BREAKPOINT, in the file's units
        (S/cm^2 for a density mechanism, uS for a point process).  A written conductance
        (e.g. gzero = ...) is read directly; a written current (ina = g (v - ena), i = g (v - Erev))
        is evaluated with each state set to one at v = 0 and v = 1 mV, the ion's reversal
        potential (reversal_name) at 0: the weight is the difference."""
        w = np.empty(self.num_states)
        for k, s in enumerate(self.states):
            vals = dict(inputs)
            vals.update({n: float(n == s) for n in self.states})
            if reversal_name:
                vals[reversal_name] = 0.0
            if self.output_kind == "conductance":
                g, _ = self._run("_breakpoint", vals)
                w[k] = g[self.output_variable]
                continue
            out = []
            for v in (0.0, 1.0):
                vals["v"] = v
                g, _ = self._run("_breakpoint", vals)
                out.append(g[self.output_variable])
            w[k] = out[1] - out[0]
        return w

    def current_reversal(self, **inputs):
        """The reversal potential a written current implies with the file's own parameters
        (i(v) = w (v - E): E = -i(0) / w), for currents of no declared species (Erev of a
        NONSPECIFIC_CURRENT).  None when the file writes a conductance."""
        if self.output_kind == "conductance":
            return None
        k = int(np.argmax(self.conductance_weights(None, **inputs)))
        vals = dict(inputs)
        vals.update({n: float(n == self.states[k]) for n in self.states})
        out = []
        for v in (0.0, 1.0):
            vals["v"] = v
            g, _ = self._run("_breakpoint", vals)
            out.append(g[self.output_variable])
        return -out[0] / (out[1] - out[0])

    @property
    def output_variable(self):
        """What BREAKPOINT writes: the ion current, else the nonspecific current, else the ion
        conductance (g<ion>)."""
        try:
            return self.current_variable
        except NmodlError:
            for ion, rw in self.ions.items():
                if f"g{ion}" in rw["write"]:
                    return f"g{ion}"
            raise NmodlError(f"{self.name} writes neither a current nor a conductance")

    @property
    def output_kind(self):
        return "conductance" if self.output_variable.startswith("g") else "current"

    @property
    def output_ion(self):
        var = self.output_variable
        if var.startswith("g"):
            return var[1:]
        return self.current_ion

    @property
    def current_variable(self):
        for ion, rw in self.ions.items():
            for name in rw["write"]:
                if name == f"i{ion}":
                    return name
        if self.nonspecific:
            return self.nonspecific[0]
        raise NmodlError(f"{self.name} writes no current")

    @property
    def current_ion(self):
        for ion, rw in self.ions.items():
            if f"i{ion}" in rw["write"]:
                return ion
        return None


# Input tables (nmodl/__init__.py:306-316 maps the voltage linearly over [-100, 100] mV and any
# other input logarithmically over [0, 1000]).  The reference's optimizer searches bucket counts
# per file for a max abs table error of 1e-4 (nmodl/__init__.py:304, optimizer.py:85-263); by
# default these fixed counts are used instead, sized for linear interpolation errors of a few
# 1e-6 per step (tests/test_nmodl.py measures them), and ``accuracy=`` runs the search
# (channels.choose_grid) from them.
LOG_SCALE = 1e-4
BUCKETS_1D = {"v": 2000, "log": 2400}
BUCKETS_2D = {"v": 400, "log": 600}


def input_grid(name, num_inputs=1):
    """The bucket mapping of one input of a kinetic scheme."""
    buckets = BUCKETS_1D if num_inputs == 1 else BUCKETS_2D
    if name == "v":
        return channels.LinearInput(-100.0, 100.0, buckets["v"])
    return channels.LogarithmicInput(0.0, 1000.0, buckets["log"], LOG_SCALE)


def load_kinetic(path, time_step=0.1, celsius=37.0, grid=None, inputs=None, accuracy=None):
    """The .mod file's kinetic scheme as an LTIScheme of its inputs (one or two: the membrane
    voltage and/or a concentration, sorted by name as the reference sorts them).  ``grid``
    replaces the voltage grid of a voltage-only scheme; ``inputs`` gives every grid;
    ``accuracy`` (e.g. the reference's 1e-4) sizes every grid by ``channels.choose_grid``
    instead, for that max abs error of the interpolated propagator."""
    km = KineticModel(path, celsius=celsius)
    if not 1 <= km.num_inputs <= 2:
        raise NmodlError(f"{km.name}: inputs {km.inputs}; the device kernel takes one or two")
    if inputs is None:
        if grid is not None and km.inputs == ["v"]:
            inputs = [grid]
        else:
            inputs = [input_grid(n, km.num_inputs) for n in km.inputs]
    scheme = channels.LTIScheme(km.name, km.states, km.derivative, km.conserve_sum, time_step,
                                inputs=inputs, input_names=km.inputs,
                                rate_matrices=km.rate_matrices)
    if accuracy is not None:
        chosen, err = channels.choose_grid(scheme, accuracy)
        scheme = channels.LTIScheme(km.name, km.states, km.derivative, km.conserve_sum,
                                    time_step, inputs=chosen, input_names=km.inputs,
                                    rate_matrices=km.rate_matrices)
        scheme.table_error = err
    scheme.model = km
    return scheme


def kinetic_mechanism(path, time_step, celsius, reversal_potentials, grid=None, accuracy=None):
    """A model mechanism (``model.LTIMechanism``) from a kinetic .mod file: the conductance of
    every state from BREAKPOINT, the reversal potential of the written ion from the model's
    species (or the file's PARAMETER default, or the reversal its nonspecific current implies).
    A POINT_PROCESS's conductance is in uS per unit occupancy and its magnitude is not scaled by
    the segment's area (nmodl/__init__.py:494-497); inputs other than the voltage are
    attributes of each instance.  ``accuracy``: as in ``load_kinetic``."""
    from .model import LTIMechanism
    scheme = load_kinetic(path, time_step, celsius, grid, accuracy=accuracy)
    km = scheme.model
    ion = km.output_ion
    e_name = f"e{ion}" if ion else None
    if ion is not None and ion in reversal_potentials:
        E = float(reversal_potentials[ion])
    elif e_name and km.parameters.get(e_name) is not None:
        E = float(km.parameters[e_name])
    elif ion is None and km.current_reversal() is not None:
        E = float(km.current_reversal())
    elif km.parameters.get("Erev") is not None:
        E = float(km.parameters["Erev"])
    else:
        raise NmodlError(f"{km.name}: no reversal potential for ion {ion!r} (declare the species)")
    w = km.conductance_weights(e_name)
    if km.point_process:
        scheme.weights = w * 1e-6                           # uS -> S
        return LTIMechanism(scheme, 1.0, E, name=km.name, point_process=True)
    scheme.weights = w
    return LTIMechanism(scheme, 1.0, E, name=km.name)
