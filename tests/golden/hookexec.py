"""Execute the reference's own `.hook` text (a GLSL subset) in FP64 NumPy.

This is the recipe behind the committed golden vectors in this directory: it reads a
shader file from the reference checkout, translates each pass's GLSL into Python
(vectorised over the pass's pixels or compute invocations), and runs the passes the way
mpv's user-shader engine chains them: a pass with //!SAVE stores a named texture, a pass
without it replaces the hooked plane; //!WIDTH/HEIGHT are evaluated as RPN;
`texture()` follows the //!TEXTURE's FILTER (NEAREST / LINEAR, clamp-to-edge) over the
FP16-rounded table (the `rgba16f` upload); NAME_tex / NAME_texOff sample bound
textures with LINEAR filtering and clamp-to-edge; compute passes run their shared-memory
loader per workgroup, then every invocation, `imageStore` scattering into the output.

It runs only where the reference checkout exists (the GPU machines never import it).
Arithmetic is FP64 throughout, with the shader's float literals taken at FP64.
"""
import re
import struct

import numpy as np

# ----------------------------------------------------------------------------- values


class V:
    """GLSL vector: a list of components, each a scalar or a NumPy array over the domain."""

    __slots__ = ("c",)
    __array_ufunc__ = None  # make NumPy defer `array * V` to V.__rmul__
    _IDX = {"x": 0, "y": 1, "z": 2, "w": 3, "r": 0, "g": 1, "b": 2, "a": 3}

    def __init__(self, comps):
        self.c = list(comps)

    def __getattr__(self, name):
        if name.startswith("_") or any(ch not in V._IDX for ch in name):
            raise AttributeError(name)
        idx = [V._IDX[ch] for ch in name]
        if len(idx) == 1:
            return self.c[idx[0]]
        return V([self.c[i] for i in idx])

    def __getitem__(self, i):
        if isinstance(i, np.ndarray):  # per-pixel component index (ravu-lite step2 `[idx]`)
            comps = np.broadcast_arrays(*[np.asarray(c, dtype=np.float64) for c in self.c], i)
            return np.choose(comps[-1], comps[:-1])
        return self.c[int(i)]

    def __setitem__(self, i, v):
        self.c[int(i)] = v

    def _bin(self, o, f):
        if isinstance(o, V):
            return V([f(a, b) for a, b in zip(self.c, o.c)])
        return V([f(a, o) for a in self.c])

    def _rbin(self, o, f):
        return V([f(o, a) for a in self.c])

    def __add__(self, o): return self._bin(o, lambda a, b: a + b)
    def __radd__(self, o): return self._rbin(o, lambda a, b: a + b)
    def __sub__(self, o): return self._bin(o, lambda a, b: a - b)
    def __rsub__(self, o): return self._rbin(o, lambda a, b: a - b)
    def __mul__(self, o): return self._bin(o, lambda a, b: a * b)
    def __rmul__(self, o): return self._rbin(o, lambda a, b: a * b)
    def __truediv__(self, o): return self._bin(o, lambda a, b: a / b)
    def __rtruediv__(self, o): return self._rbin(o, lambda a, b: a / b)
    def __neg__(self): return V([-a for a in self.c])


class M:
    """mat4x3 as 4 vec3 columns (only the operations the shaders use)."""

    __array_ufunc__ = None

    def __init__(self, cols):
        self.cols = list(cols)

    def __getitem__(self, i):
        return self.cols[int(i)]

    def __setitem__(self, i, v):
        self.cols[int(i)] = v

    def __add__(self, o):
        return M([a + b for a, b in zip(self.cols, o.cols)])


def _vec(n):
    def make(*args):
        comps = []
        for a in args:
            if isinstance(a, V):
                comps.extend(a.c)
            else:
                comps.append(a)
        if len(comps) == 1:
            comps = comps * n
        assert len(comps) == n, (n, len(comps))
        return V(comps)
    return make


def mat4x3(*args):
    if len(args) == 1 and not isinstance(args[0], V):
        return M([V([args[0] * (1.0 if i == j else 0.0) for j in range(3)]) for i in range(4)])
    assert len(args) == 4
    return M(list(args))


def _map(f, *xs):
    if any(isinstance(x, V) for x in xs):
        n = max(len(x.c) for x in xs if isinstance(x, V))
        cols = [[x.c[i] if isinstance(x, V) else x for x in xs] for i in range(n)]
        return V([f(*c) for c in cols])
    return f(*xs)


def _ew(f):
    return lambda *xs: _map(f, *xs)


def glsl_mix(a, b, t):
    if isinstance(t, (bool, np.bool_)) or (isinstance(t, np.ndarray) and t.dtype == bool):
        return _map(lambda x, y: np.where(t, y, x), a, b)
    if isinstance(t, V) and isinstance(t.c[0], (bool, np.bool_, np.ndarray)) and np.asarray(t.c[0]).dtype == bool:
        return V([np.where(tc, y, x) for x, y, tc in zip(a.c, b.c, t.c)])
    return _map(lambda x, y, s: x * (1.0 - s) + y * s, a, b, t)


def glsl_clamp(x, lo, hi):
    return _map(lambda v, l, h: np.minimum(np.maximum(v, l), h), x, lo, hi)


def glsl_mod(x, y):
    return _map(lambda a, b: a - b * np.floor(a / b), x, y)


def dot(a, b):
    s = 0.0
    for x, y in zip(a.c, b.c):
        s = s + x * y
    return s


def outerProduct(c, r):  # column vector c (vec3) times row vector r (vec4) -> mat4x3
    return M([c * r.c[j] for j in range(4)])


def matrixCompMult(a, b):
    return M([x * y for x, y in zip(a.cols, b.cols)])


def intBitsToFloat(i):
    return float(struct.unpack("<f", struct.pack("<i", int(i)))[0])


def _sqrt(x):
    with np.errstate(invalid="ignore"):
        return np.sqrt(x)


def _invsqrt(x):
    with np.errstate(invalid="ignore", divide="ignore"):
        return 1.0 / np.sqrt(x)


def glsl_int(x):
    if isinstance(x, np.ndarray):
        if x.dtype == bool:
            return x.astype(np.int64)
        return np.trunc(x).astype(np.int64)
    return int(x)


BUILTINS = {
    "vec2": _vec(2), "vec3": _vec(3), "vec4": _vec(4), "ivec2": lambda *a: V([glsl_int(v) for v in _vec(2)(*a).c]),
    "mat4x3": mat4x3, "float": lambda x: x.astype(np.float64) if isinstance(x, np.ndarray) else float(x),
    "int": glsl_int, "mix": glsl_mix, "clamp": glsl_clamp, "mod": glsl_mod, "dot": dot,
    "max": _ew(np.maximum), "min": _ew(np.minimum), "abs": _ew(np.abs), "floor": _ew(np.floor),
    "fract": _ew(lambda x: x - np.floor(x)), "sqrt": _ew(_sqrt), "inversesqrt": _ew(_invsqrt),
    "exp": _ew(np.exp), "log2": _ew(lambda x: np.log2(x)), "atan": lambda y, x: np.arctan2(y, x),
    "outerProduct": outerProduct, "matrixCompMult": matrixCompMult, "intBitsToFloat": intBitsToFloat,
}

# ----------------------------------------------------------------------------- textures


class Tex:
    """A 2D texture (H, W, 4) with a filter; sampling is clamp-to-edge."""

    def __init__(self, data, linear=True):
        self.d = np.asarray(data, dtype=np.float64)
        self.linear = linear

    @property
    def size(self):
        return V([float(self.d.shape[1]), float(self.d.shape[0])])

    def sample(self, pos):
        H, W = self.d.shape[:2]
        u = np.asarray(pos.c[0], dtype=np.float64) * W
        v = np.asarray(pos.c[1], dtype=np.float64) * H
        u, v = np.broadcast_arrays(u, v)
        if not self.linear:
            xi = np.clip(np.floor(u).astype(np.int64), 0, W - 1)
            yi = np.clip(np.floor(v).astype(np.int64), 0, H - 1)
            t = self.d[yi, xi]
            return V([t[..., k] for k in range(4)])
        # coordinates within 1e-9 texel of a texel centre snap to it, as a GPU's fixed-point
        # sub-texel position does (FP64 arithmetic on HOOKED_pos lands a hair off centres)
        x = u - 0.5
        y = v - 0.5
        x = np.where(np.abs(x - np.round(x)) < 1e-9, np.round(x), x)
        y = np.where(np.abs(y - np.round(y)) < 1e-9, np.round(y), y)
        x0 = np.floor(x)
        y0 = np.floor(y)
        fx = (x - x0)[..., None]
        fy = (y - y0)[..., None]
        x0 = x0.astype(np.int64)
        y0 = y0.astype(np.int64)
        xa, xb = np.clip(x0, 0, W - 1), np.clip(x0 + 1, 0, W - 1)
        ya, yb = np.clip(y0, 0, H - 1), np.clip(y0 + 1, 0, H - 1)
        t = ((self.d[ya, xa] * (1 - fx) + self.d[ya, xb] * fx) * (1 - fy)
             + (self.d[yb, xa] * (1 - fx) + self.d[yb, xb] * fx) * fy)
        return V([t[..., k] for k in range(4)])


def texture_gather_offset(tex, pos, off, comp):
    """textureGatherOffset: the 2x2 texel footprint of LINEAR filtering at pos + off
    texels, component `comp`, in GLSL order (i0,j1), (i1,j1), (i1,j0), (i0,j0).  The
    footprint origin snaps like a GPU's fixed-point texel coordinate (a 1e-9 nudge)."""
    H, W = tex.d.shape[:2]
    u = np.asarray(pos.c[0], dtype=np.float64) * W - 0.5
    v = np.asarray(pos.c[1], dtype=np.float64) * H - 0.5
    i0 = np.floor(u + 1e-9).astype(np.int64) + int(off.c[0])
    j0 = np.floor(v + 1e-9).astype(np.int64) + int(off.c[1])

    def at(i, j):
        return tex.d[np.clip(j, 0, H - 1), np.clip(i, 0, W - 1), int(comp)]
    return V([at(i0, j0 + 1), at(i0 + 1, j0 + 1), at(i0 + 1, j0), at(i0, j0)])


class Shared:
    """A `shared` array of one workgroup family: store (ngroups, n[, comps])."""

    def __init__(self, n, comps):
        self.n, self.comps = n, comps
        self.store = None
        self.mode = None
        self.gidx = None

    active = True   # loader-loop mask over workgroups (loops whose bound varies per group)

    def __setitem__(self, i, v):
        if self.mode != "load":
            return
        G = self.store.shape[0]
        rows = np.arange(G)
        idx = np.broadcast_to(np.asarray(i), (G,)).astype(np.int64)
        m = np.broadcast_to(np.asarray(Shared.active), (G,))
        rows, idx = rows[m], idx[m]
        if self.comps == 1:
            self.store[rows, idx] = np.broadcast_to(np.asarray(v, dtype=np.float64), (G,))[m]
        else:
            for k in range(self.comps):
                self.store[rows, idx, k] = np.broadcast_to(np.asarray(v.c[k], dtype=np.float64), (G,))[m]

    def __getitem__(self, i):
        if self.mode == "load":
            idx = np.broadcast_to(np.asarray(i), (self.store.shape[0],)).astype(np.int64)
            vals = self.store[np.arange(self.store.shape[0]), np.clip(idx, 0, self.n - 1)]
        else:
            vals = self.store[self.gidx, np.asarray(i)]
        if self.comps == 1:
            return vals
        return V([vals[..., k] for k in range(self.comps)])


# ----------------------------------------------------------------------------- parsing


def parse_hook_file(path):
    """Passes (directives + GLSL body) and textures of a hook file."""
    lines = open(path).read().split("\n")
    passes, textures = [], {}
    i = 0
    while i < len(lines):
        if lines[i].startswith("//!TEXTURE"):
            name = lines[i].split()[1]
            w, h = (int(v) for v in lines[i + 1].split()[1:3])
            filt = lines[i + 3].split()[1]
            arr = np.frombuffer(bytes.fromhex(lines[i + 4].strip()), dtype="<f4").reshape(h, w, 4)
            textures[name] = Tex(arr.astype(np.float16).astype(np.float64), linear=(filt == "LINEAR"))
            i += 5
            continue
        if lines[i].startswith("//!DESC"):
            d = {"BIND": [], "line": i + 1}
            while lines[i].startswith("//!"):
                key, _, rest = lines[i][3:].partition(" ")
                if key == "BIND":
                    d["BIND"].append(rest.strip())
                else:
                    d[key] = rest.strip()
                i += 1
            body = []
            while i < len(lines) and not lines[i].startswith("//!"):
                body.append(lines[i])
                i += 1
            d["body"] = "\n".join(body)
            passes.append(d)
            continue
        i += 1
    return passes, textures


def rpn(expr, env):
    st = []
    for tok in expr.split():
        if tok in "+-*/<>" and len(tok) == 1:
            b, a = st.pop(), st.pop()
            st.append({"+": a + b, "-": a - b, "*": a * b, "/": a / b, "<": float(a < b), ">": float(a > b)}[tok])
        elif re.match(r"^[A-Za-z_]\w*\.[wh]$", tok):
            st.append(float(env[tok]))
        else:
            st.append(float(tok))
    assert len(st) == 1, expr
    return st[0]


def expand_macros(src):
    """Apply #define macros (object- and function-like) textually; drop #pragma."""
    out, macros = [], {}
    for ln in src.split("\n"):
        ln = re.sub(r"//.*$", "", ln)
        s = ln.strip()
        if s.startswith("#pragma"):
            continue
        m = re.match(r"#define\s+(\w+)(\(([^)]*)\))?\s*(.*)$", s)
        if m:
            params = [p.strip() for p in m.group(3).split(",")] if m.group(2) else None
            macros[m.group(1)] = (params, m.group(4))
            continue
        out.append(ln)
    text = "\n".join(out)
    for _ in range(8):
        changed = False
        for name, (params, body) in macros.items():
            if params is None:
                new = re.sub(r"\b%s\b" % name, body, text)
            else:
                new = _expand_fn(text, name, params, body)
            if new != text:
                text, changed = new, True
        if not changed:
            break
    return text


def _split_args(s):
    args, depth, cur = [], 0, ""
    for ch in s:
        if ch == "," and depth == 0:
            args.append(cur)
            cur = ""
            continue
        depth += ch in "([" and 1 or (ch in ")]" and -1 or 0)
        cur += ch
    args.append(cur)
    return [a.strip() for a in args]


def _expand_fn(text, name, params, body):
    res, i = [], 0
    pat = re.compile(r"\b%s\s*\(" % name)
    while True:
        m = pat.search(text, i)
        if not m:
            res.append(text[i:])
            break
        res.append(text[i:m.start()])
        j, depth = m.end(), 1
        while depth:
            depth += {"(": 1, ")": -1}.get(text[j], 0)
            j += 1
        args = _split_args(text[m.end():j - 1])
        b = body
        for p, a in zip(params, args):
            b = re.sub(r"\b%s\b" % p, "(%s)" % a, b)
        res.append(b)
        i = j
    return "".join(res)


TYPES = r"(?:const\s+)?(?:float|int|bool|vec2|vec3|vec4|ivec2|mat4x3)"


class Translator:
    """GLSL statements -> Python source (vectorised semantics)."""

    def __init__(self):
        self.out = []

    def expr(self, e, int_ctx=False):
        e = e.strip()
        e = e.replace("&&", " & ").replace("||", " | ")
        if int_ctx:
            e = re.sub(r"(?<![/])/(?![/])", "//", e)
        return e

    def stmts(self, src):
        """Split a body into a nested statement list."""
        toks = []
        i, n = 0, len(src)
        while i < n:
            ch = src[i]
            if ch.isspace():
                i += 1
                continue
            if src.startswith("for", i) and re.match(r"for\s*\(", src[i:]):
                j = src.index("(", i) + 1
                depth = 1
                while depth:
                    depth += {"(": 1, ")": -1}.get(src[j], 0)
                    j += 1
                header = src[src.index("(", i) + 1:j - 1]
                body, i = self._body(src, j)
                toks.append(("for", header, body))
                continue
            if src.startswith("if", i) and re.match(r"if\s*\(", src[i:]):
                j = src.index("(", i) + 1
                depth = 1
                while depth:
                    depth += {"(": 1, ")": -1}.get(src[j], 0)
                    j += 1
                cond = src[src.index("(", i) + 1:j - 1]
                then, i = self._body(src, j)
                els = None
                k = i
                while k < n and src[k].isspace():
                    k += 1
                if src.startswith("else", k):
                    els, i = self._body(src, k + 4)
                toks.append(("if", cond, then, els))
                continue
            if ch == "{":
                body, i = self._body(src, i)
                toks.extend(body)
                continue
            j = src.index(";", i)
            toks.append(("s", src[i:j].strip()))
            i = j + 1
        return toks

    def _body(self, src, i):
        while src[i].isspace():
            i += 1
        if src[i] == "{":
            j, depth = i + 1, 1
            while depth:
                depth += {"{": 1, "}": -1}.get(src[j], 0)
                j += 1
            return self.stmts(src[i + 1:j - 1]), j
        j = src.index(";", i)
        return [("s", src[i:j].strip())], j + 1

    def emit(self, toks, ind):
        k = 0
        while k < len(toks):
            t = toks[k]
            if t[0] == "if":
                # every `if` in the shaders selects between returns: fold into a select
                rest = toks[k + 1:]
                e = self._ret_expr(t, rest)
                self.out.append(" " * ind + "return " + e)
                return
            if t[0] == "for":
                init, cond, step = (p.strip() for p in t[1].split(";"))
                m = re.match(TYPES + r"\s+(\w+)\s*=\s*(.*)$", init)
                var, val = m.group(1), m.group(2)
                self.out.append(" " * ind + "%s = %s" % (var, self.expr(val, True)))
                self.out.append(" " * ind + "while _loop(%s):" % self.expr(cond, True))
                self.emit(t[2], ind + 4)
                st = step
                mm = re.match(r"(\w+)\+\+$", st)
                if mm:
                    st = "%s += 1" % mm.group(1)
                self.out.append(" " * (ind + 4) + self.expr(st, True))
                self.out.append(" " * ind + "_loop_end()")
                k += 1
                continue
            self.stmt(t[1], ind)
            k += 1

    def _ret_expr(self, t, rest):
        """Expression of `if (c) A else B` where A/B (or the statements after) return."""
        _, cond, then, els = t
        a = self._branch_expr(then)
        b = self._branch_expr(els if els is not None else rest)
        return "glsl_mix(%s, %s, %s)" % (b, a, self.expr(cond))

    def _branch_expr(self, toks):
        t = toks[0]
        if t[0] == "if":
            return self._ret_expr(t, toks[1:])
        assert t[0] == "s" and t[1].startswith("return"), t
        return "(" + self.expr(t[1][len("return"):]) + ")"

    def stmt(self, s, ind):
        pad = " " * ind
        if not s:
            return
        if s.startswith("return"):
            self.out.append(pad + "return " + self.expr(s[6:]))
            return
        if s == "barrier()":
            self.out.append(pad + "barrier()")
            return
        m = re.match(r"shared\s+(float|vec3|vec4)\s+(\w+)\[(\d+)\]$", s)
        if m:
            self.out.append(pad + "%s = SHARED(%r, %s, %d)" % (m.group(2), m.group(2), m.group(3),
                                                            {"float": 1, "vec3": 3, "vec4": 4}[m.group(1)]))
            return
        m = re.match(r"(" + TYPES + r")\s+(.*)$", s)
        if m:
            typ = m.group(1)
            is_int = bool(re.search(r"\b(int|ivec2)\b", typ))
            for decl in _split_args(m.group(2)):
                dm = re.match(r"(\w+)\[(\d+)\]$", decl)
                if dm:
                    base = typ.split()[-1]
                    self.out.append(pad + "%s = [%s(0.0) for _ in range(%s)]" % (dm.group(1), base, dm.group(2)))
                    continue
                name, eq, val = decl.partition("=")
                if eq:
                    self.out.append(pad + "%s = %s" % (name.strip(), self.expr(val, is_int)))
                else:
                    self.out.append(pad + "%s = None" % name.strip())
            return
        m = re.match(r"^([\w\[\]\.\s+]+?)\s*([+\-*/])=\s*(.*)$", s)
        if m:
            lhs = m.group(1).strip()
            self.out.append(pad + "%s = %s %s (%s)" % (lhs, lhs, m.group(2), self.expr(m.group(3))))
            return
        self.out.append(pad + self.expr(s))


PY_KEYWORDS = ("lambda", "in", "is", "not", "and", "or", "pass", "from", "global", "def", "del", "yield")


def translate(body):
    """Python source of the shader: GLSL functions become defs, hook() becomes hook()."""
    src = expand_macros(body)
    for kw in PY_KEYWORDS:
        src = re.sub(r"\b%s\b" % kw, kw + "_", src)
    tr = Translator()
    i = 0
    pat = re.compile(r"^\s*(" + TYPES + r"|void)\s+(\w+)\s*\(([^)]*)\)\s*\{", re.M)
    pre = []
    while True:
        m = pat.search(src, i)
        if not m:
            pre.append(src[i:])
            break
        pre.append(src[i:m.start()])
        j, depth = m.end(), 1
        while depth:
            depth += {"{": 1, "}": -1}.get(src[j], 0)
            j += 1
        params = [re.sub(r"\[\d+\]", "", p).split()[-1] for p in m.group(3).split(",") if p.strip()]
        tr.out.append("def %s(%s):" % (m.group(2), ", ".join(params)))
        tr.emit(tr.stmts(src[m.end():j - 1]), 4)
        tr.out.append("    return None")
        i = j
    glob = "\n".join(pre)
    head = []
    for t in tr.stmts(glob) if glob.strip() else []:
        sub = Translator()
        sub.stmt(t[1], 0)
        head.extend(sub.out)
    return "\n".join(head + tr.out)


# ----------------------------------------------------------------------------- running


class _StopLoader(Exception):
    pass


def _loop(cond):
    """Loop condition, possibly an array over workgroups: run while any is true and mask
    the shared-memory writes of the finished ones."""
    if isinstance(cond, np.ndarray):
        Shared.active = cond
        return bool(cond.any())
    Shared.active = True
    return bool(cond)


def _loop_end():
    Shared.active = True


def run_hook(path, planes, target=None, fbo16=False):
    """Run every pass of the hook file on `planes` ((H,W) luma or (3,H,W)); returns the
    hooked plane(s) after the last pass.  `target` = (w, h) of mpv's OUTPUT.
    fbo16: every pass's output texture (SAVE or the replaced hooked plane) is stored in
    FP16 (round to nearest even), as mpv's default rgba16f FBOs store it; the source
    planes themselves are not rounded."""
    passes, luts = parse_hook_file(path)
    img = np.asarray(planes, dtype=np.float64)
    if img.ndim == 2:
        rgba = np.zeros(img.shape + (4,))
        rgba[..., 0] = img
        nch = 1
    else:
        rgba = np.zeros(img.shape[1:] + (4,))
        rgba[..., :3] = np.moveaxis(img, 0, -1)
        rgba[..., 3] = 1.0
        nch = 3
    hooked = Tex(rgba)
    saved = {}
    ow, oh = target if target is not None else (None, None)
    for p in passes:
        env = {"HOOKED.w": hooked.d.shape[1], "HOOKED.h": hooked.d.shape[0], "OUTPUT.w": ow, "OUTPUT.h": oh}
        out_w = int(rpn(p["WIDTH"], env)) if "WIDTH" in p else hooked.d.shape[1]
        out_h = int(rpn(p["HEIGHT"], env)) if "HEIGHT" in p else hooked.d.shape[0]
        code = translate(p["body"])
        g = dict(BUILTINS)
        g["glsl_mix"] = glsl_mix
        for b in p["BIND"]:
            t = hooked if b == "HOOKED" else (saved[b] if b in saved else luts.get(b))
            if b in luts:
                g[b] = luts[b]
            else:
                g[b + "_size"] = t.size
                g[b + "_pt"] = V([1.0 / t.size.c[0], 1.0 / t.size.c[1]])
                g[b + "_tex"] = (lambda tt: (lambda pos: tt.sample(pos)))(t)
                g[b + "_raw"] = t
                g[b + "_mul"] = 1.0
        g["texture"] = lambda lut, pos: lut.sample(pos)
        g["textureGatherOffset"] = texture_gather_offset
        g["_loop"] = _loop
        g["_loop_end"] = _loop_end
        g["HOOKED_map"] = (lambda ow_, oh_: (lambda v: V([(v.c[0] + 0.5) / ow_, (v.c[1] + 0.5) / oh_])))(out_w, out_h)
        result = np.zeros((out_h, out_w, 4))
        if "COMPUTE" in p:
            cv = [int(v) for v in p["COMPUTE"].split()]
            bw, bh, tw, th = cv if len(cv) == 4 else cv + cv   # threads default to the block
            ngx, ngy = -(-out_w // bw), -(-out_h // bh)
            G = ngx * ngy
            gid_x = np.repeat(np.arange(ngx), ngy)
            gid_y = np.tile(np.arange(ngy), ngx)
            shared = {}

            def SHARED(name, n, comps):
                if name not in shared:
                    s = Shared(n, comps)
                    s.store = np.zeros((G, n) + ((comps,) if comps > 1 else ()))
                    shared[name] = s
                return shared[name]
            g["SHARED"] = SHARED
            g["gl_WorkGroupSize"] = V([tw, th])

            exec(code, g)
            # loader: every local index, vectorised over workgroups
            g["gl_WorkGroupID"] = V([gid_x, gid_y])

            def barrier_load():
                raise _StopLoader()
            g["barrier"] = barrier_load
            for s in shared.values():
                s.mode = "load"
            for lid in range(tw * th):
                g["gl_LocalInvocationIndex"] = lid
                g["gl_LocalInvocationID"] = V([lid % tw, lid // tw])
                g["gl_GlobalInvocationID"] = V([gid_x * tw + lid % tw, gid_y * th + lid // tw])
                for s in shared.values():
                    s.mode = "load"
                try:
                    g["hook"]()
                except _StopLoader:
                    pass
            # invocations: all (group, local) pairs at once
            lids = np.arange(tw * th)
            gi = np.repeat(np.arange(G), tw * th)
            li = np.tile(lids, G)
            for s in shared.values():
                s.mode = "run"
                s.gidx = gi
            gx, gy = gid_x[gi], gid_y[gi]
            g["gl_WorkGroupID"] = V([gx, gy])
            g["gl_LocalInvocationIndex"] = 1 << 30   # the loader loop does not run again
            g["gl_LocalInvocationID"] = V([li % tw, li // tw])
            g["gl_GlobalInvocationID"] = V([gx * tw + li % tw, gy * th + li // tw])
            g["barrier"] = lambda: None
            g["out_image"] = result

            def imageStore(img_, pos, val):
                x = np.asarray(pos.c[0])
                y = np.asarray(pos.c[1])
                ok = (x >= 0) & (x < out_w) & (y >= 0) & (y < out_h)
                for k in range(4):
                    v = np.broadcast_to(np.asarray(val.c[k], dtype=np.float64), x.shape)
                    img_[y[ok], x[ok], k] = v[ok]
            g["imageStore"] = imageStore
            # loader writes are ignored in run mode; the loop still evaluates harmlessly
            for s in shared.values():
                s.mode = "run"
            g["hook"]()
        else:
            yy, xx = np.mgrid[0:out_h, 0:out_w]
            xx = xx.reshape(-1).astype(np.float64)
            yy = yy.reshape(-1).astype(np.float64)
            for b in p["BIND"]:
                if b in luts:
                    continue
                t = hooked if b == "HOOKED" else saved[b]
                pos = V([(xx + 0.5) / out_w, (yy + 0.5) / out_h])
                g[b + "_pos"] = pos
                g[b + "_texOff"] = (lambda tt, pp: (lambda off: tt.sample(
                    pp + off * V([1.0 / tt.size.c[0], 1.0 / tt.size.c[1]]))))(t, pos)
            exec(code, g)
            r = g["hook"]()
            for k in range(4):
                result[..., k] = np.broadcast_to(np.asarray(r.c[k], dtype=np.float64), xx.shape).reshape(out_h, out_w)
        if fbo16:
            result = result.astype(np.float16).astype(np.float64)
        tex = Tex(result)
        if "SAVE" in p:
            saved[p["SAVE"]] = tex
        else:
            hooked = tex
    d = hooked.d
    return d[..., 0] if nch == 1 else np.moveaxis(d[..., :3], -1, 0)
