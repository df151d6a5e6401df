"""SymPy pair potential -> HIP device functions (compiled with hiprtc).

Replaces the lambdify + numba.njit step of dompap's make_pair_potential
(potential.py:43-66): the same SymPy pipeline (``sympify``; f = d(-u)/dr;
u'' = d^2u/dr^2) but printed as C++ templates over the floating type, so one
expression yields both the FP64 parity kernels and the FP32 mixed kernels.

Generated functions (x = r / sigma):

* ``dp_pot_u<T>(x)``  u(x)          (the kernel applies the r_cut shift)
* ``dp_pot_f<T>(x)``  f(x) = -u'(x)
* ``dp_pot_c<T>(x)``  u''(x)
* ``dp_pot_gs<T>(s)`` f(x)/x as a function of s = x^2, and
  ``dp_pot_us<T>(s)`` u(x) as a function of s, when both are expressible
  without odd powers of x ("even" potentials such as Lennard-Jones or r^-n):
  the mixed kernel then needs neither sqrt nor a division per pair;
* ``dp_pot_gsk<T>(s, a)`` (f(x)/x for a polynomial in 1/s of at most four
  terms) with runtime coefficients into which the host folds the pair
  kernel's scales, and its ``// dp-gsk:`` line of exponents and coefficients.
"""
from __future__ import annotations

from dataclasses import dataclass

import sympy as sp
from sympy.printing.c import C99CodePrinter

_PRELUDE = r"""
// ---- generated by dompap_amd.codegen ----
template <int N, class T> __device__ __forceinline__ T dp_ipow(T x) {
  if constexpr (N == 0) return T(1);
  else if constexpr (N == 1) return x;
  else if constexpr (N % 2 == 0) { const T h = dp_ipow<N / 2, T>(x); return h * h; }
  else return x * dp_ipow<N - 1, T>(x);
}
// hardware reciprocal refined by one Newton step
__device__ __forceinline__ float dp_rcp(float x) {
  const float y = __builtin_amdgcn_rcpf(x);
  return fmaf(y, fmaf(-x, y, 1.0f), y);
}
#if DP_MODE == 1
// mixed mode: hardware reciprocal refined by one Newton step (the relative
// error e of the estimate becomes e^2) instead of the correctly rounded
// quotient
__device__ __forceinline__ double dp_rcp(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  return fma(y, fma(-x, y, 1.0), y);
}
#else
__device__ __forceinline__ double dp_rcp(double x) { return 1.0 / x; }
#endif
__device__ __forceinline__ float dp_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double dp_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float dp_pow(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double dp_pow(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float dp_exp(float a) { return expf(a); }
__device__ __forceinline__ double dp_exp(double a) { return exp(a); }
__device__ __forceinline__ float dp_log(float a) { return logf(a); }
__device__ __forceinline__ double dp_log(double a) { return log(a); }
__device__ __forceinline__ float dp_sin(float a) { return sinf(a); }
__device__ __forceinline__ double dp_sin(double a) { return sin(a); }
__device__ __forceinline__ float dp_cos(float a) { return cosf(a); }
__device__ __forceinline__ double dp_cos(double a) { return cos(a); }
__device__ __forceinline__ float dp_tan(float a) { return tanf(a); }
__device__ __forceinline__ double dp_tan(double a) { return tan(a); }
__device__ __forceinline__ float dp_tanh(float a) { return tanhf(a); }
__device__ __forceinline__ double dp_tanh(double a) { return tanh(a); }
__device__ __forceinline__ float dp_sinh(float a) { return sinhf(a); }
__device__ __forceinline__ double dp_sinh(double a) { return sinh(a); }
__device__ __forceinline__ float dp_cosh(float a) { return coshf(a); }
__device__ __forceinline__ double dp_cosh(double a) { return cosh(a); }
__device__ __forceinline__ float dp_abs(float a) { return fabsf(a); }
__device__ __forceinline__ double dp_abs(double a) { return fabs(a); }
__device__ __forceinline__ float dp_erf(float a) { return erff(a); }
__device__ __forceinline__ double dp_erf(double a) { return erf(a); }
__device__ __forceinline__ float dp_erfc(float a) { return erfcf(a); }
__device__ __forceinline__ double dp_erfc(double a) { return erfc(a); }
template <class T> __device__ __forceinline__ T dp_max(T a, T b) { return a > b ? a : b; }
template <class T> __device__ __forceinline__ T dp_min(T a, T b) { return a < b ? a : b; }
template <class T> __device__ __forceinline__ T dp_heaviside(T a) { return a > T(0) ? T(1) : (a < T(0) ? T(0) : T(0.5)); }
"""

_FUNCS = {
    "exp": "dp_exp", "log": "dp_log", "sin": "dp_sin", "cos": "dp_cos", "tan": "dp_tan",
    "tanh": "dp_tanh", "sinh": "dp_sinh", "cosh": "dp_cosh", "erf": "dp_erf", "erfc": "dp_erfc",
}


class _HipPrinter(C99CodePrinter):
    """C printer producing type-generic code (``T`` is the float type)."""

    def _print_Float(self, expr):
        return f"T({float(expr)!r})"

    def _print_Integer(self, expr):
        return f"T({int(expr)})"

    def _print_Rational(self, expr):
        return f"(T({int(expr.p)}) / T({int(expr.q)}))"

    def _print_NumberSymbol(self, expr):
        return f"T({float(expr)!r})"

    def _print_Exp1(self, expr):
        return f"T({float(expr)!r})"

    def _print_Pi(self, expr):
        return f"T({float(expr)!r})"

    def _print_Infinity(self, expr):
        return "T(__builtin_huge_val())"

    def _print_NegativeInfinity(self, expr):
        return "T(-__builtin_huge_val())"

    def _print_Pow(self, expr):
        base, ex = expr.as_base_exp()
        b = self.parenthesize(base, 1000) if not base.is_Atom else self._print(base)
        b = f"({self._print(base)})"
        if ex.is_Integer:
            n = int(ex)
            if n == 0:
                return "T(1)"
            if n > 0:
                return f"dp_ipow<{n}, T>{b}"
            return f"dp_rcp(dp_ipow<{-n}, T>{b})"
        if ex == sp.Rational(1, 2):
            return f"dp_sqrt{b}"
        if ex == sp.Rational(-1, 2):
            return f"dp_rcp(dp_sqrt{b})"
        if ex.is_Rational and ex.q == 2:
            n = int(ex.p)
            if n > 0:
                return f"(dp_sqrt{b} * dp_ipow<{n // 2}, T>{b})"
        return f"dp_pow({b}, {self._print(ex)})"

    def _print_Mul(self, expr):
        # a * b**-1 -> a * dp_rcp(b); keep the default structure otherwise
        return super()._print_Mul(expr)

    def _print_Function(self, expr):
        name = type(expr).__name__.lower()
        if name in _FUNCS:
            return f"{_FUNCS[name]}({', '.join(self._print(a) for a in expr.args)})"
        return super()._print_Function(expr)

    def _print_Abs(self, expr):
        return f"dp_abs({self._print(expr.args[0])})"

    def _print_Max(self, expr):
        args = [self._print(a) for a in expr.args]
        out = args[0]
        for a in args[1:]:
            out = f"dp_max({out}, {a})"
        return out

    def _print_Min(self, expr):
        args = [self._print(a) for a in expr.args]
        out = args[0]
        for a in args[1:]:
            out = f"dp_min({out}, {a})"
        return out

    def _print_DiracDelta(self, expr):
        return "T(0)"

    def _print_sign(self, expr):
        a = self._print(expr.args[0])
        return f"(({a}) > T(0) ? T(1) : (({a}) < T(0) ? T(-1) : T(0)))"

    def _print_Heaviside(self, expr):
        return f"dp_heaviside({self._print(expr.args[0])})"

    def _print_Piecewise(self, expr):
        out = "T(0)"
        for e, c in reversed(expr.args):
            if c == sp.true:
                out = self._print(e)
            else:
                out = f"(({self._print(c)}) ? ({self._print(e)}) : ({out}))"
        return out

    def _print_BooleanTrue(self, expr):
        return "true"

    def _print_BooleanFalse(self, expr):
        return "false"


_PRINTER = _HipPrinter()


def _emit(name: str, expr: sp.Expr, var: sp.Symbol) -> str:
    """Device function ``name<T>(var)`` evaluating ``expr``; a polynomial in
    1/var is evaluated in Horner form from one reciprocal."""
    expr = sp.sympify(expr)
    body_lines = []
    q = sp.Symbol("dp_q", positive=True)
    inv = sp.expand(expr.subs(var, 1 / q))
    use_inv = False
    if expr.free_symbols == {var} and inv.is_polynomial(q) and not expr.is_polynomial(var):
        use_inv = True
        poly = sp.horner(sp.Poly(inv, q).as_expr(), wrt=q)
    if use_inv:
        body_lines.append(f"const T dp_q = dp_rcp({var.name});")
        result = _PRINTER.doprint(poly)
    else:
        repl, reduced = sp.cse(expr, symbols=sp.numbered_symbols("dp_t"))
        for sym, sub in repl:
            body_lines.append(f"const T {sym} = {_PRINTER.doprint(sub)};")
        result = _PRINTER.doprint(reduced[0])
    if var.name not in result and not any(var.name in line for line in body_lines):
        body_lines.insert(0, f"(void){var.name};")
    body = "\n  ".join(body_lines + [f"return {result};"])
    return f"template <class T> __device__ __forceinline__ T {name}(T {var.name}) {{\n  {body}\n}}\n"


def _emit_scaled(name: str, expr: sp.Expr, var: sp.Symbol):
    """Device function ``name<T>(var)`` evaluating ``expr / C`` and the
    rational C: for a polynomial in 1/var, C is the content of its
    coefficients (48 q^7 - 24 q^4 becomes 24 * q^4 (2 q^3 - 1), whose inner
    factor is one FMA with inline constants); C = 1 otherwise.  The pair
    kernels fold C into their per-type-pair prefactor."""
    expr = sp.sympify(expr)
    q = sp.Symbol("dp_q", positive=True)
    inv = sp.expand(expr.subs(var, 1 / q))
    if expr.free_symbols == {var} and inv.is_polynomial(q) and not expr.is_polynomial(var):
        content, prim = sp.Poly(inv, q).primitive()
        if content != 0 and prim.LC() < 0:
            content, prim = -content, -prim
        if content != 0:
            poly = sp.horner(prim.as_expr(), wrt=q)
            body = f"const T dp_q = dp_rcp({var.name});\n  return {_PRINTER.doprint(poly)};"
            src = f"template <class T> __device__ __forceinline__ T {name}(T {var.name}) {{\n  {body}\n}}\n"
            return src, sp.Rational(content)
    return _emit(name, expr, var), sp.Integer(1)


def _emit_folded(name: str, expr: sp.Expr, var: sp.Symbol):
    """For ``expr`` a polynomial in q = 1/var, the device function
    ``name<T>(var, a)`` = sum_k a[k] * z^e_k with z = 1/var, evaluated as
    z^e_0 (a[0] + z^(e_1 - e_0) (a[1] + ...)), and the comment line
    ``// dp-gsk: n e_0 c_0 e_1 c_1 ...`` giving the exponents e_k and the
    coefficients c_k of expr = sum_k c_k var^-e_k.  The host folds its
    per-type-pair scales into a[k] = ke * c_k * ks^-e_k (ks: the scale of
    var, ke: the prefactor), so that the kernel evaluates the scaled term
    with neither the scale multiply nor the prefactor multiply.  None when
    expr is not such a polynomial or has more than 4 terms."""
    expr = sp.sympify(expr)
    q = sp.Symbol("dp_q", positive=True)
    inv = sp.expand(expr.subs(var, 1 / q))
    if not (expr.free_symbols == {var} and inv.is_polynomial(q) and not expr.is_polynomial(var)):
        return None
    content, prim = sp.Poly(inv, q).primitive()
    if content == 0:
        return None
    if prim.LC() < 0:
        prim = -prim
    terms = sorted((int(m[0]), sp.Rational(c)) for m, c in prim.terms() if c != 0)
    sign = 1 if sp.Poly(inv, q).LC() * prim.LC() > 0 else -1
    if not terms or len(terms) > 4:
        return None
    z = "dp_z"
    inner = f"a[{len(terms) - 1}]"
    for k in range(len(terms) - 2, -1, -1):
        de = terms[k + 1][0] - terms[k][0]
        inner = f"(a[{k}] + dp_ipow<{de}, T>({z}) * {inner})"
    e0 = terms[0][0]
    body = inner if e0 == 0 else f"dp_ipow<{e0}, T>({z}) * {inner}"
    # the coefficients of expr itself (content times primitive part)
    meta = " ".join(f"{e} {float(sign * content * c)!r}" for e, c in terms)
    src = (f"#define DP_GSK_N {len(terms)}\n"
           f"// dp-gsk: {len(terms)} {meta}\n"
           f"template <class T> __device__ __forceinline__ T {name}(T {var.name}, const T* a) {{\n"
           f"  const T {z} = dp_rcp({var.name});\n  return {body};\n}}\n")
    return src


def _is_even_form(expr_s: sp.Expr, s: sp.Symbol) -> bool:
    for p in sp.preorder_traversal(expr_s):
        if isinstance(p, sp.Pow) and p.base.has(s) and not p.exp.is_Integer:
            return False
    return True


@dataclass(frozen=True)
class PotentialCode:
    source: str        # device source for dp_load_potential
    even: bool         # dp_pot_gs / dp_pot_us defined
    expression: str    # the potential string it was made from


HARMONIC_REPULSIVE = "Harmonic repulsive"


def generate(pair_potential_str: str) -> PotentialCode:
    """Device source for ``pair_potential_str`` (a SymPy expression in r, or
    the reference's hard-coded name "Harmonic repulsive", potential.py:10-40,
    whose curvature function returns 0.0)."""
    # plain symbols, as the reference (potential.py:47): assumptions would let
    # SymPy rewrite e.g. sqrt((1-r)**2) as Abs(1-r)
    r = sp.Symbol("r")
    x = sp.Symbol("x")
    s = sp.Symbol("s", positive=True)
    if pair_potential_str == HARMONIC_REPULSIVE:
        u = sp.Rational(1, 2) * (1 - x) ** 2
        f = 1 - x
        c = sp.Integer(0)
    else:
        u0 = sp.sympify(pair_potential_str)
        f0 = sp.diff(-u0, r)
        c0 = sp.diff(u0, r, 2)
        u, f, c = (e.subs(r, x) for e in (u0, f0, c0))
    parts = [_PRELUDE,
             _emit("dp_pot_u", u, x),
             _emit("dp_pot_f", f, x),
             _emit("dp_pot_c", c, x)]
    g_s = sp.expand(sp.powsimp(sp.expand(f / x).subs(x, sp.sqrt(s)), force=True))
    u_s = sp.expand(sp.powsimp(sp.expand(u).subs(x, sp.sqrt(s)), force=True))
    even = _is_even_form(g_s, s) and _is_even_form(u_s, s)
    if even:
        parts.append(_emit("dp_pot_gs", g_s, s))
        parts.append(_emit("dp_pot_us", u_s, s))
        # mixed-mode force term: gs / DP_GS_SCALE (the scale goes into the
        # kernels' prefactor)
        src, scale = _emit_scaled("dp_pot_gsn", g_s, s)
        parts.append(f"#define DP_GS_SCALE ({float(scale)!r})\n" + src)
        # the same term with the scales folded into its coefficients
        folded = _emit_folded("dp_pot_gsk", g_s, s)
        if folded is not None:
            parts.append(folded)
    return PotentialCode(source="\n".join(parts), even=even, expression=pair_potential_str)
