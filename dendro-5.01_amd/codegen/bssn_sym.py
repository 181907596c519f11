"""Symbolic statement of the BSSN right-hand side and its C/HIP emitters.

This is the project's own generator for the 24 BSSN evolution equations that
Dendro-5 evaluates per grid point inside ``bssnrhs``.  It states the same
system the reference states in ``CodeGen/bssn.py:41-88`` (helpers in
``GR/rhs_scripts/bssn/dendro.py``) with the constant-eta, standard-gauge
choices that ``BSSN_GR/src/rhs.cpp:228-247`` selects by default:

* 1+log lapse with advection (``a_rhs``), Gamma-driver shift (``b_rhs``),
* conformal metric / conformal factor / traceless extrinsic curvature,
  trace K, conformal connection ``Gt`` and ``B`` with constant damping
  ``eta`` (``B_rhs = Gt_rhs - eta*B + l3 beta.adv(B) - l4 beta.adv(Gt)``).

Derivative symbols follow the reference naming so that every derivative is
the same discrete operator in the same place:

* ``grad_i_f``   first derivative (deriv42_{x,y,z}),
* ``grad2_i_j_f`` second derivative, i<=j, mixed ones by composition,
* ``agrad_i_f``  upwinded first derivative (deriv42adv_{x,y,z}).

Every ``lie``/advection term uses ``agrad``; every other first derivative
uses ``grad`` (``dendro.py`` ``lie`` / ``vec_j_ad_j``).

Two emitters exist:

``emit_cse``      SymPy common-subexpression elimination over the fully
                  expanded system, the same method the reference generator
                  uses (``CodeGen/dendro.py:536-624``).  Used by the CPU
                  oracle so that its evaluation order stays close to the
                  reference's.
``emit_staged``   A tensor-staged form (inverse metric, Christoffels, Ricci
                  pieces as named intermediates, advection pre-contracted
                  with beta) for the GPU kernels; fewer live values.

Run ``python bssn_sym.py --out-oracle ... --out-io ... --out-physcon-io ...`` to regenerate.
"""
from __future__ import annotations

import argparse
import sys
import time

import sympy as sp
from sympy.printing.c import C99CodePrinter
from sympy.printing.precedence import PRECEDENCE

E_I = [0, 1, 2]
E_IJ = [(i, j) for i in E_I for j in E_I]
SYM_PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
SYM_IDX = {p: n for n, p in enumerate(SYM_PAIRS)}

# Variable order of the reference's VAR enum (BSSN_GR/include/grDef.h:31).
VAR_NAMES = ["alpha", "chi", "K", "Gt0", "Gt1", "Gt2", "beta0", "beta1",
             "beta2", "B0", "B1", "B2", "gt0", "gt1", "gt2", "gt3", "gt4",
             "gt5", "At0", "At1", "At2", "At3", "At4", "At5"]

# Output names in VAR order (rhs arrays in rhs.cpp:125-148).
RHS_NAMES = ["a_rhs", "chi_rhs", "K_rhs", "Gt_rhs0", "Gt_rhs1", "Gt_rhs2",
             "b_rhs0", "b_rhs1", "b_rhs2", "B_rhs0", "B_rhs1", "B_rhs2",
             "gt_rhs00", "gt_rhs01", "gt_rhs02", "gt_rhs11", "gt_rhs12",
             "gt_rhs22", "At_rhs00", "At_rhs01", "At_rhs02", "At_rhs11",
             "At_rhs12", "At_rhs22"]

# Variables whose first derivatives enter the equations
# (BSSN_GR/scripts/bssnrhs_derivs.h: 24 vars get grad, but At and B first
# derivatives are used only by the Sommerfeld boundary terms).
GRAD_VARS = ["alpha", "beta0", "beta1", "beta2", "chi", "Gt0", "Gt1", "Gt2",
             "K", "gt0", "gt1", "gt2", "gt3", "gt4", "gt5"]
# Variables with second derivatives (bssnrhs_derivs.h grad2_*).
GRAD2_VARS = ["gt0", "gt1", "gt2", "gt3", "gt4", "gt5", "chi", "alpha",
              "beta0", "beta1", "beta2"]


def _s(name):
    return sp.Symbol(name, real=True)


def grad(i, f):
    return _s("grad_%d_%s" % (i, f.name))


def grad2(i, j, f):
    if i > j:
        i, j = j, i
    return _s("grad2_%d_%d_%s" % (i, j, f.name))


def agrad(i, f):
    return _s("agrad_%d_%s" % (i, f.name))


def sym3(name):
    m = [_s("%s%d" % (name, n)) for n in range(6)]
    return sp.Matrix([[m[0], m[1], m[2]], [m[1], m[3], m[4]], [m[2], m[4], m[5]]])


class BSSNSystem:
    """The BSSN equations as SymPy expressions (reference math, own code)."""

    def __init__(self):
        self.l = [_s("lambda%d" % i) for i in range(4)]
        self.lf = [_s("lambda_f%d" % i) for i in range(2)]
        self.eta = _s("eta")
        self.a = _s("alpha")
        self.chi = _s("chi")
        self.K = _s("K")
        self.Gt = tuple(_s("Gt%d" % i) for i in E_I)
        self.b = tuple(_s("beta%d" % i) for i in E_I)
        self.B = tuple(_s("B%d" % i) for i in E_I)
        self.gt = sym3("gt")
        self.At = sym3("At")

    # -- metric quantities (dendro.py get_inverse_metric / christoffels) --
    def build(self):
        gt, chi = self.gt, self.chi
        # inverse via adjugate (dendro.py:345, metric.inv('ADJ') simplified)
        self.igt = sp.simplify(gt.inv("ADJ"))
        igt = self.igt
        half = sp.Float(0.5)
        C1 = {}
        for k in E_I:
            for i in E_I:
                for j in E_I:
                    C1[k, i, j] = half * (grad(j, gt[k, i]) + grad(i, gt[k, j])
                                          - grad(k, gt[i, j]))
        C2 = {}
        for a in E_I:
            for d in E_I:
                for e in E_I:
                    C2[a, d, e] = sum(igt[a, b] * C1[b, d, e] for b in E_I)
        C3 = {}
        for i in E_I:
            for j in E_I:
                for k in E_I:
                    C3[i, j, k] = C2[i, j, k] - half / chi * (
                        sp.KroneckerDelta(i, j) * grad(k, chi)
                        + sp.KroneckerDelta(i, k) * grad(j, chi)
                        - gt[j, k] * sum(igt[i, m] * grad(m, chi) for m in E_I))
        self.C1, self.C2, self.C3 = C1, C2, C3

        # Ricci tensor (dendro.py compute_ricci)
        CalGt = [sum(igt[k, l] * C2[i, k, l] for k, l in E_IJ) for i in E_I]
        Gt = self.Gt
        Rt = sp.zeros(3, 3)
        for i, j in E_IJ:
            Rt[i, j] = (-half * sum(igt[l, m] * grad2(l, m, gt[i, j]) for l, m in E_IJ)
                        + half * sum(gt[k, i] * grad(j, Gt[k]) + gt[k, j] * grad(i, Gt[k]) for k in E_I)
                        + half * sum(CalGt[k] * (C1[i, j, k] + C1[j, i, k]) for k in E_I)
                        + sum(igt[l, m] * (C2[k, l, i] * C1[j, k, m] + C2[k, l, j] * C1[i, k, m]
                                           + C2[k, i, m] * C1[k, l, j])
                              for k in E_I for l, m in E_IJ))
        Rphi = sp.zeros(3, 3)
        for i, j in E_IJ:
            x = (1 / (2 * chi) * (grad2(i, j, chi) - sum(C2[k, j, i] * grad(k, chi) for k in E_I))
                 - 1 / (4 * chi * chi) * grad(i, chi) * grad(j, chi))
            y = 1 / (2 * chi) * gt[i, j] * (
                sum(igt[k, l] * (grad2(k, l, chi) - 3 / (2 * chi) * grad(k, chi) * grad(l, chi))
                    for k, l in E_IJ)
                - sum(CalGt[m] * grad(m, chi) for m in E_I))
            Rphi[i, j] = x + y
        self.CalGt = CalGt
        self.R = Rt + Rphi
        return self

    # -- operators (dendro.py lie / DiDj / laplacian / trace_free / sqr) --
    def lie_scalar(self, f, weight=0):
        b = self.b
        return sum(b[i] * agrad(i, f) for i in E_I) + weight * f * sum(grad(i, b[i]) for i in E_I)

    def lie_sym(self, A, weight):
        b = self.b
        m = sp.zeros(3, 3)
        for i, j in E_IJ:
            m[i, j] = sum(b[k] * agrad(k, A[i, j]) + A[i, k] * grad(j, b[k])
                          + A[k, j] * grad(i, b[k]) + weight * A[i, j] * grad(k, b[k])
                          for k in E_I)
        return m

    def adv(self, f):
        return sum(self.b[i] * agrad(i, f) for i in E_I)

    def DiDj(self, f):
        m = sp.zeros(3, 3)
        for i, j in E_IJ:
            m[i, j] = grad2(i, j, f) - sum(self.C3[l, i, j] * grad(l, f) for l in E_I)
        return m

    def laplacian(self, f):
        full = self.gt / self.chi
        ifull = sp.simplify(full.inv("ADJ"))
        return sum(ifull[i, j] * (grad2(i, j, f) - sum(self.C3[l, i, j] * grad(l, f) for l in E_I))
                   for i, j in E_IJ)

    def trace_free(self, x):
        tr = sum(self.igt[i, j] * x[i, j] for i, j in E_IJ)
        m = sp.zeros(3, 3)
        for i, j in E_IJ:
            m[i, j] = x[i, j] - self.gt[i, j] * tr / 3
        return m

    def sqr(self, A):
        igt = self.igt
        return sum(A[i, j] * sum(igt[i, k] * igt[j, l] * A[k, l] for k in E_I for l in E_I)
                   for i, j in E_IJ)

    def up_up(self, A):
        igt = self.igt
        m = sp.zeros(3, 3)
        for i, j in E_IJ:
            m[i, j] = sum(igt[i, k] * igt[j, l] * A[k, l] for k, l in E_IJ)
        return m

    # -- the 24 right-hand sides (CodeGen/bssn.py:60-88, eta constant) --
    def equations(self):
        a, chi, K, Gt, b, B, gt, At = (self.a, self.chi, self.K, self.Gt,
                                       self.b, self.B, self.gt, self.At)
        l1, l2, l3, l4 = self.l
        lf0, lf1 = self.lf
        igt, C2 = self.igt, self.C2
        weight = -sp.Rational(2, 3)

        a_rhs = l1 * self.lie_scalar(a) - 2 * a * K
        b_rhs = [sp.Rational(3, 4) * (lf0 + lf1 * a) * B[i] + l2 * self.adv(b[i]) for i in E_I]
        gt_rhs = self.lie_sym(gt, weight) - 2 * a * At
        chi_rhs = self.lie_scalar(chi, weight) + sp.Rational(2, 3) * (chi * a * K)
        AikAkj = sp.zeros(3, 3)
        for i, j in E_IJ:
            AikAkj[i, j] = sum(At[i, k] * sum(igt[k, l] * At[l, j] for l in E_I) for k in E_I)
        At_rhs = (self.lie_sym(At, weight) + chi * self.trace_free(a * self.R - self.DiDj(a))
                  + a * (K * At - 2 * AikAkj))
        K_rhs = self.lie_scalar(K) - self.laplacian(a) + a * (K * K / 3 + self.sqr(At))
        AUU = self.up_up(At)
        CalGt = self.CalGt
        Gt_rhs = []
        for i in E_I:
            e = sum(b[j] * agrad(j, Gt[i]) for j in E_I)
            e -= sum(CalGt[j] * grad(j, b[i]) for j in E_I)
            e += sp.Rational(2, 3) * CalGt[i] * sum(grad(j, b[j]) for j in E_I)
            e += sum(igt[j, k] * grad2(j, k, b[i]) + igt[i, j] * grad2(j, k, b[k]) / 3 for j, k in E_IJ)
            e -= sum(2 * AUU[i, j] * grad(j, a) for j in E_I)
            e += sum(2 * a * C2[i, j, k] * AUU[j, k] for j, k in E_IJ)
            e -= sum(a * (3 / chi * AUU[i, j] * grad(j, chi)
                          + sp.Rational(4, 3) * igt[i, j] * grad(j, K)) for j in E_I)
            Gt_rhs.append(e)
        B_rhs = [Gt_rhs[i] - self.eta * B[i] + l3 * self.adv(B[i]) - l4 * self.adv(Gt[i])
                 for i in E_I]

        out = {}
        out["a_rhs"] = a_rhs
        out["chi_rhs"] = chi_rhs
        out["K_rhs"] = K_rhs
        for i in E_I:
            out["Gt_rhs%d" % i] = Gt_rhs[i]
            out["b_rhs%d" % i] = b_rhs[i]
            out["B_rhs%d" % i] = B_rhs[i]
        for (i, j) in SYM_PAIRS:
            out["gt_rhs%d%d" % (i, j)] = gt_rhs[i, j]
            out["At_rhs%d%d" % (i, j)] = At_rhs[i, j]
        return [out[n] for n in RHS_NAMES]


def input_names():
    """Ordered list of point inputs used by the equations (oracle layout)."""
    names = list(VAR_NAMES)
    for v in GRAD_VARS:
        for i in E_I:
            names.append("grad_%d_%s" % (i, v))
    for v in GRAD2_VARS:
        for (i, j) in SYM_PAIRS:
            names.append("grad2_%d_%d_%s" % (i, j, v))
    for v in VAR_NAMES:
        for i in E_I:
            names.append("agrad_%d_%s" % (i, v))
    return names


PARAM_NAMES = ["lambda0", "lambda1", "lambda2", "lambda3", "lambda_f0", "lambda_f1", "eta"]


class _Printer(C99CodePrinter):
    """C printer: integer powers as products, doubles everywhere."""

    def _print_Pow(self, expr):
        b, e = expr.as_base_exp()
        if e.is_Integer:
            n = int(e)
            if n == 2:
                s = self.parenthesize(b, PRECEDENCE["Mul"])
                return "(%s*%s)" % (s, s)
            if n == -1:
                return "(1.0/%s)" % self.parenthesize(b, PRECEDENCE["Atom"])
        return "pow(%s, %s)" % (self._print(b), self._print(sp.Float(e) if e.is_Rational else e))

    def _print_Rational(self, expr):
        return "(%d.0/%d.0)" % (expr.p, expr.q)

    def _print_Integer(self, expr):
        return "%d" % int(expr)

    def _print_Float(self, expr):
        return repr(float(expr))


def _ccode(e):
    return _Printer().doprint(e)


def emit_cse(sys_, fname, header_comment):
    """Fully expanded system -> SymPy CSE -> a C++ point function."""
    t0 = time.time()
    exprs = sys_.equations()
    t1 = time.time()
    syms = sp.numbered_symbols(prefix="T")
    repl, red = sp.cse(exprs, symbols=syms, optimizations="basic")
    t2 = time.time()
    ops = sum(sp.count_ops(v) for _, v in repl) + sum(sp.count_ops(e) for e in red)
    lines = []
    lines.append("// " + header_comment)
    lines.append("// generated by dendro-5.01_amd/codegen/bssn_sym.py (emit_cse); do not edit.")
    lines.append("// temporaries: %d, reduced ops: %d" % (len(repl), ops))
    lines.append("// point inputs are plain local doubles named as in bssn_sym.input_names();")
    lines.append("// outputs are written to out[0..23] in VAR order.")
    for lhs, rhs in repl:
        lines.append("const double %s = %s;" % (lhs, _ccode(rhs)))
    for n, e in enumerate(red):
        lines.append("out[%d] = %s;  // %s" % (n, _ccode(e), RHS_NAMES[n]))
    with open(fname, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("emit_cse: build %.1fs cse %.1fs temps %d ops %d" % (t1 - t0, t2 - t1, len(repl), ops),
          file=sys.stderr)
    return len(repl), ops


def emit_io_header(fname):
    """X-macro list of the point inputs / outputs (shared by the oracle and
    the reference harness in oracle/)."""
    names = input_names()
    lines = ["/* generated by dendro-5.01_amd/codegen/bssn_sym.py (emit_io_header); do not edit.",
             " * Point inputs of the BSSN equations in a fixed order: 24 state",
             " * variables (VAR order, BSSN_GR/include/grDef.h:31), 45 first",
             " * derivatives, 66 second derivatives, 72 upwinded derivatives. */",
             "#ifndef DGR_BSSN_POINT_IO_H", "#define DGR_BSSN_POINT_IO_H",
             "#define DGR_POINT_NIN %d" % len(names),
             "#define DGR_POINT_INPUTS(X) \\"]
    for k, n in enumerate(names):
        lines.append("  X(%s, %d) \\" % (n, k))
    lines.append("")
    lines.append("#define DGR_POINT_OUTPUTS(X) \\")
    for k, n in enumerate(RHS_NAMES):
        lines.append("  X(%s, %d) \\" % (n, k))
    lines.append("")
    lines.append("#endif")
    with open(fname, "w") as f:
        f.write("\n".join(lines) + "\n")


# Point inputs of the constraint equations (BSSN_GR/src/physconeqs.cpp,
# derivatives listed in BSSN_GR/scripts/constraint_derivs.h): state values,
# the point coordinates, first derivatives of 17 variables and second
# derivatives of chi and the conformal metric.
PHYSCON_STATE = ["chi", "K", "Gt0", "Gt1", "Gt2", "gt0", "gt1", "gt2", "gt3", "gt4", "gt5",
                 "At0", "At1", "At2", "At3", "At4", "At5"]
PHYSCON_GRAD2 = ["chi", "gt0", "gt1", "gt2", "gt3", "gt4", "gt5"]
PHYSCON_OUT = ["ham", "mom0", "mom1", "mom2", "psi4_real", "psi4_img"]


def physcon_input_names():
    names = list(PHYSCON_STATE) + ["px", "py", "pz"]
    for v in PHYSCON_STATE:
        for i in E_I:
            names.append("grad_%d_%s" % (i, v))
    for v in PHYSCON_GRAD2:
        for (i, j) in SYM_PAIRS:
            names.append("grad2_%d_%d_%s" % (i, j, v))
    return names


def emit_physcon_io(fname):
    """X-macro list of the constraint point inputs / outputs (oracle and
    reference harness)."""
    names = physcon_input_names()
    lines = ["/* generated by dendro-5.01_amd/codegen/bssn_sym.py (emit_physcon_io); do not edit.",
             " * Point inputs of the constraint equations in a fixed order: 17 state",
             " * values, the coordinates px, py, pz, 51 first derivatives, 42 second",
             " * derivatives.  Outputs in VAR_CONSTRAINT order (grDef.h:34). */",
             "#ifndef DGR_PHYSCON_POINT_IO_H", "#define DGR_PHYSCON_POINT_IO_H",
             "#define DGR_PHYSCON_NIN %d" % len(names),
             "#define DGR_PHYSCON_INPUTS(X) \\"]
    for k, n in enumerate(names):
        lines.append("  X(%s, %d) \\" % (n, k))
    lines.append("")
    lines.append("#define DGR_PHYSCON_OUTPUTS(X) \\")
    for k, n in enumerate(PHYSCON_OUT):
        lines.append("  X(%s, %d) \\" % (n, k))
    lines.append("")
    lines.append("#endif")
    with open(fname, "w") as f:
        f.write("\n".join(lines) + "\n")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out-oracle", help="CSE point body for the CPU oracle")
    ap.add_argument("--out-io", help="X-macro header of point inputs/outputs")
    ap.add_argument("--out-physcon-io", help="X-macro header of the constraint point inputs/outputs")
    args = ap.parse_args(argv)
    if args.out_io:
        emit_io_header(args.out_io)
    if args.out_physcon_io:
        emit_physcon_io(args.out_physcon_io)
    if not args.out_oracle:
        return
    s = BSSNSystem().build()
    if args.out_oracle:
        emit_cse(s, args.out_oracle,
                 "BSSN point equations (constant eta, standard gauge), CSE form")


if __name__ == "__main__":
    main()
