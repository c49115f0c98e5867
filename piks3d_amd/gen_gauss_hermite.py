"""Generate the Gauss-Hermite node/weight tables used by the velocity grid.

The velocity set of the DVM solver is built from 1-D quadratures
(reference: src/velocityGrid.F90:102-155 selects a table, :166-170 scales the
nodes by sqrt(2)).  Two families are used:

* full range: the n positive nodes of the 2n-point Gauss rule for the weight
  exp(-x^2)/sqrt(pi) on the whole real line (weights of the positive half sum
  to 1/2);
* half range: the n-point Gauss rule for the weight exp(-x^2)/sqrt(pi) on
  [0, inf) (weights sum to 1/2).

Both are computed here from first principles with mpmath at 250 digits
(Golub-Welsch on the Jacobi matrix, nodes polished by Newton's method on the
three-term recurrence) and rounded once to IEEE double, so every entry is the
correctly rounded value of the exact quadrature.  The full-range recurrence is
the Hermite one (alpha_k = 0, beta_k = k/2); the half-range recurrence comes
from the modified moments m_k = Gamma((k+1)/2) / (2 sqrt(pi)) through the
Chebyshev algorithm.

The reference ships its own decimal tables for n in {2, 4, 6, 8, 12, 16}
(src/gaussHermite.F90:56-329).  They are not all the exact rules: n = 2, 4, 8,
12 agree with the exact values to about one ulp, the half-range n = 6 and 16
tables differ by 1e-13..3e-12, and the full-range weights for n >= 6 lack the
1/sqrt(pi) normalisation.  Results identical to the reference's need its
doubles, so for those n the header carries the reference's values (parsed
from that file's text when it is available, correctly rounded like the
Fortran compiler does) as the default table, and the exact values next to it.
n = 3 (a 6^3 velocity grid) is not in the reference; only the exact rule is
available for it.

Run:  python -m piks3d_amd.gen_gauss_hermite [REFERENCE_SRC_DIR]
      (rewrites csrc/gauss_hermite_tables.h)
"""
from __future__ import annotations

import os
import sys

import mpmath as mp

NS = (2, 3, 4, 6, 8, 12, 16)
DPS = 250


def _jacobi_from_moments(moments, n):
    """Chebyshev algorithm: recurrence coefficients alpha[0..n-1], beta[0..n-1]
    of the monic orthogonal polynomials from ordinary moments (Gautschi 2004,
    Alg. 2.1).  beta[0] = m_0."""
    sig_prev = [mp.mpf(0)] * (2 * n)
    sig = [moments[l] for l in range(2 * n)]
    alpha = [moments[1] / moments[0]]
    beta = [moments[0]]
    for k in range(1, n):
        sig_next = [mp.mpf(0)] * (2 * n)
        for l in range(k, 2 * n - k):
            sig_next[l] = sig[l + 1] - alpha[k - 1] * sig[l] - beta[k - 1] * sig_prev[l]
        alpha.append(sig_next[k + 1] / sig_next[k] - sig[k] / sig[k - 1])
        beta.append(sig_next[k] / sig[k - 1])
        sig_prev, sig = sig, sig_next
    return alpha, beta


def _gauss(alpha, beta):
    """Golub-Welsch: nodes and weights of the Gauss rule for the recurrence."""
    n = len(alpha)
    J = mp.matrix(n, n)
    for i in range(n):
        J[i, i] = alpha[i]
        if i + 1 < n:
            J[i, i + 1] = J[i + 1, i] = mp.sqrt(beta[i + 1])
    evals, evecs = mp.eigsy(J)
    nodes = []
    weights = []
    for i in range(n):
        x = evals[i]
        # Newton polish on p_n(x) from the recurrence
        for _ in range(6):
            p_prev, p = mp.mpf(0), mp.mpf(1)
            dp_prev, dp = mp.mpf(0), mp.mpf(0)
            for k in range(n):
                bk = beta[k] if k > 0 else mp.mpf(0)
                p_new = (x - alpha[k]) * p - bk * p_prev
                dp_new = p + (x - alpha[k]) * dp - bk * dp_prev
                p_prev, p = p, p_new
                dp_prev, dp = dp, dp_new
            x = x - p / dp
        nodes.append(x)
    # weights from the Christoffel numbers: w_i = 1 / sum_k p_k(x_i)^2 / h_k
    for x in nodes:
        s = mp.mpf(0)
        p_prev, p = mp.mpf(0), mp.mpf(1)
        h = beta[0]
        s += p * p / h
        for k in range(n - 1):
            bk = beta[k] if k > 0 else mp.mpf(0)
            p_new = (x - alpha[k]) * p - bk * p_prev
            p_prev, p = p, p_new
            h = h * beta[k + 1]
            s += p * p / h
        weights.append(1 / s)
    order = sorted(range(n), key=lambda i: nodes[i])
    return [nodes[i] for i in order], [weights[i] for i in order]


def half_range(n):
    mp.mp.dps = DPS
    mom = [mp.gamma(mp.mpf(k + 1) / 2) / (2 * mp.sqrt(mp.pi)) for k in range(2 * n)]
    a, b = _jacobi_from_moments(mom, n)
    return _gauss(a, b)


def full_range(n):
    """Positive half of the 2n-point Hermite rule, weight exp(-x^2)/sqrt(pi)."""
    mp.mp.dps = DPS
    m = 2 * n
    alpha = [mp.mpf(0)] * m
    beta = [mp.mpf(1)] + [mp.mpf(k) / 2 for k in range(1, m)]
    x, w = _gauss(alpha, beta)
    pos = [(xi, wi) for xi, wi in zip(x, w) if xi > 0]
    return [p[0] for p in pos], [p[1] for p in pos]


def _hexd(v):
    return float(v).hex()


REF_NS = (2, 4, 6, 8, 12, 16)


def reference_tables(src_dir):
    """Parse the reference's decimal tables (text) into doubles."""
    import re
    text = open(os.path.join(src_dir, "gaussHermite.F90")).read()
    out = {}
    for fam, prefix in (("half", "h"), ("full", "")):
        for n in REF_NS:
            vals = []
            for kind in ("xi", "wi"):
                m = re.search(r"::\s*%s%s%d\s*=\s*\(/&(.*?)/\)" % (prefix, kind, n), text, re.S)
                nums = re.findall(r"[-+]?\d+\.\d+[dD][-+]\d+", m.group(1))
                vals.append([float(t.replace("d", "e").replace("D", "e")) for t in nums])
            out[(fam, n)] = tuple(vals)
    return out


def render(ref) -> str:
    out = []
    out.append("// Generated by piks3d_amd/gen_gauss_hermite.py -- do not edit.")
    out.append("// 1-D Gauss-Hermite nodes and weights for the weight exp(-x^2)/sqrt(pi).")
    out.append("// piks_gh_{half,full}_{x,w}<n>: the table the solver uses.  For the n the")
    out.append("// reference tabulates (2, 4, 6, 8, 12, 16) these are the reference's doubles")
    out.append("// (src/gaussHermite.F90:56-329), kept identical so results match it bit for")
    out.append("// bit; for n = 3 they are the exact rule.  piks_ghx_*: the exact rules,")
    out.append("// correctly rounded (method in the generator).")
    out.append("#pragma once")
    out.append("")
    for fam, fn in (("half", half_range), ("full", full_range)):
        for n in NS:
            x, w = fn(n)
            ex = [float(v) for v in x]
            ew = [float(v) for v in w]
            rx, rw = ref.get((fam, n), (ex, ew))
            out.append("static const double piks_gh_%s_x%d[%d] = {%s};" % (fam, n, n, ", ".join(_hexd(v) for v in rx)))
            out.append("static const double piks_gh_%s_w%d[%d] = {%s};" % (fam, n, n, ", ".join(_hexd(v) for v in rw)))
            out.append("static const double piks_ghx_%s_x%d[%d] = {%s};" % (fam, n, n, ", ".join(_hexd(v) for v in ex)))
            out.append("static const double piks_ghx_%s_w%d[%d] = {%s};" % (fam, n, n, ", ".join(_hexd(v) for v in ew)))
    out.append("")
    out.append("// Returns 0 and sets *x, *w for a supported n, -1 otherwise.  exact != 0")
    out.append("// selects the exact rule instead of the reference-compatible one.")
    out.append("static inline int piks_gh_table(int half, int n, int exact, const double** x, const double** w) {")
    out.append("  switch (n) {")
    for n in NS:
        out.append("    case %d:" % n)
        out.append("      if (exact) { *x = half ? piks_ghx_half_x%d : piks_ghx_full_x%d; *w = half ? piks_ghx_half_w%d : piks_ghx_full_w%d; }" % (n, n, n, n))
        out.append("      else { *x = half ? piks_gh_half_x%d : piks_gh_full_x%d; *w = half ? piks_gh_half_w%d : piks_gh_full_w%d; }" % (n, n, n, n))
        out.append("      return 0;")
    out.append("    default: return -1;")
    out.append("  }")
    out.append("}")
    out.append("")
    out.append("// n values the reference itself supports (src/velocityGrid.F90:103-154).")
    out.append("static inline int piks_gh_in_reference(int n) {")
    out.append("  return n == 2 || n == 4 || n == 6 || n == 8 || n == 12 || n == 16;")
    out.append("}")
    return "\n".join(out) + "\n"


def main(argv=None):
    here = os.path.dirname(os.path.abspath(__file__))
    path = os.path.join(here, "csrc", "gauss_hermite_tables.h")
    src = argv[0] if argv else "/root/reference/src"
    if not os.path.exists(os.path.join(src, "gaussHermite.F90")):
        raise SystemExit("reference tables not found under %s; the committed header "
                         "already holds them" % src)
    text = render(reference_tables(src))
    with open(path, "w") as fh:
        fh.write(text)
    print("wrote", path)


if __name__ == "__main__":
    main(sys.argv[1:])
