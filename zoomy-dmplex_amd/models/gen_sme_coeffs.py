"""Generate the exact moment-coupling tables of the shallow-water moment equations (SME).

The reference builds its model headers with ``create_model.py`` from the external
``zoomy_core`` package (create_model.py:32-37, 80-101), which is not available.  This script
derives the three coefficient tables the SME of Kowalski & Torrilhon (2019, "Moment
approximations and model cascades for shallow flow") need, exactly, with sympy, and writes
them to ``sme_coeffs.H`` as constexpr tables for levels 0..LMAX.

Basis: phi_j(z) = P_j(1 - 2 z) on z in [0, 1] (P_j Legendre), so phi_j(0) = 1 (bottom) and
int_0^1 phi_i phi_j dz = delta_ij / (2 j + 1).

  A[i][j][k] = (2i+1) int_0^1 phi_i phi_j phi_k dz                (advection of moments)
  B[i][j][k] = (2i+1) int_0^1 phi_i' (int_0^z phi_j) phi_k dz     (vertical coupling)
  C[i][j]    =        int_0^1 phi_i' phi_j' dz                    (internal viscous stress)

Run:  python gen_sme_coeffs.py  (rewrites sme_coeffs.H next to this file)
"""
import os
import sympy as sp

LMAX = 3


def tables(level):
    z = sp.symbols("z")
    phi = [sp.expand(sp.legendre(j, 1 - 2 * z)) for j in range(level + 1)]
    dphi = [sp.diff(p, z) for p in phi]
    iphi = [sp.integrate(p, (z, 0, z)) for p in phi]
    n = level + 1
    A = [[[sp.integrate((2 * i + 1) * phi[i] * phi[j] * phi[k], (z, 0, 1)) for k in range(n)]
          for j in range(n)] for i in range(n)]
    B = [[[sp.integrate((2 * i + 1) * dphi[i] * iphi[j] * phi[k], (z, 0, 1)) for k in range(n)]
          for j in range(n)] for i in range(n)]
    C = [[sp.integrate(dphi[i] * dphi[j], (z, 0, 1)) for j in range(n)] for i in range(n)]
    return A, B, C


def fmt(x):
    x = sp.Rational(x)
    if x.q == 1:
        return "%d.0" % x.p
    return "%d.0/%d.0" % (x.p, x.q)


def main():
    out = []
    out.append("// Generated by gen_sme_coeffs.py -- exact SME moment tables (do not edit).")
    out.append("// Index 0 is the mean (phi_0 = 1); entries are rationals written as p.0/q.0.")
    out.append("#pragma once")
    out.append("")
    out.append("template <int L> struct SmeCoeffs;")
    for level in range(LMAX + 1):
        A, B, C = tables(level)
        n = level + 1
        out.append("")
        out.append("template <> struct SmeCoeffs<%d> {" % level)
        out.append("    static constexpr int N = %d;" % n)
        a = ", ".join(fmt(A[i][j][k]) for i in range(n) for j in range(n) for k in range(n))
        b = ", ".join(fmt(B[i][j][k]) for i in range(n) for j in range(n) for k in range(n))
        c = ", ".join(fmt(C[i][j]) for i in range(n) for j in range(n))
        out.append("    // A[(i*N + j)*N + k]")
        out.append("    static constexpr double A[%d] = {%s};" % (n ** 3, a))
        out.append("    // B[(i*N + j)*N + k]")
        out.append("    static constexpr double B[%d] = {%s};" % (n ** 3, b))
        out.append("    // C[i*N + j]")
        out.append("    static constexpr double C[%d] = {%s};" % (n ** 2, c))
        out.append("};")
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sme_coeffs.H")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")
    print("wrote", path)


if __name__ == "__main__":
    main()
