// Generated by gen_sme_coeffs.py -- exact SME moment tables (do not edit).
// Index 0 is the mean (phi_0 = 1); entries are rationals written as p.0/q.0.
#pragma once

template <int L> struct SmeCoeffs;

template <> struct SmeCoeffs<0> {
    static constexpr int N = 1;
    // A[(i*N + j)*N + k]
    static constexpr double A[1] = {1.0};
    // B[(i*N + j)*N + k]
    static constexpr double B[1] = {0.0};
    // C[i*N + j]
    static constexpr double C[1] = {0.0};
};

template <> struct SmeCoeffs<1> {
    static constexpr int N = 2;
    // A[(i*N + j)*N + k]
    static constexpr double A[8] = {1.0, 0.0, 0.0, 1.0/3.0, 0.0, 1.0, 1.0, 0.0};
    // B[(i*N + j)*N + k]
    static constexpr double B[8] = {0.0, 0.0, 0.0, 0.0, -3.0, 1.0, -1.0, 0.0};
    // C[i*N + j]
    static constexpr double C[4] = {0.0, 0.0, 0.0, 4.0};
};

template <> struct SmeCoeffs<2> {
    static constexpr int N = 3;
    // A[(i*N + j)*N + k]
    static constexpr double A[27] = {1.0, 0.0, 0.0, 0.0, 1.0/3.0, 0.0, 0.0, 0.0, 1.0/5.0, 0.0, 1.0, 0.0, 1.0, 0.0, 2.0/5.0, 0.0, 2.0/5.0, 0.0, 0.0, 0.0, 1.0, 0.0, 2.0/3.0, 0.0, 1.0, 0.0, 2.0/7.0};
    // B[(i*N + j)*N + k]
    static constexpr double B[27] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -3.0, 1.0, 0.0, -1.0, 0.0, 1.0/5.0, 0.0, -1.0/5.0, 0.0, 5.0, -5.0, 2.0, 0.0, -1.0, 0.0, -1.0, 0.0, -1.0/7.0};
    // C[i*N + j]
    static constexpr double C[9] = {0.0, 0.0, 0.0, 0.0, 4.0, 0.0, 0.0, 0.0, 12.0};
};

template <> struct SmeCoeffs<3> {
    static constexpr int N = 4;
    // A[(i*N + j)*N + k]
    static constexpr double A[64] = {1.0, 0.0, 0.0, 0.0, 0.0, 1.0/3.0, 0.0, 0.0, 0.0, 0.0, 1.0/5.0, 0.0, 0.0, 0.0, 0.0, 1.0/7.0, 0.0, 1.0, 0.0, 0.0, 1.0, 0.0, 2.0/5.0, 0.0, 0.0, 2.0/5.0, 0.0, 9.0/35.0, 0.0, 0.0, 9.0/35.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 2.0/3.0, 0.0, 3.0/7.0, 1.0, 0.0, 2.0/7.0, 0.0, 0.0, 3.0/7.0, 0.0, 4.0/21.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 3.0/5.0, 0.0, 0.0, 3.0/5.0, 0.0, 4.0/15.0, 1.0, 0.0, 4.0/15.0, 0.0};
    // B[(i*N + j)*N + k]
    static constexpr double B[64] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, -3.0, 1.0, 0.0, 0.0, -1.0, 0.0, 1.0/5.0, 0.0, 0.0, -1.0/5.0, 0.0, 3.0/35.0, 0.0, 0.0, -3.0/35.0, 0.0, 5.0, -5.0, 2.0, 0.0, 0.0, -1.0, 0.0, 3.0/7.0, -1.0, 0.0, -1.0/7.0, 0.0, 0.0, -2.0/7.0, 0.0, -1.0/21.0, -7.0, 7.0, -7.0, 3.0, 0.0, 0.0, -6.0/5.0, 0.0, 0.0, -4.0/5.0, 0.0, -2.0/15.0, -1.0, 0.0, -1.0/5.0, 0.0};
    // C[i*N + j]
    static constexpr double C[16] = {0.0, 0.0, 0.0, 0.0, 0.0, 4.0, 0.0, 4.0, 0.0, 0.0, 12.0, 0.0, 0.0, 4.0, 0.0, 24.0};
};
