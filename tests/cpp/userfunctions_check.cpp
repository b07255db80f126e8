// Check of include/zoomy/UserFunctions.H (eigenvalues, eigensystem, solve) for
// tests/test_userfunctions.py.  Reads cases from stdin, one per line:
//   E n a00 a01 ... (n*n values)       -> "E" then the eigenvalues, R (n*n), L (n*n)
//   S n a00 ... (n*n values) b0 ... bn  -> "S" then the solution x
// and calls the header's variadic kernels exactly as a generated Numerics.H does (one call per
// output index, the matrix entries as separate arguments).
#include <cstdio>
#include <iostream>
#include <string>
#include <utility>
#include <vector>

#include "zoomy/UserFunctions.H"

template <std::size_t... K>
double eig_at(int idx, const double* a, std::index_sequence<K...>) {
    return eigenvalues(idx, a[K]...);
}
template <std::size_t... K>
double sys_at(int idx, const double* a, std::index_sequence<K...>) {
    return eigensystem(idx, a[K]...);
}
template <std::size_t... K>
double solve_at(int idx, const double* a, std::index_sequence<K...>) {
    return solve(idx, a[K]...);
}

template <int N>
void eig_case(const std::vector<double>& a) {
    std::printf("E");
    for (int i = 0; i < N; ++i) std::printf(" %.17g", eig_at(i, a.data(), std::make_index_sequence<N * N>{}));
    for (int i = 0; i < N + 2 * N * N; ++i)
        std::printf(" %.17g", sys_at(i, a.data(), std::make_index_sequence<N * N>{}));
    std::printf("\n");
}

template <int N>
void solve_case(const std::vector<double>& a) {
    std::printf("S");
    for (int i = 0; i < N; ++i) std::printf(" %.17g", solve_at(i, a.data(), std::make_index_sequence<N * N + N>{}));
    std::printf("\n");
}

template <int N>
void dispatch(char kind, const std::vector<double>& a) {
    if (kind == 'E')
        eig_case<N>(a);
    else
        solve_case<N>(a);
}

int main() {
    std::string kind;
    int n;
    while (std::cin >> kind >> n) {
        const int m = kind == "E" ? n * n : n * n + n;
        std::vector<double> a(m);
        for (double& v : a) {
            std::string tok;
            std::cin >> tok;
            v = std::stod(tok);
        }
        switch (n) {
            case 1: dispatch<1>(kind[0], a); break;
            case 2: dispatch<2>(kind[0], a); break;
            case 3: dispatch<3>(kind[0], a); break;
            case 4: dispatch<4>(kind[0], a); break;
            case 6: dispatch<6>(kind[0], a); break;
            case 8: dispatch<8>(kind[0], a); break;
            default: std::printf("X unsupported size\n");
        }
        std::fflush(stdout);
    }
    return 0;
}
