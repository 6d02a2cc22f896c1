// Gauss quadrature rules for the partial-coherence integrals of the lenses.
//
// Nodes and weights come from the three-term recurrence of the weight's
// monic orthogonal polynomials (Golub and Welsch, Math. Comp. 23 (1969)
// 221-230): the nodes are the eigenvalues of the symmetric tridiagonal
// Jacobi matrix, the weights mu0 * (first component of the unit
// eigenvector)^2.  The reference obtains the same rules from its cgqf
// routine (Quadrature kinds 25 and 26, src/quadrature.hpp:32-126).
#pragma once

#include <vector>

namespace mtg {

struct GaussRule {
    std::vector<double> x, w;   // nodes ascending
};

// int_{-inf}^{inf} |x - a|^alpha exp(-b (x - a)^2) f(x) dx ~ sum w_i f(x_i)
GaussRule gauss_gen_hermite(int n, double alpha, double a, double b);

// int_a^{inf} (x - a)^alpha exp(-b (x - a)) f(x) dx ~ sum w_i f(x_i)
GaussRule gauss_gen_laguerre(int n, double alpha, double a, double b);

} // namespace mtg
