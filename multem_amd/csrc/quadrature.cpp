// Gauss rules from the Jacobi matrix, solved by cyclic Jacobi rotations.
#include "quadrature.hpp"

#include <algorithm>
#include <cmath>

namespace mtg {

namespace {

// Eigenvalues of the symmetric tridiagonal matrix (diag d, off-diag e) by
// cyclic Jacobi rotations on the dense matrix.  n is small (the number of
// integration points), so O(n^3) per sweep is immaterial.
std::vector<double> jacobi_eigenvalues(const std::vector<double> &d, const std::vector<double> &e)
{
    const int n = (int)d.size();
    std::vector<double> A((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) {
        A[(size_t)i * n + i] = d[i];
        if (i + 1 < n) A[(size_t)i * n + i + 1] = A[(size_t)(i + 1) * n + i] = e[i];
    }
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; ++i) {
            diag += A[(size_t)i * n + i] * A[(size_t)i * n + i];
            for (int j = i + 1; j < n; ++j) off += A[(size_t)i * n + j] * A[(size_t)i * n + j];
        }
        if (off <= 1e-34 * diag || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[(size_t)p * n + q];
                if (apq == 0.0) continue;
                const double app = A[(size_t)p * n + p], aqq = A[(size_t)q * n + q];
                // rotation angle that zeroes A[p][q]: t = tan(theta), the smaller root
                const double th = (aqq - app) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    const double akp = A[(size_t)k * n + p], akq = A[(size_t)k * n + q];
                    A[(size_t)k * n + p] = c * akp - s * akq;
                    A[(size_t)k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    const double apk = A[(size_t)p * n + k], aqk = A[(size_t)q * n + k];
                    A[(size_t)p * n + k] = c * apk - s * aqk;
                    A[(size_t)q * n + k] = s * apk + c * aqk;
                }
                A[(size_t)p * n + q] = A[(size_t)q * n + p] = 0.0;
            }
    }
    std::vector<double> lam(n);
    for (int i = 0; i < n; ++i) lam[i] = A[(size_t)i * n + i];
    return lam;
}

// Nodes from the eigenvalues; weights from the Christoffel numbers
// 1 / sum_k p_k(x)^2 of the orthonormal polynomials p_k (recurrence
// b_{k+1} p_{k+1} = (x - a_k) p_k - b_k p_{k-1}, p_0 = 1/sqrt(mu0)), which
// keeps full relative accuracy for the tiny weights of the outer nodes (the
// squared eigenvector components do not).
GaussRule golub_welsch(const std::vector<double> &d, const std::vector<double> &e, double mu0)
{
    std::vector<double> lam = jacobi_eigenvalues(d, e);
    const int n = (int)d.size();
    std::sort(lam.begin(), lam.end());
    GaussRule r;
    for (int i = 0; i < n; ++i) {
        const double x = lam[i];
        double pm = 0.0, p = 1.0 / std::sqrt(mu0), s = p * p;
        for (int k = 0; k + 1 < n; ++k) {
            const double pn = ((x - d[k]) * p - (k > 0 ? e[k - 1] * pm : 0.0)) / e[k];
            pm = p;
            p = pn;
            s += p * p;
        }
        r.x.push_back(x);
        r.w.push_back(1.0 / s);
    }
    return r;
}

} // namespace

GaussRule gauss_gen_hermite(int n, double alpha, double a, double b)
{
    // weight |t|^alpha exp(-t^2): a_k = 0, b_k = (k + alpha [k odd]) / 2,
    // mu0 = Gamma((alpha + 1) / 2); then t = sqrt(b) (x - a)
    std::vector<double> d(n, 0.0), e(n > 0 ? n - 1 : 0);
    for (int k = 1; k < n; ++k) e[k - 1] = std::sqrt((k + ((k & 1) ? alpha : 0.0)) / 2.0);
    GaussRule r = golub_welsch(d, e, std::tgamma((alpha + 1.0) / 2.0));
    const double sl = 1.0 / std::sqrt(b), p = std::pow(sl, alpha + 1.0);
    for (int i = 0; i < n; ++i) {
        r.x[i] = a + sl * r.x[i];
        r.w[i] *= p;
    }
    return r;
}

GaussRule gauss_gen_laguerre(int n, double alpha, double a, double b)
{
    // weight t^alpha exp(-t): a_k = 2k - 1 + alpha, b_k = k (k + alpha),
    // mu0 = Gamma(alpha + 1); then t = b (x - a)
    std::vector<double> d(n), e(n > 0 ? n - 1 : 0);
    for (int k = 1; k <= n; ++k) d[k - 1] = 2.0 * k - 1.0 + alpha;
    for (int k = 1; k < n; ++k) e[k - 1] = std::sqrt(k * (k + alpha));
    GaussRule r = golub_welsch(d, e, std::tgamma(alpha + 1.0));
    const double sl = 1.0 / b, p = std::pow(sl, alpha + 1.0);
    for (int i = 0; i < n; ++i) {
        r.x[i] = a + sl * r.x[i];
        r.w[i] *= p;
    }
    return r;
}

} // namespace mtg
