// Host-side symmetric tridiagonal eigensolver of the duct's z-wall operator
// (plain C++, no HIP: also compiled on its own by the CPU tests).
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace dfp {

// Eigenvalues and orthonormal eigenvectors of the symmetric tridiagonal
// matrix with diagonal d[0..n) and off-diagonal e[i] (coupling rows i and
// i+1) by implicit QL sweeps with Wilkinson shifts (Golub & Van Loan,
// Matrix Computations sec. 8.3, in the QL orientation of EISPACK's tql2):
// on return d holds the eigenvalues in ascending order and column m of v
// (n x n, column-major) the eigenvector of d[m].  false if an eigenvalue
// takes more than 60 sweeps.
inline bool symtri_eigen(int n, std::vector<double>& d, std::vector<double> e, std::vector<double>& v) {
    v.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) v[(size_t)i * n + i] = 1.0;
    e.resize(n, 0.0);
    e[n - 1] = 0.0;
    const double eps = std::numeric_limits<double>::epsilon();
    for (int l = 0; l < n; ++l) {
        for (int sweep = 0;; ++sweep) {
            // the first negligible off-diagonal at or below row l closes the block
            int m = l;
            while (m < n - 1 && std::fabs(e[m]) > eps * (std::fabs(d[m]) + std::fabs(d[m + 1]))) ++m;
            if (m == l) break;
            if (sweep == 60) return false;
            // shift: the eigenvalue of the leading 2x2 block nearer d[l]
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
            double s = 1.0, c = 1.0, p = 0.0;
            bool split = false;
            // chase the bulge from the bottom of the block up to row l
            for (int i = m - 1; i >= l; --i) {
                const double f = s * e[i], b = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {  // the block splits: restart on the smaller one
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    split = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                double* vi = v.data() + (size_t)i * n;
                double* vj = v.data() + (size_t)(i + 1) * n;
                for (int k = 0; k < n; ++k) {
                    const double a1 = vi[k], a2 = vj[k];
                    vj[k] = s * a1 + c * a2;
                    vi[k] = c * a1 - s * a2;
                }
            }
            if (split) continue;
            d[l] -= p;
            e[l] = g;
            e[m] = 0.0;
        }
    }
    // ascending order, eigenvectors along
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        for (int j = i + 1; j < n; ++j)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            std::swap(d[i], d[k]);
            for (int q = 0; q < n; ++q) std::swap(v[(size_t)i * n + q], v[(size_t)k * n + q]);
        }
    }
    return true;
}

}  // namespace dfp
