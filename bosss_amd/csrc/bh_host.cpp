// Host side of libbosss_hip.so: context, error state, basis / quadrature
// tables and the Cartesian grid.  Written from the mathematics of the
// reference's components (cited per function); no reference code is used.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "bh_internal.hpp"

namespace bh {

static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }
const std::string& last_error() { return g_last_error; }

// ------------------------------------------------------------------ basis
// Np = binomial(k+D, D) (full polynomial space of F/Basis.cs)
int basis_count(int D, int k) {
    int n = 1;
    for (int i = 1; i <= D; ++i) n = n * (k + i) / i;
    return n;
}

// Degree-graded mode order of the reference polynomial tables: within degree
// d, ascending x-exponent, then ascending y-exponent (F/Square.cs:197-260,
// F/Cube_polynomials.cs:20-60).
void basis_exponents(int D, int k, std::vector<int>& ex) {
    ex.clear();
    for (int deg = 0; deg <= k; ++deg) {
        if (D == 2) {
            for (int a = 0; a <= deg; ++a) { ex.push_back(a); ex.push_back(deg - a); }
        } else {
            for (int a = 0; a <= deg; ++a)
                for (int b = 0; b <= deg - a; ++b) { ex.push_back(a); ex.push_back(b); ex.push_back(deg - a - b); }
        }
    }
}

// Orthonormal Legendre factors sqrt(n + 1/2) P_n and derivatives (Bonnet recurrence).
static void legendre_n(int nmax, double x, double* p, double* dp) {
    double P[40], dP[40];
    P[0] = 1; dP[0] = 0;
    if (nmax > 0) { P[1] = x; dP[1] = 1; }
    for (int n = 2; n <= nmax; ++n) {
        P[n] = ((2.0 * n - 1.0) * x * P[n - 1] - (n - 1.0) * P[n - 2]) / n;
        dP[n] = n * P[n - 1] + x * dP[n - 1];
    }
    for (int n = 0; n <= nmax; ++n) {
        double s = std::sqrt(n + 0.5);
        p[n] = s * P[n];
        dp[n] = s * dP[n];
    }
}

void basis_eval(int D, int k, const double* xi, double* val, double* grad) {
    std::vector<int> ex;
    basis_exponents(D, k, ex);
    int Np = (int)ex.size() / D;
    double p[3][40], dp[3][40];
    for (int d = 0; d < D; ++d) legendre_n(k, xi[d], p[d], dp[d]);
    for (int n = 0; n < Np; ++n) {
        const int* e = &ex[n * D];
        if (val) {
            double v = 1;
            for (int d = 0; d < D; ++d) v *= p[d][e[d]];
            val[n] = v;
        }
        if (grad)
            for (int d = 0; d < D; ++d) {
                double g = 1;
                for (int dd = 0; dd < D; ++dd) g *= (dd == d) ? dp[dd][e[dd]] : p[dd][e[dd]];
                grad[n * D + d] = g;
            }
    }
}

// Gauss-Legendre rule by the Golub-Welsch eigenvalue formulation solved with
// a symmetric tridiagonal QL iteration (F/LineQuadRules/Line{n}.txt hold the
// same rules; ascending nodes).
void gauss_rule(int n, double* x, double* w) {
    std::vector<long double> d(n, 0.0L), e(n, 0.0L), z0(n, 0.0L);
    for (int i = 1; i < n; ++i) e[i - 1] = i / std::sqrt(4.0L * i * i - 1.0L);
    // eigenvectors: track only first component (QL with implicit shifts)
    std::vector<long double> zv(n * n, 0.0L);
    for (int i = 0; i < n; ++i) zv[i * n + i] = 1.0L;
    for (int l = 0; l < n; ++l) {
        int iter = 0, m;
        do {
            for (m = l; m < n - 1; ++m) {
                long double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
                if (std::fabs(e[m]) <= 1e-30L * dd || std::fabs(e[m]) < 1e-300L) break;
            }
            if (m != l) {
                if (iter++ == 200) break;
                long double g = (d[l + 1] - d[l]) / (2.0L * e[l]);
                long double r = std::hypot(g, 1.0L);
                g = d[m] - d[l] + e[l] / (g + (g >= 0 ? std::fabs(r) : -std::fabs(r)));
                long double s = 1, c = 1, p = 0;
                int i;
                for (i = m - 1; i >= l; --i) {
                    long double f = s * e[i], b = c * e[i];
                    r = std::hypot(f, g);
                    e[i + 1] = r;
                    if (r == 0) { d[i + 1] -= p; e[m] = 0; break; }
                    s = f / r; c = g / r;
                    g = d[i + 1] - p;
                    r = (d[i] - g) * s + 2.0L * c * b;
                    p = s * r;
                    d[i + 1] = g + p;
                    g = c * r - b;
                    for (int kk = 0; kk < n; ++kk) {
                        long double t = zv[kk * n + i + 1];
                        zv[kk * n + i + 1] = s * zv[kk * n + i] + c * t;
                        zv[kk * n + i] = c * zv[kk * n + i] - s * t;
                    }
                }
                if (r == 0 && i >= l) continue;
                d[l] -= p; e[l] = g; e[m] = 0;
            }
        } while (m != l);
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) { return d[a] < d[b]; });
    for (int i = 0; i < n; ++i) {
        int j = order[i];
        x[i] = (double)d[j];
        w[i] = (double)(2.0L * zv[0 * n + j] * zv[0 * n + j]);
    }
    // symmetrise (the exact rule is symmetric about 0)
    for (int i = 0; i < n / 2; ++i) {
        double xs = 0.5 * (x[n - 1 - i] - x[i]), ws = 0.5 * (w[i] + w[n - 1 - i]);
        x[i] = -xs; x[n - 1 - i] = xs; w[i] = ws; w[n - 1 - i] = ws;
    }
    if (n % 2) x[n / 2] = 0.0;
}

// Tables at tensor Gauss nodes (volume: F/Square.cs:110-128 tensor rule;
// faces: same rule in the D-1 tangential coordinates, ascending dimension).
void SipTables::build(int D_, int k_, int nq1_) {
    D = D_; k = k_; nq1 = nq1_;
    Np = basis_count(D, k);
    std::vector<double> x(nq1), w(nq1);
    gauss_rule(nq1, x.data(), w.data());
    nqv = 1; nqf = 1;
    for (int d = 0; d < D; ++d) nqv *= nq1;
    for (int d = 0; d < D - 1; ++d) nqf *= nq1;
    vw.assign(nqv, 1.0); vxi.assign(nqv * D, 0.0);
    G.assign((size_t)nqv * D * Np, 0.0);
    std::vector<double> val(Np), grad(Np * D);
    for (int q = 0; q < nqv; ++q) {
        int r = q;
        for (int d = 0; d < D; ++d) { int id = r % nq1; r /= nq1; vxi[q * D + d] = x[id]; vw[q] *= w[id]; }
        basis_eval(D, k, &vxi[q * D], val.data(), grad.data());
        for (int d = 0; d < D; ++d)
            for (int n = 0; n < Np; ++n) G[((size_t)q * D + d) * Np + n] = grad[n * D + d];
    }
    Kv.assign((size_t)D * Np * Np, 0.0);
    for (int d = 0; d < D; ++d)
        for (int n = 0; n < Np; ++n)
            for (int m = 0; m < Np; ++m) {
                double s = 0;
                for (int q = 0; q < nqv; ++q) s += vw[q] * G[((size_t)q * D + d) * Np + n] * G[((size_t)q * D + d) * Np + m];
                Kv[((size_t)d * Np + n) * Np + m] = s;
            }
    fw.assign(nqf, 1.0); fxi.assign((size_t)nqf * std::max(D - 1, 1), 0.0);
    for (int q = 0; q < nqf; ++q) {
        int r = q;
        for (int d = 0; d < D - 1; ++d) { int id = r % nq1; r /= nq1; fxi[q * (D - 1) + d] = x[id]; fw[q] *= w[id]; }
    }
    T.assign((size_t)2 * D * nqf * Np, 0.0);
    Nn.assign((size_t)2 * D * nqf * Np, 0.0);
    for (int f = 0; f < 2 * D; ++f) {
        int d = f / 2, s = f % 2;
        for (int q = 0; q < nqf; ++q) {
            double xi[3];
            int m = 0;
            for (int dd = 0; dd < D; ++dd) xi[dd] = (dd == d) ? (s ? 1.0 : -1.0) : fxi[q * (D - 1) + m++];
            basis_eval(D, k, xi, val.data(), grad.data());
            for (int n = 0; n < Np; ++n) {
                T[((size_t)f * nqf + q) * Np + n] = val[n];
                Nn[((size_t)f * nqf + q) * Np + n] = grad[n * D + d];
            }
        }
    }
}

}  // namespace bh

// ------------------------------------------------------------------ ctx
void bh_ctx::upload(void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    BH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
    BH_HIP(hipStreamSynchronize(stream));
}
void bh_ctx::download(void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    BH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream));
    BH_HIP(hipStreamSynchronize(stream));
}
void bh_ctx::sync() { BH_HIP(hipStreamSynchronize(stream)); }
