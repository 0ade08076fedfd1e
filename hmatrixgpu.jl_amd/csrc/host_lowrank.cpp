// host_lowrank.cpp -- ACA+ with partial pivoting and QR+SVD recompression,
// FP64 on the host (compiled with -ffp-contract=off).
//
//   aca_plus        ACA_plus, aca.jl:36-106 (argmax_not_in_list aca.jl:9-17)
//   svd_recompress  SVD_recompress, aca.jl:123-143
//
// The QR is Householder (reflector H = I - tau v v^T with v(0) = 1) and the
// SVD of the small k x k core R_U R_V^T is one-sided cyclic Jacobi.  LAPACK is
// not used: the same arithmetic sequence is reproduced deterministically on any
// host, so ranks and factors are identical from run to run and thread count.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "hmx_host.hpp"

namespace hmx {

namespace {

inline bool jisless(double a, double b) {
    if (std::isnan(a)) return false;
    if (std::isnan(b)) return true;
    if (a < b) return true;
    return a == b && std::signbit(a) && !std::signbit(b);
}

// First entry of the stable descending sortperm not yet used: the maximum under
// isless, the lowest index on ties (aca.jl:9-17).
int64_t pick_pivot(const double* absval, int64_t n, const std::vector<char>& used) {
    int64_t best = -1;
    for (int64_t i = 0; i < n; i++) {
        if (used[i]) continue;
        if (best < 0 || jisless(absval[best], absval[i])) best = i;
    }
    return best;
}

}  // namespace

// Dot product in the fixed order shared with the device recompression: 64
// lane partials (lane l sums x[l]*y[l], x[l+64]*y[l+64], ... in index order),
// then p[l] += p[l+off] for off = 32, 16, ..., 1.
double dot_lanes64(const double* x, const double* y, int64_t n) {
    double p[64] = {};
    for (int64_t i = 0; i < n; i++) {
        const double pr = x[i] * y[i];
        p[i % 64] = p[i % 64] + pr;
    }
    for (int off = 32; off >= 1; off >>= 1)
        for (int l = 0; l < off; l++) p[l] = p[l] + p[l + off];
    return p[0];
}

// Sum of squares in the fixed order shared with the device ACA+ kernel: 256
// lane partials (lane l sums x[l], x[l+256], ... in index order), then a binary
// tree p[l] += p[l+off] for off = 128, 64, ..., 1.  (Julia's sum in aca.jl:95 is
// blocked pairwise with SIMD reassociation, machine dependent; this order is
// this library's definition of the stop test's rounding.)
double sumsq_lanes(const double* x, int64_t n) {
    double p[kSumLanes] = {};
    for (int64_t i = 0; i < n; i++) {
        const double sq = x[i] * x[i];
        p[i % kSumLanes] = p[i % kSumLanes] + sq;
    }
    for (int off = kSumLanes / 2; off >= 1; off >>= 1)
        for (int l = 0; l < off; l++) p[l] = p[l] + p[l + off];
    return p[0];
}

namespace {

// Householder QR of an m x k column-major panel (m >= k), in place.  The
// reflector sums use the fixed 64-lane order of dot_lanes64.
void householder_qr(int64_t m, int64_t k, double* A, double* tau) {
    for (int64_t j = 0; j < k; j++) {
        double* a = A + j * m;
        const double alpha = a[j];
        const double xs = dot_lanes64(a + j + 1, a + j + 1, m - j - 1);
        if (xs == 0.0) {
            tau[j] = 0.0;
            continue;
        }
        const double nrm = std::sqrt(alpha * alpha + xs);
        const double beta = alpha >= 0.0 ? -nrm : nrm;
        tau[j] = (beta - alpha) / beta;
        const double scale = 1.0 / (alpha - beta);
        for (int64_t i = j + 1; i < m; i++) a[i] = a[i] * scale;
        a[j] = beta;
        for (int64_t c = j + 1; c < k; c++) {
            double* b = A + c * m;
            double w = b[j] + dot_lanes64(a + j + 1, b + j + 1, m - j - 1);
            w = w * tau[j];
            b[j] = b[j] - w;
            for (int64_t i = j + 1; i < m; i++) b[i] = b[i] - a[i] * w;
        }
    }
}

// B <- Q B, Q = H_0 ... H_{k-1}
void apply_q(int64_t m, int64_t k, const double* A, const double* tau, int64_t ncols, double* B) {
    for (int64_t j = k - 1; j >= 0; j--) {
        if (tau[j] == 0.0) continue;
        const double* a = A + j * m;
        for (int64_t c = 0; c < ncols; c++) {
            double* b = B + c * m;
            double w = b[j] + dot_lanes64(a + j + 1, b + j + 1, m - j - 1);
            w = w * tau[j];
            b[j] = b[j] - w;
            for (int64_t i = j + 1; i < m; i++) b[i] = b[i] - a[i] * w;
        }
    }
}

// One-sided Jacobi SVD: M (k x k, column-major) = W diag(sig) Z^T.  On return
// M holds W; sig is descending (stable order).  Round-robin sweeps.
void jacobi_svd(int64_t k, double* M, double* sig, double* Z) {
    const double tol = 2.220446049250313e-16;
    std::fill(Z, Z + k * k, 0.0);
    for (int64_t i = 0; i < k; i++) Z[i + i * k] = 1.0;
    // Round-robin (circle) schedule: k' = k rounded up to even players, player 0
    // fixed, 1 .. k'-1 rotated by the round; the k'/2 pairs of a round touch
    // disjoint columns, so their order inside the round does not change a bit
    // (the device rotates them in parallel).
    const int64_t kp = k + (k & 1);
    for (int sweep = 0; sweep < 80; sweep++) {
        bool rotated = false;
        for (int64_t rnd = 0; rnd + 1 < kp; rnd++) {
            for (int64_t pi = 0; pi < kp / 2; pi++) {
                const int64_t a = pi == 0 ? 0 : 1 + (pi - 1 + rnd) % (kp - 1);
                const int64_t b = 1 + (kp - 2 - pi + rnd) % (kp - 1);
                const int64_t p = std::min(a, b), q = std::max(a, b);
                if (q >= k) continue;  // the dummy column of an odd k
                double* ap = M + p * k;
                double* aq = M + q * k;
                double al = 0.0, be = 0.0, ga = 0.0;
                for (int64_t i = 0; i < k; i++) {
                    al += ap[i] * ap[i];
                    be += aq[i] * aq[i];
                    ga += ap[i] * aq[i];
                }
                if (ga == 0.0 || std::fabs(ga) <= tol * std::sqrt(al * be)) continue;
                rotated = true;
                const double zeta = (be - al) / (2.0 * ga);
                const double t =
                    (zeta >= 0.0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
                const double c = 1.0 / std::sqrt(1.0 + t * t);
                const double s = c * t;
                for (int64_t i = 0; i < k; i++) {
                    const double x = ap[i], y = aq[i];
                    ap[i] = c * x - s * y;
                    aq[i] = s * x + c * y;
                }
                double* zp = Z + p * k;
                double* zq = Z + q * k;
                for (int64_t i = 0; i < k; i++) {
                    const double x = zp[i], y = zq[i];
                    zp[i] = c * x - s * y;
                    zq[i] = s * x + c * y;
                }
            }
        }
        if (!rotated) break;
    }
    std::vector<double> nrm(k);
    std::vector<int64_t> ord(k);
    for (int64_t j = 0; j < k; j++) {
        double s = 0.0;
        for (int64_t i = 0; i < k; i++) s += M[i + j * k] * M[i + j * k];
        nrm[j] = std::sqrt(s);
        ord[j] = j;
    }
    std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) { return nrm[a] > nrm[b]; });
    std::vector<double> W(k * k), Z2(k * k);
    for (int64_t j = 0; j < k; j++) {
        const int64_t o = ord[j];
        sig[j] = nrm[o];
        for (int64_t i = 0; i < k; i++) {
            W[i + j * k] = nrm[o] > 0.0 ? M[i + o * k] / nrm[o] : 0.0;
            Z2[i + j * k] = Z[i + o * k];
        }
    }
    std::memcpy(M, W.data(), sizeof(double) * k * k);
    std::memcpy(Z, Z2.data(), sizeof(double) * k * k);
}

}  // namespace

LowRank aca_plus(int64_t m, int64_t n, const RowFn& row, const RowFn& col, double eps,
                 int64_t max_iter) {
    if (max_iter <= 0) max_iter = m;
    const int64_t kmax = std::min(std::min(m, n), max_iter);  // aca.jl:60
    std::vector<double> left, right;  // left[q*m + i], right[q*n + j]
    std::vector<char> used_r(m, 0), used_c(n, 0);
    std::vector<double> rrow(n, 0.0), rcol(m, 0.0), arow(n), acol(m);
    int64_t k = 0;

    auto residual_row = [&](int64_t I) {  // calc_residual_rows, aca.jl:43-49
        row(I, rrow.data());
        for (int64_t q = 0; q < k; q++) {
            const double uq = left[q * m + I];
            const double* vq = right.data() + q * n;
            for (int64_t j = 0; j < n; j++) rrow[j] = rrow[j] - uq * vq[j];
        }
    };
    auto residual_col = [&](int64_t J) {  // calc_residual_cols, aca.jl:52-58
        col(J, rcol.data());
        for (int64_t q = 0; q < k; q++) {
            const double vq = right[q * n + J];
            const double* uq = left.data() + q * m;
            for (int64_t i = 0; i < m; i++) rcol[i] = rcol[i] - uq[i] * vq;
        }
    };

    for (int64_t it = 0; it < kmax; it++) {
        for (int64_t j = 0; j < n; j++) arow[j] = std::fabs(rrow[j]);
        for (int64_t i = 0; i < m; i++) acol[i] = std::fabs(rcol[i]);
        int64_t rp = 0, cp = 0;  // first step: pivot (1,1), aca.jl:69-70
        if (it > 0) {
            cp = pick_pivot(arow.data(), n, used_c);
            rp = pick_pivot(acol.data(), m, used_r);
        }
        const double cval = arow[cp], rval = acol[rp];
        if (rval > cval) {  // aca.jl:79-87
            residual_row(rp);
            for (int64_t j = 0; j < n; j++) arow[j] = std::fabs(rrow[j]);
            cp = pick_pivot(arow.data(), n, used_c);
            residual_col(cp);
        } else {
            residual_col(cp);
            for (int64_t i = 0; i < m; i++) acol[i] = std::fabs(rcol[i]);
            rp = pick_pivot(acol.data(), m, used_r);
            residual_row(rp);
        }
        used_r[rp] = 1;
        used_c[cp] = 1;
        const double piv = rrow[cp];
        right.resize((k + 1) * n);
        left.resize((k + 1) * m);
        double* v = right.data() + k * n;
        double* u = left.data() + k * m;
        for (int64_t j = 0; j < n; j++) v[j] = rrow[j] / piv;  // aca.jl:92
        std::memcpy(u, rcol.data(), sizeof(double) * m);       // aca.jl:93
        k++;
        const double su = sumsq_lanes(u, m), sv = sumsq_lanes(v, n);  // aca.jl:95
        if (std::sqrt(su * sv) < eps) break;                           // aca.jl:97-99
    }
    LowRank out;
    out.m = m;
    out.n = n;
    out.k = k;
    out.U.assign(left.begin(), left.begin() + m * k);  // hcat(left...) : m x k
    out.V.resize(k * n);                               // hcat(right...)' : k x n
    for (int64_t q = 0; q < k; q++)
        for (int64_t j = 0; j < n; j++) out.V[q + j * k] = right[q * n + j];
    return out;
}

LowRank svd_recompress(const LowRank& in, double eps) {
    const int64_t m = in.m, n = in.n, k = in.k;
    LowRank out;
    out.m = m;
    out.n = n;
    if (k == 0) return out;
    std::vector<double> QU(in.U.begin(), in.U.begin() + m * k), QV(n * k), tu(k), tv(k);
    for (int64_t q = 0; q < k; q++)
        for (int64_t j = 0; j < n; j++) QV[j + q * n] = in.V[q + j * k];  // V'
    householder_qr(m, k, QU.data(), tu.data());
    householder_qr(n, k, QV.data(), tv.data());
    std::vector<double> M(k * k);
    for (int64_t i = 0; i < k; i++)
        for (int64_t j = 0; j < k; j++) {  // (R_U R_V')[i,j]
            double s = 0.0;
            for (int64_t l = std::max(i, j); l < k; l++) s += QU[i + l * m] * QV[j + l * n];
            M[i + j * k] = s;
        }
    std::vector<double> sig(k), Z(k * k);
    jacobi_svd(k, M.data(), sig.data(), Z.data());
    // reverse(sqrt.(cumsum(reverse(SIG.^2)))) and the first index below eps
    int64_t r = k;
    {
        std::vector<double> tail(k);
        double acc = 0.0;
        for (int64_t i = k - 1; i >= 0; i--) {
            const double s2 = sig[i] * sig[i];
            acc = (i == k - 1) ? s2 : acc + s2;
            tail[i] = std::sqrt(acc);
        }
        for (int64_t i = 0; i < k; i++)
            if (tail[i] < eps) {
                r = i;
                break;
            }
    }
    out.k = r;
    out.U.assign(m * r, 0.0);
    for (int64_t j = 0; j < r; j++)
        for (int64_t i = 0; i < k; i++) out.U[i + j * m] = M[i + j * k] * sig[j];
    apply_q(m, k, QU.data(), tu.data(), r, out.U.data());
    std::vector<double> C(n * r, 0.0);
    for (int64_t j = 0; j < r; j++)
        for (int64_t i = 0; i < k; i++) C[i + j * n] = Z[i + j * k];
    apply_q(n, k, QV.data(), tv.data(), r, C.data());
    out.V.resize(r * n);
    for (int64_t j = 0; j < r; j++)
        for (int64_t c = 0; c < n; c++) out.V[j + c * r] = C[c + j * n];
    return out;
}

}  // namespace hmx
