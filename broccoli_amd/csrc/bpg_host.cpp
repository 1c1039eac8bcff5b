// bpg_host.cpp -- host-side pieces of the path: FP64 model construction,
// permutation / sign generators, design detection and the CLI's model
// precompute.  Pure C++ (no device code).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unordered_set>
#include <vector>

#include "bpg_internal.h"

namespace bpg {

// Householder QR of a column-major m x n matrix A (m >= n).  Q (m x m,
// column-major, full) and R (m x n) may be null.
int householder_qr(int m, int n, const double* A, double* Q, double* R) {
    std::vector<double> a(A, A + (size_t)m * n), q((size_t)m * m, 0.0), v(m);
    for (int i = 0; i < m; i++) q[i + (size_t)i * m] = 1.0;
    const int steps = std::min(n, m - 1);
    for (int k = 0; k < steps; k++) {
        double nrm = 0.0;
        for (int i = k; i < m; i++) nrm += a[i + (size_t)k * m] * a[i + (size_t)k * m];
        nrm = std::sqrt(nrm);
        if (nrm == 0.0) continue;
        const double alpha = a[k + (size_t)k * m] > 0 ? -nrm : nrm;
        std::fill(v.begin(), v.end(), 0.0);
        for (int i = k; i < m; i++) v[i] = a[i + (size_t)k * m];
        v[k] -= alpha;
        double vv = 0.0;
        for (int i = k; i < m; i++) vv += v[i] * v[i];
        if (vv == 0.0) continue;
        for (int j = k; j < n; j++) {
            double d = 0.0;
            for (int i = k; i < m; i++) d += v[i] * a[i + (size_t)j * m];
            d = 2.0 * d / vv;
            for (int i = k; i < m; i++) a[i + (size_t)j * m] -= d * v[i];
        }
        for (int i = 0; i < m; i++) {
            double d = 0.0;
            for (int l = k; l < m; l++) d += q[i + (size_t)l * m] * v[l];
            d = 2.0 * d / vv;
            for (int l = k; l < m; l++) q[i + (size_t)l * m] -= d * v[l];
        }
    }
    if (Q) std::copy(q.begin(), q.end(), Q);
    if (R)
        for (int j = 0; j < n; j++)
            for (int i = 0; i < m; i++) R[i + (size_t)j * m] = (i <= j) ? a[i + (size_t)j * m] : 0.0;
    return 0;
}

namespace {

// Thin QR of X (N x p): Q1 (N x p) and upper-triangular R1 (p x p).
// Returns false when X is rank deficient.
bool thin_qr(int N, int p, const std::vector<double>& X, std::vector<double>& Q1,
             std::vector<double>& R1) {
    std::vector<double> Q((size_t)N * N), R((size_t)N * p);
    householder_qr(N, p, X.data(), Q.data(), R.data());
    Q1.assign((size_t)N * p, 0.0);
    R1.assign((size_t)p * p, 0.0);
    double dmax = 0.0;
    for (int j = 0; j < p; j++) dmax = std::max(dmax, std::fabs(R[j + (size_t)j * N]));
    for (int j = 0; j < p; j++) {
        for (int i = 0; i < N; i++) Q1[i + (size_t)j * N] = Q[i + (size_t)j * N];
        for (int i = 0; i <= j; i++) R1[i + (size_t)j * p] = R[i + (size_t)j * N];
        if (!(std::fabs(R[j + (size_t)j * N]) > 1e-10 * dmax)) return false;
    }
    return true;
}

// Solves R1' w = b (forward substitution), R1 upper triangular p x p.
void solve_rt(int p, const std::vector<double>& R1, const double* b, double* w) {
    for (int i = 0; i < p; i++) {
        double acc = b[i];
        for (int j = 0; j < i; j++) acc -= R1[j + (size_t)i * p] * w[j];
        w[i] = acc / R1[i + (size_t)i * p];
    }
}

// Gauss-Jordan inverse with partial pivoting (n x n column-major, in place).
bool invert(int n, std::vector<double>& A) {
    std::vector<double> B((size_t)n * n, 0.0);
    for (int i = 0; i < n; i++) B[i + (size_t)i * n] = 1.0;
    for (int k = 0; k < n; k++) {
        int piv = k;
        for (int i = k + 1; i < n; i++)
            if (std::fabs(A[i + (size_t)k * n]) > std::fabs(A[piv + (size_t)k * n])) piv = i;
        if (A[piv + (size_t)k * n] == 0.0) return false;
        if (piv != k)
            for (int j = 0; j < n; j++) {
                std::swap(A[k + (size_t)j * n], A[piv + (size_t)j * n]);
                std::swap(B[k + (size_t)j * n], B[piv + (size_t)j * n]);
            }
        const double d = A[k + (size_t)k * n];
        for (int j = 0; j < n; j++) { A[k + (size_t)j * n] /= d; B[k + (size_t)j * n] /= d; }
        for (int i = 0; i < n; i++) {
            if (i == k) continue;
            const double f = A[i + (size_t)k * n];
            if (f == 0.0) continue;
            for (int j = 0; j < n; j++) {
                A[i + (size_t)j * n] -= f * A[k + (size_t)j * n];
                B[i + (size_t)j * n] -= f * B[k + (size_t)j * n];
            }
        }
    }
    A = B;
    return true;
}

std::vector<double> to_double(const float* a, size_t n) {
    std::vector<double> r(n);
    for (size_t i = 0; i < n; i++) r[i] = (double)a[i];
    return r;
}

// The constant vector is the one direction every row permutation leaves in
// place (P'1 = 1).  When it lies in the nuisance space, the residualised
// data are orthogonal to it, so its coordinate 1'P'y*/sqrt(N) is zero for
// every permutation and contributes nothing to |z|^2.  The nuisance columns
// of Qt (which span the same space as Qn) are then rotated so that one of
// them is 1/sqrt(N), and that column is dropped from the permuted fit; the
// degrees of freedom stay N - p.
void drop_invariant_axis(int N, int p, MapModel& m) {
    const int r = m.r, nn = p - r;
    if (nn < 1) return;
    const double e = 1.0 / std::sqrt((double)N);
    // coordinates of 1/sqrt(N) in the nuisance columns of Qt
    std::vector<double> a(nn);
    double in = 0.0;
    for (int j = 0; j < nn; j++) {
        double acc = 0.0;
        for (int s = 0; s < N; s++) acc += m.Qt[s + (size_t)(r + j) * N] * e;
        a[j] = acc;
        in += acc * acc;
    }
    if (!(std::fabs(in - 1.0) <= 1e-10)) return;  // not in the nuisance space
    // orthogonal U with first column +-a: the rotated columns Qt_n U keep
    // the span, their first one is +-1/sqrt(N)
    std::vector<double> U((size_t)nn * nn);
    householder_qr(nn, 1, a.data(), U.data(), nullptr);
    std::vector<double> Qt((size_t)N * (p - 1), 0.0);
    std::copy(m.Qt.begin(), m.Qt.begin() + (size_t)N * r, Qt.begin());
    for (int c = 1; c < nn; c++)
        for (int s = 0; s < N; s++) {
            double acc = 0.0;
            for (int j = 0; j < nn; j++) acc += m.Qt[s + (size_t)(r + j) * N] * U[j + (size_t)c * nn];
            Qt[s + (size_t)(r + c - 1) * N] = acc;
        }
    m.Qt.swap(Qt);
    m.pfit = p - 1;
}

}  // namespace

int build_map_model(const bpg_study_desc& d, int map, MapModel& m, std::string& err) {
    const int N = d.N;
    if (d.test == BPG_GROUP_MEAN) {
        // K2:1458-1559: OLS of (sign o y) on the first column of X (the CLI
        // passes all ones, RGL:841-848), t over N - 1 degrees of freedom.
        double nrm = 0.0;
        for (int s = 0; s < N; s++) nrm += (double)d.X[s] * d.X[s];
        nrm = std::sqrt(nrm);
        if (!(nrm > 0.0)) { err = "group mean: the design column is zero"; return BPG_ERR_DESIGN; }
        m.r = 1; m.q = 0; m.pfit = 1; m.dof = N - 1;
        m.Qt.resize(N);
        m.dscale.resize(N);
        for (int s = 0; s < N; s++) m.Qt[s] = m.dscale[s] = (double)d.X[s] / nrm;
        m.Qn.clear();
        m.sign_operand = true;
        return BPG_OK;
    }
    const int p = d.p;
    std::vector<double> X = to_double(d.X, (size_t)N * p);
    std::vector<double> Q1, R1;
    if (!thin_qr(N, p, X, Q1, R1)) { err = "design matrix is rank deficient"; return BPG_ERR_DESIGN; }
    // contrast rows of this map
    const int r = (d.test == BPG_TTEST) ? 1 : d.nc;
    std::vector<double> Crow((size_t)r * p);  // Crow[j + i*p] = C(i, j)
    for (int i = 0; i < r; i++)
        for (int j = 0; j < p; j++)
            Crow[j + (size_t)i * p] = (double)d.C[j + (size_t)((d.test == BPG_TTEST) ? map : i) * p];
    // W' = R1^-T C' (p x r); its column space, rotated into Q1, is the
    // contrast space X (X'X)^-1 C'.
    std::vector<double> Wt((size_t)p * r);
    for (int i = 0; i < r; i++) solve_rt(p, R1, &Crow[(size_t)i * p], &Wt[(size_t)i * p]);
    std::vector<double> U((size_t)p * p), RW((size_t)p * r);
    householder_qr(p, r, Wt.data(), U.data(), RW.data());
    double wmax = 0.0;
    for (int i = 0; i < r; i++) wmax = std::max(wmax, std::fabs(Wt[i + (size_t)i * p]));
    for (int i = 0; i < r; i++) {
        double cn = 0.0;
        for (int j = 0; j < p; j++) cn += Wt[j + (size_t)i * p] * Wt[j + (size_t)i * p];
        if (!(cn > 0.0)) { err = "contrast is zero"; return BPG_ERR_DESIGN; }
    }
    for (int i = 0; i < r; i++) {
        double nrm = 0.0;
        for (int j = 0; j < p; j++) nrm += Wt[j + (size_t)i * p] * Wt[j + (size_t)i * p];
        if (!(std::fabs(RW[i + (size_t)i * p]) > 1e-10 * std::sqrt(nrm))) {
            err = "F contrast matrix is rank deficient";
            return BPG_ERR_DESIGN;
        }
    }
    if (r == 1 && RW[0] < 0.0)  // orient the t axis along +c b
        for (int j = 0; j < p; j++) U[j] = -U[j];
    m.r = r;
    m.pfit = p;
    m.dof = N - p;
    m.Qt.assign((size_t)N * p, 0.0);
    for (int c = 0; c < p; c++)
        for (int s = 0; s < N; s++) {
            double acc = 0.0;
            for (int k = 0; k < p; k++) acc += Q1[s + (size_t)k * N] * U[k + (size_t)c * p];
            m.Qt[s + (size_t)c * N] = acc;
        }
    // nuisance space X null(C) (LIB:16654-16727): null(C) from the full QR
    // of C' (p x r), columns r..p-1.
    const bool transform = (d.test == BPG_TTEST && p > 1) || (d.test == BPG_FTEST && p > d.nc);
    m.q = transform ? p - r : 0;
    m.Qn.clear();
    if (m.q > 0) {
        std::vector<double> Qc((size_t)p * p);
        householder_qr(p, r, Crow.data(), Qc.data(), nullptr);
        std::vector<double> W2((size_t)N * m.q);
        for (int j = 0; j < m.q; j++)
            for (int s = 0; s < N; s++) {
                double acc = 0.0;
                for (int k = 0; k < p; k++) acc += X[s + (size_t)k * N] * Qc[k + (size_t)(r + j) * p];
                W2[s + (size_t)j * N] = acc;
            }
        std::vector<double> Qn, Rn;
        if (!thin_qr(N, m.q, W2, Qn, Rn)) { err = "nuisance regressors are rank deficient"; return BPG_ERR_DESIGN; }
        m.Qn = Qn;
        drop_invariant_axis(N, p, m);
    }
    m.dscale.assign(N, 1.0);
    m.sign_operand = false;
    return BPG_OK;
}

// First-level model: X = Q1 R1 (thin QR in FP64), z = Q1' x gives
// c'beta = (R1^-T c)' z and c'(X'X)^-1 c = |R1^-T c|^2, so the t statistic
// row is d / |d| with d = R1^-T c.  For F, D = C R1^-1 (rows R1^-T C_i) and
// C(X'X)^-1C' = D D' = L L' (Cholesky), so (C beta)'(C(X'X)^-1C')^-1(C beta)
// = |L^-1 D z|^2 and the rows are L^-1 D.
int build_fl_model(int T, int p, int nc, int ftest, const float* X, const float* C, int pq, FlModel& m,
                   std::string& err) {
    if (pq < p) { err = "internal: design width"; return BPG_ERR_ARG; }
    const std::vector<double> Xd = [&] {
        std::vector<double> r((size_t)T * p);
        for (size_t i = 0; i < r.size(); i++) r[i] = (double)X[i];  // X[t + r*T]
        return r;
    }();
    std::vector<double> Q1, R1;
    if (!thin_qr(T, p, Xd, Q1, R1)) { err = "the design matrix is rank deficient"; return BPG_ERR_ARG; }
    m = FlModel{};
    m.T = T; m.p = p; m.pq = pq; m.ftest = ftest != 0;
    m.Q.assign((size_t)T * pq, 0.0);
    m.u.assign(pq, 0.0);
    for (int t = 0; t < T; t++)
        for (int j = 0; j < p; j++) {
            m.Q[(size_t)t * pq + j] = Q1[t + (size_t)j * T];
            m.u[j] += Q1[t + (size_t)j * T];
        }
    {
        // |P_X 1|^2 = |Q1' 1|^2 equals T exactly when 1 is in col(X)
        double uu = 0.0;
        for (int j = 0; j < p; j++) uu += m.u[j] * m.u[j];
        m.const_in_span = (double)T - uu <= 1e-9 * (double)T;
    }
    // pinv(X) = R1^-1 Q1'
    std::vector<double> Ri((size_t)p * p, 0.0);  // column-major R1^-1
    for (int c = 0; c < p; c++) {
        for (int r = p - 1; r >= 0; r--) {
            double acc = (r == c) ? 1.0 : 0.0;
            for (int j = r + 1; j < p; j++) acc -= R1[r + (size_t)j * p] * Ri[j + (size_t)c * p];
            Ri[r + (size_t)c * p] = acc / R1[r + (size_t)r * p];
        }
    }
    m.Xp.assign((size_t)p * T, 0.0);
    m.X.assign((size_t)T * p, 0.0);
    for (int t = 0; t < T; t++)
        for (int r = 0; r < p; r++) {
            double acc = 0.0;
            for (int j = 0; j < p; j++) acc += Ri[r + (size_t)j * p] * Q1[t + (size_t)j * T];
            m.Xp[(size_t)r * T + t] = acc;
            m.X[(size_t)t * p + r] = Xd[t + (size_t)r * T];
        }
    std::vector<std::vector<double>> D(nc, std::vector<double>(p));
    for (int c = 0; c < nc; c++) {
        std::vector<double> cc(p);
        for (int r = 0; r < p; r++) cc[r] = (double)C[r + (size_t)c * p];
        solve_rt(p, R1, cc.data(), D[c].data());
    }
    if (!ftest) {
        m.nmaps = nc;
        for (int c = 0; c < nc; c++) {
            double n2 = 0.0;
            for (double v : D[c]) n2 += v * v;
            if (!(n2 > 0.0)) { err = "a contrast is zero"; return BPG_ERR_ARG; }
            std::vector<double> row(pq, 0.0);
            for (int j = 0; j < p; j++) row[j] = D[c][j] / std::sqrt(n2);
            m.A.push_back(row);
            m.nrows.push_back(1);
        }
    } else {
        m.nmaps = 1;
        std::vector<double> L((size_t)nc * nc, 0.0);  // row-major lower
        for (int i = 0; i < nc; i++)
            for (int j = 0; j <= i; j++) {
                double g = 0.0;
                for (int q = 0; q < p; q++) g += D[i][q] * D[j][q];
                for (int k = 0; k < j; k++) g -= L[(size_t)i * nc + k] * L[(size_t)j * nc + k];
                if (i == j) {
                    if (!(g > 1e-12 * std::max(1.0, std::fabs(g)))) {
                        err = "the F contrasts are linearly dependent";
                        return BPG_ERR_ARG;
                    }
                    L[(size_t)i * nc + i] = std::sqrt(g);
                } else {
                    L[(size_t)i * nc + j] = g / L[(size_t)j * nc + j];
                }
            }
        std::vector<double> rows((size_t)nc * pq, 0.0);
        for (int i = 0; i < nc; i++)
            for (int q = 0; q < p; q++) {
                double acc = D[i][q];
                for (int k = 0; k < i; k++) acc -= L[(size_t)i * nc + k] * rows[(size_t)k * pq + q];
                rows[(size_t)i * pq + q] = acc / L[(size_t)i * nc + i];
            }
        m.A.push_back(rows);
        m.nrows.push_back(nc);
    }
    return BPG_OK;
}

int build_obs_model(const bpg_study_desc& d, ObsModel& o, std::string& err) {
    const int N = d.N;
    const bool mean = d.test == BPG_GROUP_MEAN;
    o.p = mean ? 1 : d.p;
    o.nc = mean ? 1 : d.nc;
    o.ftest = d.test == BPG_FTEST;
    o.nmaps = (d.test == BPG_TTEST) ? d.nc : 1;
    const int p = o.p, nc = o.nc;
    o.X = to_double(d.X, (size_t)N * p);
    o.C = to_double(d.C, (size_t)p * nc);
    std::vector<double> Q1, R1;
    if (!thin_qr(N, p, o.X, Q1, R1)) { err = "design matrix is rank deficient"; return BPG_ERR_DESIGN; }
    // pinv(X) = R1^-1 Q1'
    o.Xp.assign((size_t)p * N, 0.0);
    for (int s = 0; s < N; s++) {
        double b[BPG_MAX_REGRESSORS];
        for (int r = 0; r < p; r++) b[r] = Q1[s + (size_t)r * N];
        for (int r = p - 1; r >= 0; r--) {
            double acc = b[r];
            for (int j = r + 1; j < p; j++) acc -= R1[r + (size_t)j * p] * b[j];
            b[r] = acc / R1[r + (size_t)r * p];
        }
        for (int r = 0; r < p; r++) o.Xp[r + (size_t)s * p] = b[r];
    }
    // W_i = R1^-T c_i': c_i (X'X)^-1 c_j' = W_i . W_j
    std::vector<double> Wt((size_t)p * nc);
    for (int i = 0; i < nc; i++) solve_rt(p, R1, &o.C[(size_t)i * p], &Wt[(size_t)i * p]);
    if (!o.ftest) {
        o.sc.resize(nc);
        for (int i = 0; i < nc; i++) {
            double acc = 0.0;
            for (int j = 0; j < p; j++) acc += Wt[j + (size_t)i * p] * Wt[j + (size_t)i * p];
            o.sc[i] = acc;
        }
    } else {
        std::vector<double> M((size_t)nc * nc);
        for (int a = 0; a < nc; a++)
            for (int b = 0; b < nc; b++) {
                double acc = 0.0;
                for (int j = 0; j < p; j++) acc += Wt[j + (size_t)a * p] * Wt[j + (size_t)b * p];
                M[a + (size_t)b * nc] = acc;
            }
        if (!invert(nc, M)) { err = "F contrast matrix is rank deficient"; return BPG_ERR_DESIGN; }
        o.sc = M;
    }
    return BPG_OK;
}

}  // namespace bpg

// ========================================================================
// C ABI: generators, design detection, model precompute
// ========================================================================

namespace {

struct SplitMix64 {
    unsigned long long s;
    unsigned long long next() {
        unsigned long long z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    // uniform integer in [0, n), rejection-free multiply-shift on 32 bits
    unsigned bounded(unsigned n) { return (unsigned)(((next() >> 32) * (unsigned long long)n) >> 32); }
};

// glibc's rand() (random_r, TYPE_3: additive feedback r[i] = r[i-31] +
// r[i-3] over a 34-word history seeded by the Park-Miller recurrence
// 16807*x mod (2^31-1), 310 warm-up draws, output r[i] >> 1), as a private
// state.  The reference draws from the process-global rand() without ever
// seeding it (LIB:16925-17122), i.e. from seed 1; a private copy reproduces
// that stream whatever else in the process (e.g. a GPU runtime) calls
// rand().
class GlibcRand {
public:
    GlibcRand() { seed(1); }
    void seed(unsigned s) {
        if (s == 0) s = 1;
        int32_t r[34];
        r[0] = (int32_t)s;
        for (int i = 1; i < 31; i++) {
            const int64_t hi = r[i - 1] / 127773, lo = r[i - 1] % 127773;
            int64_t w = 16807 * lo - 2836 * hi;
            if (w < 0) w += 2147483647;
            r[i] = (int32_t)w;
        }
        for (int i = 31; i < 34; i++) r[i] = r[i - 31];
        for (int i = 0; i < 34; i++) h_[i] = (uint32_t)r[i];
        pos_ = 34;
        for (int i = 34; i < 344; i++) step();
    }
    int next() { return (int)(step() >> 1); }

private:
    uint32_t step() {
        const uint32_t v = h_[(pos_ - 31) % 34] + h_[(pos_ - 3) % 34];
        h_[pos_ % 34] = v;
        pos_++;
        return v;
    }
    uint32_t h_[34];
    uint64_t pos_ = 0;
};

GlibcRand& compat_stream() {
    static GlibcRand g;
    return g;
}

// libstdc++ std::random_shuffle(first, last) with rand():
// for i = 1..n-1, j = rand() % (i+1), swap unless i == j.
template <class T>
void compat_shuffle(T* a, int n) {
    GlibcRand& g = compat_stream();
    for (int i = 1; i < n; i++) {
        const int j = g.next() % (i + 1);
        if (i != j) std::swap(a[i], a[j]);
    }
}

template <class T>
void seeded_shuffle(T* a, int n, SplitMix64& g) {
    for (int i = n - 1; i > 0; i--) std::swap(a[i], a[g.bounded((unsigned)i + 1)]);
}

template <class T>
std::string key_of(const T* a, int n) {
    return std::string(reinterpret_cast<const char*>(a), sizeof(T) * (size_t)n);
}

// Number of distinct vectors, saturated at 2^62.
double distinct_count(int kind, int N, int n1) {
    if (kind == 0) return std::ldexp(1.0, std::min(N, 62));
    if (kind == 1) return std::round(std::exp(std::lgamma(N + 1.0) - std::lgamma(n1 + 1.0) - std::lgamma(N - n1 + 1.0)));
    return std::exp(std::lgamma(N + 1.0));
}

void fill_two_sample(const int* g, int N, int n1, unsigned short* out) {
    int a = 0, b = 0;
    for (int i = 0; i < N; i++) out[i] = (unsigned short)(g[i] == 1 ? a++ : n1 + b++);
}

}  // namespace

extern "C" int bpg_gen_sign_matrix(int N, long K, int compat, unsigned long long seed, float* sign) {
    if (N <= 0 || N > BPG_MAX_SUBJECTS || K <= 0 || !sign) return BPG_ERR_ARG;
    if ((double)K > distinct_count(0, N, 0)) return BPG_ERR_ARG;
    std::vector<int> flips(N, 1);
    std::unordered_set<std::string> seen;
    seen.insert(key_of(flips.data(), N));
    for (int i = 0; i < N; i++) sign[i] = 1.0f;
    SplitMix64 g{seed};
    if (compat && seed) compat_stream().seed((unsigned)seed);
    for (long k = 1; k < K; k++) {
        for (;;) {
            if (compat) {
                // cumulative flips, LIB:17093-17097
                for (int i = 0; i < N; i++) flips[i] *= (2 * (int)(compat_stream().next() % 2) - 1);
            } else {
                for (int i = 0; i < N; i++) flips[i] = (g.next() >> 63) ? 1 : -1;
            }
            if (seen.insert(key_of(flips.data(), N)).second) break;
        }
        for (int i = 0; i < N; i++) sign[i + (size_t)k * N] = (float)flips[i];
    }
    return BPG_OK;
}

extern "C" int bpg_gen_perm_two_sample(int N, int n1, int n2, long K, int compat,
                                       unsigned long long seed, unsigned short* perm) {
    if (N <= 0 || N > BPG_MAX_SUBJECTS || n1 < 0 || n2 < 0 || n1 + n2 != N || K <= 0 || !perm)
        return BPG_ERR_ARG;
    if ((double)K > distinct_count(1, N, n1)) return BPG_ERR_ARG;
    std::vector<int> g(N);
    for (int i = 0; i < N; i++) g[i] = i < n1 ? 1 : -1;
    std::unordered_set<std::string> seen;
    seen.insert(key_of(g.data(), N));
    fill_two_sample(g.data(), N, n1, perm);
    SplitMix64 rng{seed};
    if (compat && seed) compat_stream().seed((unsigned)seed);
    for (long k = 1; k < K; k++) {
        for (;;) {
            if (compat) compat_shuffle(g.data(), N);
            else seeded_shuffle(g.data(), N, rng);
            if (seen.insert(key_of(g.data(), N)).second) break;
        }
        fill_two_sample(g.data(), N, n1, perm + (size_t)k * N);
    }
    return BPG_OK;
}

extern "C" int bpg_gen_perm_correlation(int N, long K, int compat, unsigned long long seed,
                                        unsigned short* perm) {
    if (N <= 0 || N > BPG_MAX_SUBJECTS || K <= 0 || !perm) return BPG_ERR_ARG;
    if ((double)K > distinct_count(2, N, 0)) return BPG_ERR_ARG;
    std::vector<unsigned short> cur(N);
    for (int i = 0; i < N; i++) cur[i] = (unsigned short)i;
    std::unordered_set<std::string> seen;
    seen.insert(key_of(cur.data(), N));
    std::copy(cur.begin(), cur.end(), perm);
    SplitMix64 rng{seed};
    if (compat && seed) compat_stream().seed((unsigned)seed);
    for (long k = 1; k < K; k++) {
        for (;;) {
            if (compat) compat_shuffle(cur.data(), N);
            else seeded_shuffle(cur.data(), N, rng);
            if (seen.insert(key_of(cur.data(), N)).second) break;
        }
        std::copy(cur.begin(), cur.end(), perm + (size_t)k * N);
    }
    return BPG_OK;
}

extern "C" int bpg_gen_perm_first_level(int T, long K, int compat, unsigned long long seed,
                                        unsigned short* perm) {
    if (T <= 1 || T > 65535 || K <= 0 || !perm) return BPG_ERR_ARG;
    if ((double)K + 1.0 > distinct_count(2, T, 0)) return BPG_ERR_ARG;
    std::vector<unsigned short> cur(T);
    for (int i = 0; i < T; i++) cur[i] = (unsigned short)i;
    std::unordered_set<std::string> seen;
    seen.insert(key_of(cur.data(), T));  // the identity is never a row
    SplitMix64 rng{seed};
    if (compat && seed) compat_stream().seed((unsigned)seed);
    for (long k = 0; k < K; k++) {
        for (;;) {
            if (compat) compat_shuffle(cur.data(), T);
            else seeded_shuffle(cur.data(), T, rng);
            if (seen.insert(key_of(cur.data(), T)).second) break;
        }
        std::copy(cur.begin(), cur.end(), perm + (size_t)k * T);
    }
    return BPG_OK;
}

extern "C" int bpg_detect_designs(int N, int p, int nc, const float* X, const float* C,
                                  int ftest, int* designs, int* n1, int* n2) {
    if (N <= 0 || p <= 0 || nc <= 0 || !X || !C || !designs || !n1 || !n2) return BPG_ERR_ARG;
    // X*c' in double, compared as float (RGL:918-940, RGL:987-1003)
    std::vector<float> v((size_t)N * nc);
    for (int c = 0; c < nc; c++)
        for (int s = 0; s < N; s++) {
            double acc = 0.0;
            for (int r = 0; r < p; r++) acc += (double)X[s + (size_t)r * N] * (double)C[r + (size_t)c * p];
            v[s + (size_t)c * N] = (float)acc;
        }
    const int groups = ftest ? 1 : nc;
    for (int gi = 0; gi < groups; gi++) {
        std::vector<float> uniq;
        for (int s = 0; s < N; s++)
            for (int c = ftest ? 0 : gi; c < (ftest ? nc : gi + 1); c++) {
                const float x = v[s + (size_t)c * N];
                if (std::find(uniq.begin(), uniq.end(), x) == uniq.end()) uniq.push_back(x);
            }
        designs[gi] = uniq.size() == 2 ? BPG_TWOSAMPLE : uniq.size() > 2 ? BPG_CORRELATION : BPG_MEAN_DESIGN;
        n1[gi] = n2[gi] = 0;
        if (uniq.size() == 2) {
            const int col = ftest ? 0 : gi;
            for (int s = 0; s < N; s++) n1[gi] += (v[s + (size_t)col * N] == uniq[0]);
            n2[gi] = N - n1[gi];
        }
    }
    return BPG_OK;
}

extern "C" double bpg_max_permutations(int test, int design, int N, int n2) {
    if (test == BPG_GROUP_MEAN) return std::pow(2.0, (double)N);
    if (design == BPG_TWOSAMPLE)
        return std::round(std::exp(std::lgamma(N + 1) - std::lgamma(N - n2 + 1) - std::lgamma(n2 + 1)));
    if (design == BPG_CORRELATION) {
        // the CLI's Stirling "factorial" with pi ~ 3.14, e ~ 2.7183 (RGL:38-42)
        const double n = (double)N;
        return (N <= 1) ? 1.0 : std::round(std::sqrt(2.0 * 3.14 * n) * std::pow(n / 2.7183, n));
    }
    return 1.0;
}

extern "C" int bpg_model_precompute(int N, int p, int nc, int ftest, const float* X,
                                    const float* C, float* xtxxt, float* ctxtxc) {
    if (N <= 0 || p <= 0 || p > BPG_MAX_REGRESSORS || nc <= 0 || !X || !C || !xtxxt || !ctxtxc)
        return BPG_ERR_ARG;
    if (ftest && nc > BPG_MAX_FROWS) return BPG_ERR_ARG;
    bpg_study_desc d{};
    d.test = ftest ? BPG_FTEST : BPG_TTEST;
    d.N = N; d.p = p; d.nc = nc; d.X = X; d.C = C;
    bpg::ObsModel o;
    std::string err;
    const int rc = bpg::build_obs_model(d, o, err);
    if (rc) return rc;
    for (int s = 0; s < N; s++)
        for (int r = 0; r < p; r++) xtxxt[s + (size_t)r * N] = (float)o.Xp[r + (size_t)s * p];
    const int ns = ftest ? nc * nc : nc;
    for (int i = 0; i < ns; i++) ctxtxc[i] = (float)o.sc[i];
    return BPG_OK;
}
