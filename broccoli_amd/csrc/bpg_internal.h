// bpg_internal.h -- shared declarations of the bpg library (not installed).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "bpg.h"

namespace bpg {

// FP64 model of one statistical map (one t contrast, the F contrast, or the
// group mean).  The permuted fit of the reference (K2:1562-1669: OLS of the
// residualised data on the row-permuted design) is rewritten with an
// orthonormal basis Qt of col(X) whose first r columns span the contrast
// space X (X'X)^-1 C':
//   z   = Qt' P' y*                 (p values per voxel and permutation)
//   SSE = |y*|^2 - |z|^2
//   t   = z_0 * sqrt(dof / SSE)                  (= c b / sqrt(s2 c(X'X)^-1 c'))
//   F   = (z_0^2 + .. + z_{r-1}^2) * dof / (r SSE)
// Qn spans the nuisance space X null(C) projected out before the loop
// (LIB:16654-16727).  For the group mean the operand rows are scaled by
// Qt(:,0) and the permutation operand holds the exact signs.
struct MapModel {
    int r = 1;      // contrast rows
    int q = 0;      // nuisance dimension (0: no residualisation)
    int pfit = 1;   // columns of Qt in the permuted fit (p, or p - 1 without the constant axis)
    int dof = 1;    // N - p
    std::vector<double> Qt;     // N x pfit, column-major
    std::vector<double> Qn;     // N x q, column-major
    std::vector<double> dscale; // N, operand row scale
    bool sign_operand = false;  // group mean: the operand is the sign matrix
};

// Observed-map model (K1:51-128, K1:1456-1673) in FP64.
struct ObsModel {
    int p = 1, nc = 1, nmaps = 1;
    bool ftest = false;
    std::vector<double> X;    // N x p column-major
    std::vector<double> Xp;   // p x N: Xp[r + s*p] = pinv(X)(r,s)
    std::vector<double> C;    // C[r + c*p]
    std::vector<double> sc;   // t: c(X'X)^-1c' per contrast; F: (C(X'X)^-1C')^-1
};

// First-level model (kernelStatistics3.cpp:2107-2190,
// kernelStatistics5.cpp:1943-2060) in FP64: an orthonormal basis Q of
// col(X), Q'1 and, per statistical map, the rows A with t = A_0 z / sd and
// F = |A z|^2 / (var nrows) for z = Q'x.  Also pinv(X) and X for the
// regression of the full design (LIB:12144-12194).
struct FlModel {
    int T = 0, p = 0, pq = 0, nmaps = 1;
    bool ftest = false;
    bool const_in_span = false;          // 1 in col(X): the residuals sum to 0
    std::vector<double> Q;               // [t*pq + j], zero beyond p
    std::vector<double> u;               // [pq]
    std::vector<std::vector<double>> A;  // per map: [i*pq + j]
    std::vector<int> nrows;              // per map
    std::vector<double> Xp;              // [r*T + t] = pinv(X)(r, t)
    std::vector<double> X;               // [t*p + r]
};

int householder_qr(int m, int n, const double* A, double* Q, double* R);
int build_fl_model(int T, int p, int nc, int ftest, const float* X, const float* C, int pq, FlModel& m,
                   std::string& err);
int build_map_model(const bpg_study_desc& d, int map, MapModel& out, std::string& err);
int build_obs_model(const bpg_study_desc& d, ObsModel& out, std::string& err);

}  // namespace bpg
