// ref_harness.cpp -- exposes CPU-only routines of the reference library,
// compiled from the reference's own unmodified sources (oracle/Makefile,
// target "ref"), so the tests can check the oracle against them.
//
// TEST INFRASTRUCTURE ONLY.  Output goes to oracle/_ref/ and never ships.
//
// What is reachable without an OpenCL device:
//   - GenerateSignMatrixSecondLevel            (LIB:17069-17122)
//   - GeneratePermutationMatrixSecondLevelTwoSample (LIB:16925-17011)
//   - GeneratePermutationMatrixSecondLevelCorrelation (LIB:17014-17066)
//   - GeneratePermutationMatrixFirstLevel      (LIB:16877-16922)
//   - Clusterize, the host 26-connected BFS     (LIB:18163-18297)
//   - pinv, the free function                  (LIB:16561-16586)
// These members are private in broccoli_lib.h (HDR:469-547); the harness
// widens access for its own translation unit only, after the system and
// Eigen headers (the class layout and the compiled reference code are
// unchanged).  BROCCOLI_LIB() (LIB:130-134)
// initialises members without touching OpenCL.  Objects are never deleted:
// the destructor (LIB:160-163) releases OpenCL handles that were never
// created.
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>
#include <opencl.h>
#include <Dense>
#include "broccoli_constants.h"
#define private public
#include "broccoli_lib.h"
#undef private

Eigen::MatrixXd pinv(const Eigen::MatrixXd mat);

extern "C" {

int ref_gen_sign_matrix(int N, long K, float* sign) {
    BROCCOLI_LIB* b = new BROCCOLI_LIB();
    static size_t kc[1];
    kc[0] = (size_t)K;
    b->SetNumberOfSubjects((size_t)N);
    b->SetNumberOfGroupPermutations(kc);
    b->SetSignMatrix(sign);
    b->GenerateSignMatrixSecondLevel();
    return 0;
}

int ref_gen_perm_two_sample(int N, int n1, int n2, long K, unsigned short* perm) {
    BROCCOLI_LIB* b = new BROCCOLI_LIB();
    static size_t kc[1];
    static int g1[1], g2[1];
    static unsigned short* mats[1];
    kc[0] = (size_t)K;
    g1[0] = n1;
    g2[0] = n2;
    mats[0] = perm;
    b->SetNumberOfSubjects((size_t)N);
    b->SetNumberOfSubjectsGroup1(g1);
    b->SetNumberOfSubjectsGroup2(g2);
    b->SetNumberOfGroupPermutations(kc);
    b->SetPermutationMatrices(mats);
    b->GeneratePermutationMatrixSecondLevelTwoSample(0);
    return 0;
}

int ref_gen_perm_correlation(int N, long K, unsigned short* perm) {
    BROCCOLI_LIB* b = new BROCCOLI_LIB();
    static size_t kc[1];
    static unsigned short* mats[1];
    kc[0] = (size_t)K;
    mats[0] = perm;
    b->SetNumberOfSubjects((size_t)N);
    b->SetNumberOfGroupPermutations(kc);
    b->SetPermutationMatrices(mats);
    b->GeneratePermutationMatrixSecondLevelCorrelation(0);
    return 0;
}

// Time-point permutations of the first-level test: K rows of T indices.
int ref_gen_perm_first_level(int T, long K, unsigned short* perm) {
    BROCCOLI_LIB* b = new BROCCOLI_LIB();
    b->SetEPITimepoints((size_t)T);
    b->SetNumberOfPermutations((size_t)K);
    b->SetPermutationMatrix(perm);
    b->GeneratePermutationMatrixFirstLevel();
    return 0;
}

// labels: 1..n in discovery order (0 = background), as LIB:18163-18297.
int ref_clusterize(int W, int H, int D, float* data, float thr, float* mask,
                   int* labels, int* max_size, int* n_clusters) {
    BROCCOLI_LIB* b = new BROCCOLI_LIB();
    int maxsz = 0, ncl = 0;
    float maxmass = 0.0f;
    b->Clusterize(labels, maxsz, maxmass, ncl, data, thr, mask, (size_t)W, (size_t)H,
                  (size_t)D, 1, 0);
    *max_size = maxsz;
    *n_clusters = ncl;
    return 0;
}

// out[r + c*cols_of_result] column-major, result = pinv(mat) (cols x rows).
int ref_pinv(int rows, int cols, const double* mat, double* out) {
    Eigen::MatrixXd m(rows, cols);
    for (int j = 0; j < cols; j++)
        for (int i = 0; i < rows; i++) m(i, j) = mat[i + (size_t)j * rows];
    Eigen::MatrixXd pm = pinv(m);
    for (int j = 0; j < pm.cols(); j++)
        for (int i = 0; i < pm.rows(); i++) out[i + (size_t)j * pm.rows()] = pm(i, j);
    return 0;
}

}  // extern "C"
