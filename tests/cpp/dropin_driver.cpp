// dropin_driver.cpp -- drives the BROCCOLI_LIB drop-in (include/broccoli_lib.h)
// with the call sequence of the reference CLI
// (code/Bash_Wrapper/RandomiseGroupLevel.cpp): design detection
// (RGL:908-1044), permutation-count clamp (RGL:1048-1134), FP64 model
// precompute (RGL:1137-1194), setters (RGL:1488-1524), one Perform call
// (RGL:1529-1546), error arrays (RGL:1555-1583).
//
// Usage: dropin_driver <dir>
//   <dir>/study.txt : W H D N p nc test inference K cdt use_perm_file precision
//   <dir>/Y.bin (N*W*H*D f32), mask.bin, X.bin (N*p, X[s + r*N]), C.bin
//   (p*nc), optionally perm.bin (K*N u16, 0-based) / sign.bin (K*N f32).
// Writes maps.bin, pvalues.bin, dist.bin (per map, K floats) and
// perm_used.bin / sign_used.bin; prints the error arrays.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "bpg.h"
#include "broccoli_lib.h"

template <class T>
static bool read_bin(const std::string& path, std::vector<T>& v, size_t n) {
    std::ifstream f(path, std::ios::binary);
    if (!f) return false;
    v.resize(n);
    f.read(reinterpret_cast<char*>(v.data()), n * sizeof(T));
    return (size_t)f.gcount() == n * sizeof(T);
}

template <class T>
static void write_bin(const std::string& path, const T* p, size_t n) {
    std::ofstream f(path, std::ios::binary);
    f.write(reinterpret_cast<const char*>(p), n * sizeof(T));
}

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s <dir>\n", argv[0]);
        return 2;
    }
    const std::string dir = argv[1];
    std::ifstream st(dir + "/study.txt");
    size_t W, H, D, N, p, nc, K;
    int test, inference, use_file, precision;
    float cdt;
    if (!(st >> W >> H >> D >> N >> p >> nc >> test >> inference >> K >> cdt >> use_file >> precision)) {
        fprintf(stderr, "bad study.txt\n");
        return 2;
    }
    const size_t G = W * H * D;
    std::vector<float> Y, mask, X, C;
    if (!read_bin(dir + "/Y.bin", Y, N * G) || !read_bin(dir + "/mask.bin", mask, G) ||
        !read_bin(dir + "/X.bin", X, N * p) || !read_bin(dir + "/C.bin", C, p * nc)) {
        fprintf(stderr, "missing inputs\n");
        return 2;
    }
    const bool ftest = test == 1, mean = test == 2;
    const size_t nmaps = test == 0 ? nc : 1;

    // design detection and permutation counts (RGL:908-1134)
    std::vector<int> designs(nmaps, BPG_CORRELATION), n1(nmaps, 0), n2(nmaps, 0);
    std::vector<size_t> Kc(nmaps, K);
    if (!mean) {
        std::vector<int> d(nc), a(nc), b(nc);
        bpg_detect_designs((int)N, (int)p, (int)nc, X.data(), C.data(), ftest, d.data(), a.data(), b.data());
        for (size_t c = 0; c < nmaps; c++) {
            designs[c] = d[c];
            n1[c] = a[c];
            n2[c] = b[c];
            const double mx = bpg_max_permutations(test, designs[c], (int)N, n2[c]);
            if (designs[c] == BPG_MEAN_DESIGN) Kc[c] = 1;
            else if ((double)K > mx) Kc[c] = (size_t)mx;
        }
    } else {
        const double mx = bpg_max_permutations(2, 0, (int)N, 0);
        if ((double)K > mx) Kc[0] = (size_t)mx;
    }
    // FP64 model precompute cast to float (RGL:1137-1194)
    std::vector<float> xtxxt(N * p), ctxtxc(ftest ? nc * nc : nc);
    if (bpg_model_precompute((int)N, (int)p, (int)nc, ftest, X.data(), C.data(), xtxxt.data(), ctxtxc.data())) {
        fprintf(stderr, "model precompute failed\n");
        return 3;
    }
    // permutation / sign matrices (RGL:1204-1304)
    std::vector<std::vector<unsigned short>> pm(nmaps);
    std::vector<unsigned short*> pmp(nmaps);
    std::vector<float> sign(Kc[0] * N);
    std::vector<std::vector<float>> dist(nmaps);
    std::vector<float*> distp(nmaps);
    for (size_t c = 0; c < nmaps; c++) {
        pm[c].resize(Kc[c] * N);
        pmp[c] = pm[c].data();
        dist[c].resize(Kc[c]);
        distp[c] = dist[c].data();
    }
    if (use_file) {
        std::vector<unsigned short> fp;
        std::vector<float> fs;
        if (!mean && read_bin(dir + "/perm.bin", fp, Kc[0] * N)) pm[0] = fp, pmp[0] = pm[0].data();
        if (mean && read_bin(dir + "/sign.bin", fs, Kc[0] * N)) sign = fs;
    }
    std::vector<float> maps(nmaps * G), pv(nmaps * G);

    BROCCOLI_LIB BROCCOLI(0, 0, 2, false);
    if (!BROCCOLI.GetOpenCLInitiated()) {
        printf("Initialization error is \"%s\" \n", BROCCOLI.GetOpenCLInitializationError().c_str());
        return 4;
    }
    BROCCOLI.SetInputFirstLevelResults(Y.data());
    BROCCOLI.SetInputMNIBrainMask(mask.data());
    BROCCOLI.SetMNIWidth(W);
    BROCCOLI.SetMNIHeight(H);
    BROCCOLI.SetMNIDepth(D);
    BROCCOLI.SetAllocatedHostMemory(0);
    BROCCOLI.SetInferenceMode(inference);
    BROCCOLI.SetClusterDefiningThreshold(cdt);
    BROCCOLI.SetSignificanceLevel(0.05f);
    BROCCOLI.SetNumberOfSubjects(N);
    BROCCOLI.SetNumberOfSubjectsGroup1(n1.data());
    BROCCOLI.SetNumberOfSubjectsGroup2(n2.data());
    BROCCOLI.SetNumberOfPermutations(K);
    BROCCOLI.SetNumberOfGroupPermutations(Kc.data());
    BROCCOLI.SetNumberOfGLMRegressors(p);
    BROCCOLI.SetNumberOfContrasts(nc);
    BROCCOLI.SetDesignMatrix(X.data(), xtxxt.data());
    BROCCOLI.SetContrasts(C.data());
    BROCCOLI.SetGLMScalars(ctxtxc.data());
    BROCCOLI.SetOutputStatisticalMapsMNI(maps.data());
    BROCCOLI.SetOutputPermutationDistributions(distp.data());
    BROCCOLI.SetOutputPValuesMNI(pv.data());
    BROCCOLI.SetDoAllPermutations(false);
    BROCCOLI.SetPermutationFileUsage(use_file != 0);
    BROCCOLI.SetPrint(false);
    BROCCOLI.SetGroupDesigns(designs.data());
    BROCCOLI.SetPermutationPrecision(precision);
    if (mean) {
        BROCCOLI.SetSignMatrix(sign.data());
        BROCCOLI.SetStatisticalTest(2);
        BROCCOLI.PerformMeanSecondLevelPermutationWrapper();
    } else if (!ftest) {
        BROCCOLI.SetPermutationMatrices(pmp.data());
        BROCCOLI.SetStatisticalTest(0);
        BROCCOLI.PerformGLMTTestSecondLevelPermutationWrapper();
    } else {
        BROCCOLI.SetPermutationMatrices(pmp.data());
        BROCCOLI.SetStatisticalTest(1);
        BROCCOLI.PerformGLMFTestSecondLevelPermutationWrapper();
    }
    int errors = 0;
    const int* eb = BROCCOLI.GetOpenCLCreateBufferErrors();
    const int* er = BROCCOLI.GetOpenCLRunKernelErrors();
    for (int i = 0; i < BROCCOLI.GetNumberOfOpenCLKernels(); i++) {
        if (eb[i]) printf("Create buffer error for kernel '%s' is '%s' \n", BROCCOLI.GetOpenCLKernelName(i),
                          BROCCOLI.GetOpenCLErrorMessage(eb[i]));
        if (er[i]) printf("Run kernel error for kernel '%s' is '%s' \n", BROCCOLI.GetOpenCLKernelName(i),
                          BROCCOLI.GetOpenCLErrorMessage(er[i]));
        errors += (eb[i] != 0) + (er[i] != 0);
    }
    if (errors) printf("%s\n", BROCCOLI.GetOpenCLError());
    write_bin(dir + "/maps.bin", maps.data(), maps.size());
    write_bin(dir + "/pvalues.bin", pv.data(), pv.size());
    for (size_t c = 0; c < nmaps; c++)
        write_bin(dir + "/dist" + std::to_string(c) + ".bin", dist[c].data(), dist[c].size());
    write_bin(dir + "/perm_used.bin", pm[0].data(), pm[0].size());
    write_bin(dir + "/sign_used.bin", sign.data(), sign.size());
    long pties = 0, cties = 0, r32 = 0, r64 = 0;
    BROCCOLI.GetPermutationTies(&pties, &cties);
    BROCCOLI.GetPermutationRepeats(&r32, &r64);
    printf("device %s, K %zu, errors %d\n", BROCCOLI.GetOpenCLDeviceName(), Kc[0], errors);
    printf("pvalue ties %ld, cluster tie maps %ld, repeated fp32 %ld fp64 %ld\n", pties, cties, r32, r64);
    return errors ? 1 : 0;
}
