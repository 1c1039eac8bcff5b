// randomise_group_level.cpp -- the RandomiseGroupLevel command line front end
// (code/Bash_Wrapper/RandomiseGroupLevel.cpp) on the MI355X library: the
// same options, input formats (NIfTI volumes, NumRegressors / NumSubjects
// design files, NumRegressors / NumContrasts contrast files, 1-based
// permutation or sign files), checks, messages, outputs
// (<name>_perm_tvalues / _perm_fvalues / _perm_pvalues, permutation values
// and vectors as text) and exit codes, driving the BROCCOLI_LIB drop-in with
// the reference's call sequence (RGL:1414-1546).
//
// Differences:
//  - BROCCOLI_DIR is optional: when set, the build-info files are written
//    under it as RGL:1423-1462 does; the kernels are built ahead of time.
//  - -precision fp32|bf16 (default bf16, the split-BF16 MFMA form) and
//    -connectivity 26|6 (default 26).
//  - Output names get their suffix before the first '.' of the file name
//    part; the reference uses the first '.' of the whole path.
//  - When the library reports an error the outputs are still written, as in
//    the reference, but the exit code is EXIT_FAILURE (the reference exits
//    with EXIT_SUCCESS after printing the error arrays).
#include <sys/stat.h>
#include <sys/time.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "bpg.h"
#include "broccoli_lib.h"
#include "nifti_io.h"

using bpg_cli::Volume;

namespace {

double wall() {
    struct timeval t;
    gettimeofday(&t, nullptr);
    return t.tv_sec + t.tv_usec * 1e-6;
}

void usage() {
    printf("\nThe function performs permutation testing for group analyses.\n\n");
    printf("General usage:\n\n");
    printf("RandomiseGroupLevel volumes.nii -design design.mat -contrasts design.con [options]\n\n");
    printf("Testing a group mean:\n\n");
    printf("RandomiseGroupLevel volumes.nii -groupmean [options]\n\n");
    printf("Options:\n\n");
    printf(" -platform                  The platform to use (default 0, the only one) \n");
    printf(" -device                    The GPU to use (HIP device ordinal, default 0) \n");
    printf(" -design                    The design matrix to apply in each permutation \n");
    printf(" -contrasts                 The contrast vector(s) to apply to the estimated beta values \n");
    printf(" -groupmean                 Test for group mean, using sign flipping (design and contrast not needed) \n");
    printf(" -mask                      A mask that defines which voxels to permute (default none) \n");
    printf(" -permutations              Number of permutations to use (default 5,000) \n");
    printf(" -teststatistics            Test statistics to use, 0 = GLM t-test, 1 = GLM F-test  (default 0) \n");
    printf(" -inferencemode             Inference mode to use, 0 = voxel, 1 = cluster extent, 2 = cluster mass, 3 = TFCE (default 1) \n");
    printf(" -cdt                       Cluster defining threshold for cluster inference (default 2.5) \n");
    printf(" -significance              The significance level to calculate the threshold for (default 0.05) \n");
    printf(" -output                    Set output filename (default volumes_perm_tvalues.nii and volumes_perm_pvalues.nii) \n");
    printf(" -writepermutationvalues    Write all the permutation values to a text file \n");
    printf(" -writepermutations         Write all the random permutations (or sign flips) to a text file \n");
    printf(" -permutationfile           Use a specific permutation file or sign flipping file (e.g. from FSL) \n");
    printf(" -precision                 fp32 or bf16 (split-BF16 MFMA, default) \n");
    printf(" -connectivity              26 (default) or 6 for cluster inference \n");
    printf(" -quiet                     Don't print anything to the terminal (default false) \n");
    printf(" -verbose                   Print extra stuff (default false) \n");
    printf("\n\n");
}

// strtol / strtod with the reference's "must be an integer / a float" test
bool parse_int(const char* s, long& v) {
    char* p;
    v = strtol(s, &p, 10);
    return isspace((unsigned char)*p) || *p == 0;
}
bool parse_float(const char* s, float& v) {
    char* p;
    v = (float)strtod(s, &p);
    return isspace((unsigned char)*p) || *p == 0;
}

// rank of a rows x cols matrix (row-major), Gaussian elimination with full
// pivoting in double, the test FullPivLU applies at RGL:897-907
int matrix_rank(std::vector<double> a, int rows, int cols) {
    double mx = 0.0;
    for (double v : a) mx = std::max(mx, std::fabs(v));
    const double tol = mx * 1e-12 * std::max(rows, cols);
    int rank = 0;
    std::vector<bool> used(rows, false);
    for (int c = 0; c < cols && rank < rows; c++) {
        int piv = -1;
        double best = tol;
        for (int r = 0; r < rows; r++)
            if (!used[r] && std::fabs(a[r * cols + c]) > best) best = std::fabs(a[r * cols + c]), piv = r;
        if (piv < 0) continue;
        used[piv] = true;
        rank++;
        for (int r = 0; r < rows; r++) {
            if (r == piv) continue;
            const double f = a[r * cols + c] / a[piv * cols + c];
            for (int k = c; k < cols; k++) a[r * cols + k] -= f * a[piv * cols + k];
        }
    }
    return rank;
}

bool file_exists(const char* path) {
    FILE* f = fopen(path, "r");
    if (!f) return false;
    fclose(f);
    return true;
}

}  // namespace

int main(int argc, char** argv) {
    int platform = 0, device = 0;
    bool print = true, verbose = false;
    size_t p = 1, nc = 1;
    float cdt = 2.5f, alpha = 0.05f;
    long K = 5000;
    int test = 0, inference = 1, precision = BPG_PREC_SPLIT_BF16, connectivity = 26;
    const char *mask_name = nullptr, *design_file = nullptr, *contrasts_file = nullptr;
    const char *perm_file = nullptr, *values_file = nullptr, *vectors_file = nullptr, *output_name = nullptr;
    bool found_design = false, found_contrasts = false, group_mean = false;

    if (argc == 1) {
        usage();
        return EXIT_SUCCESS;
    }
    {
        std::string ext;
        if (!bpg_cli::extension_ok(argv[1], ext)) {
            printf("File extension is not .nii or .nii.gz, %s is not allowed!\n", ext.c_str());
            return EXIT_FAILURE;
        }
        if (!file_exists(argv[1])) {
            printf("Could not open file %s !\n", argv[1]);
            return EXIT_FAILURE;
        }
    }
    // options (RGL:200-472)
    for (int i = 2; i < argc;) {
        const std::string o = argv[i];
        auto need = [&](const char* what) {
            if (i + 1 >= argc) {
                printf("Unable to read %s after %s !\n", what, o.c_str());
                return false;
            }
            return true;
        };
        long iv;
        float fv;
        if (o == "-platform" || o == "-device") {
            if (!need("value")) return EXIT_FAILURE;
            const bool plat = o == "-platform";
            if (!parse_int(argv[i + 1], iv)) {
                printf("OpenCL %s must be an integer! You provided %s \n", plat ? "platform" : "device", argv[i + 1]);
                return EXIT_FAILURE;
            }
            if (iv < 0) {
                printf("OpenCL %s must be >= 0!\n", plat ? "platform" : "device");
                return EXIT_FAILURE;
            }
            (plat ? platform : device) = (int)iv;
            i += 2;
        } else if (o == "-design") {
            if (!need("value")) return EXIT_FAILURE;
            design_file = argv[i + 1];
            found_design = true;
            i += 2;
        } else if (o == "-contrasts") {
            if (!need("value")) return EXIT_FAILURE;
            contrasts_file = argv[i + 1];
            found_contrasts = true;
            i += 2;
        } else if (o == "-groupmean") {
            group_mean = found_design = found_contrasts = true;
            i += 1;
        } else if (o == "-permutations") {
            if (!need("value")) return EXIT_FAILURE;
            if (!parse_int(argv[i + 1], iv)) {
                printf("Number of permutations must be an integer! You provided %s \n", argv[i + 1]);
                return EXIT_FAILURE;
            }
            if (iv <= 0) {
                printf("Number of permutations must be > 0!\n");
                return EXIT_FAILURE;
            }
            K = iv;
            i += 2;
        } else if (o == "-teststatistics") {
            if (!need("value")) return EXIT_FAILURE;
            if (!parse_int(argv[i + 1], iv)) {
                printf("Test statistics must be an integer! You provided %s \n", argv[i + 1]);
                return EXIT_FAILURE;
            }
            if (iv != 0 && iv != 1) {
                printf("Test statistics must be 0 or 1!\n");
                return EXIT_FAILURE;
            }
            test = (int)iv;
            i += 2;
        } else if (o == "-inferencemode") {
            if (!need("value")) return EXIT_FAILURE;
            if (!parse_int(argv[i + 1], iv)) {
                printf("Inference mode must be an integer! You provided %s \n", argv[i + 1]);
                return EXIT_FAILURE;
            }
            if (iv < 0 || iv > 3) {
                printf("Inference mode must be 0, 1, 2 or 3 !\n");
                return EXIT_FAILURE;
            }
            inference = (int)iv;
            i += 2;
            if (inference == 3) {
                printf("TFCE is currently turned off!\n");
                return EXIT_FAILURE;
            }
        } else if (o == "-cdt") {
            if (!need("value")) return EXIT_FAILURE;
            if (!parse_float(argv[i + 1], fv)) {
                printf("Cluster defining threshold must be a float! You provided %s \n", argv[i + 1]);
                return EXIT_FAILURE;
            }
            cdt = fv;
            i += 2;
        } else if (o == "-significance") {
            if (!need("value")) return EXIT_FAILURE;
            if (!parse_float(argv[i + 1], fv)) {
                printf("Significance level must be a float! You provided %s \n", argv[i + 1]);
                return EXIT_FAILURE;
            }
            if (fv <= 0.0f || fv >= 1.0f) {
                printf("Significance level must be between %f and %f ! You provided %f \n", 0.0f, 1.0f, fv);
                return EXIT_FAILURE;
            }
            alpha = fv;
            i += 2;
        } else if (o == "-mask") {
            if (!need("name")) return EXIT_FAILURE;
            mask_name = argv[i + 1];
            i += 2;
        } else if (o == "-debug") {
            i += 1;
        } else if (o == "-quiet") {
            print = false;
            i += 1;
        } else if (o == "-verbose") {
            verbose = true;
            i += 1;
        } else if (o == "-output") {
            if (!need("name")) return EXIT_FAILURE;
            output_name = argv[i + 1];
            i += 2;
        } else if (o == "-writepermutationvalues") {
            if (!need("name")) return EXIT_FAILURE;
            values_file = argv[i + 1];
            i += 2;
        } else if (o == "-writepermutations") {
            if (!need("name")) return EXIT_FAILURE;
            vectors_file = argv[i + 1];
            i += 2;
        } else if (o == "-permutationfile") {
            if (!need("name")) return EXIT_FAILURE;
            perm_file = argv[i + 1];
            i += 2;
        } else if (o == "-precision") {
            if (!need("value")) return EXIT_FAILURE;
            const std::string v = argv[i + 1];
            if (v == "fp32") precision = BPG_PREC_FP32;
            else if (v == "bf16") precision = BPG_PREC_SPLIT_BF16;
            else {
                printf("Precision must be fp32 or bf16! You provided %s \n", argv[i + 1]);
                return EXIT_FAILURE;
            }
            i += 2;
        } else if (o == "-connectivity") {
            if (!need("value")) return EXIT_FAILURE;
            if (!parse_int(argv[i + 1], iv) || (iv != 6 && iv != 26)) {
                printf("Connectivity must be 6 or 26! You provided %s \n", argv[i + 1]);
                return EXIT_FAILURE;
            }
            connectivity = (int)iv;
            i += 2;
        } else {
            printf("Unrecognized option! %s \n", argv[i]);
            return EXIT_FAILURE;
        }
    }
    if (!found_design) {
        printf("No design file detected, aborting! \n");
        return EXIT_FAILURE;
    }
    if (!found_contrasts) {
        printf("No contrasts file detected, aborting! \n");
        return EXIT_FAILURE;
    }
    if (group_mean && test == 1) {
        printf("Cannot use F-test for group mean, aborting! \n");
        return EXIT_FAILURE;
    }
    const bool ftest = !group_mean && test == 1;

    // volumes (RGL:511-593)
    double t0 = wall();
    Volume data, maskv;
    std::string err;
    if (!bpg_cli::read_volume(argv[1], data, err)) {
        printf("Could not open volumes!\n");
        if (verbose) printf("%s\n", err.c_str());
        return EXIT_FAILURE;
    }
    if (mask_name) {
        std::string ext;
        if (!bpg_cli::extension_ok(mask_name, ext)) {
            printf("File extension is not .nii or .nii.gz, %s is not allowed!\n", ext.c_str());
            return EXIT_FAILURE;
        }
        if (!bpg_cli::read_volume(mask_name, maskv, err)) {
            printf("Could not open mask volume!\n");
            if (verbose) printf("%s\n", err.c_str());
            return EXIT_FAILURE;
        }
    } else {
        printf("\nWarning: No mask being used, doing permutations for all voxels.\n\n");
    }
    if (verbose) printf("It took %f seconds to read the nifti file(s)\n", (float)(wall() - t0));
    const size_t W = data.nx, H = data.ny, D = data.nz, N = data.nt;
    if (N <= 1) {
        printf("Input data is a single volume, nothing to permute! \n");
        return EXIT_FAILURE;
    }
    if (mask_name && (maskv.nx != W || maskv.ny != H || maskv.nz != D)) {
        printf("Input data has the dimensions %zu x %zu %zu, while the mask volume has the dimensions %zu x %zu x %zu. "
               "Aborting! \n", W, H, D, maskv.nx, maskv.ny, maskv.nz);
        return EXIT_FAILURE;
    }

    // design and contrast files (RGL:599-741, RGL:815-895)
    std::ifstream design, contrasts;
    if (!group_mean) {
        design.open(design_file);
        if (!design.good()) {
            printf("Unable to open design file %s. Aborting! \n", design_file);
            return EXIT_FAILURE;
        }
        std::string word;
        long n;
        design >> word;
        if (word != "NumRegressors") {
            printf("First element of the design file should be the string 'NumRegressors', but it is %s. Aborting! \n",
                   word.c_str());
            return EXIT_FAILURE;
        }
        design >> n;
        if (n <= 0) {
            printf("Number of regressors must be > 0 ! You provided %ld regressors in the design file. Aborting! \n", n);
            return EXIT_FAILURE;
        }
        if (n > BPG_MAX_REGRESSORS) {
            printf("Number of regressors must be <= 25 ! You provided %ld regressors in the design file. Aborting! \n", n);
            return EXIT_FAILURE;
        }
        if ((size_t)n > N) {
            printf("Number of regressors must be smaller or equal to the number of subjects! You provided %ld regressors "
                   "in the design file and there are %zu subjects. Aborting! \n", n, N);
            return EXIT_FAILURE;
        }
        p = (size_t)n;
        design >> word;
        if (word != "NumSubjects") {
            printf("Third element of the design file should be the string 'NumSubjects', but it is %s. Aborting! \n",
                   word.c_str());
            return EXIT_FAILURE;
        }
        design >> n;
        if (n <= 0) {
            printf("Number of subjects must be > 0 ! You provided %ld in the design file. Aborting! \n", n);
            return EXIT_FAILURE;
        }
        if ((size_t)n != N) {
            printf("Input data contains %zu volumes, while the design file says %ld subjects. Aborting! \n", N, n);
            return EXIT_FAILURE;
        }
        contrasts.open(contrasts_file);
        if (!contrasts.good()) {
            printf("Unable to open contrasts file %s. Aborting! \n", contrasts_file);
            return EXIT_FAILURE;
        }
        contrasts >> word;
        if (word != "NumRegressors") {
            printf("First element of the contrasts file should be the string 'NumRegressors', but it is %s. Aborting! \n",
                   word.c_str());
            return EXIT_FAILURE;
        }
        contrasts >> n;
        if ((size_t)n != p) {
            printf("Design file says that number of regressors is %zu, while contrast file says there are %ld regressors. "
                   "Aborting! \n", p, n);
            return EXIT_FAILURE;
        }
        contrasts >> word;
        if (word != "NumContrasts") {
            printf("Third element of the contrasts file should be the string 'NumContrasts', but it is %s. Aborting! \n",
                   word.c_str());
            return EXIT_FAILURE;
        }
        contrasts >> n;
        if (n <= 0) {
            printf("Number of contrasts must be > 0 ! You provided %ld in the contrasts file. Aborting! \n", n);
            return EXIT_FAILURE;
        }
        if (ftest && n > BPG_MAX_FROWS) {
            printf("Number of contrasts must currently be <= 10 for F-test! You provided %ld contrasts in the contrasts "
                   "file. Aborting! \n", n);
            return EXIT_FAILURE;
        }
        nc = (size_t)n;
    }
    if (print) {
        printf("Data size: %zu x %zu x %zu x %zu \n", W, H, D, N);
        printf("Number of permutations: %ld \n", K);
        printf("Number of regressors: %zu \n", p);
        printf("Number of contrasts: %zu \n", nc);
        if (group_mean) {
        } else if (!ftest) {
            printf("Performing %zu t-tests \n", nc);
        } else {
            printf("Performing one F-test \n");
        }
    }
    if (verbose) printf("Using a cluster defining threshold of %f \n", cdt);

    const size_t G = W * H * D;
    std::vector<float> X(N * p), C(p * nc);
    if (!group_mean) {
        for (size_t s = 0; s < N; s++)
            for (size_t r = 0; r < p; r++)
                if (!(design >> X[N * r + s])) {
                    printf("Could not read all values of the design file %s, aborting! Please check if the number of "
                           "regressors and subjects are correct. \n", design_file);
                    return EXIT_FAILURE;
                }
        for (size_t c = 0; c < nc; c++)
            for (size_t r = 0; r < p; r++)
                if (!(contrasts >> C[r + c * p])) {
                    printf("Could not read all values of the contrasts file %s, aborting! Please check if the number of "
                           "regressors and contrasts are correct. \n", contrasts_file);
                    return EXIT_FAILURE;
                }
    } else {
        for (size_t s = 0; s < N; s++) X[s] = 1.0f;
        C[0] = 1.0f;
    }
    if (ftest) {
        std::vector<double> cm(nc * p);
        for (size_t c = 0; c < nc; c++)
            for (size_t r = 0; r < p; r++) cm[c * p + r] = C[r + c * p];
        if (matrix_rank(cm, (int)nc, (int)p) < (int)nc) {
            printf("Contrast matrix does not have full rank, at least one contrast can be written as a linear "
                   "combination of other contrasts, not OK for F-test, aborting!\n");
            return EXIT_FAILURE;
        }
    }

    // design detection and permutation counts (RGL:908-1134)
    const size_t nmaps = group_mean ? 1 : (ftest ? 1 : nc);
    std::vector<int> designs(std::max<size_t>(nc, 1), 100), n1(std::max<size_t>(nc, 1), 0),
        n2(std::max<size_t>(nc, 1), 0);
    std::vector<size_t> Kc(nmaps, (size_t)K);
    if (!group_mean) {
        bpg_detect_designs((int)N, (int)p, (int)nc, X.data(), C.data(), ftest ? 1 : 0, designs.data(), n1.data(),
                           n2.data());
        for (size_t c = 0; c < nmaps; c++) {
            if (designs[c] == BPG_TWOSAMPLE) {
                if (!ftest)
                    printf("Two sample t-test design detected for t-contrast %zu, %i subjects in group 1 and %i subjects "
                           "in group 2\n", c + 1, n1[c], n2[c]);
                else
                    printf("\nTwo sample t-test design detected for F-test, %i subjects in group 1 and %i subjects in "
                           "group 2\n\n", n1[0], n2[0]);
            } else if (designs[c] == BPG_CORRELATION) {
                if (!ftest) printf("Correlation design detected for t-contrast %zu\n", c + 1);
                else printf("\nCorrelation design detected for F-test\n\n");
            } else if (designs[c] == BPG_MEAN_DESIGN && !ftest) {
                printf("Mean design detected for t-contrast %zu\n", c + 1);
            }
        }
        for (size_t c = 0; c < nmaps; c++) {
            if (designs[c] == BPG_MEAN_DESIGN) {
                printf("Warning: Contrast %zu leads to a simple mean value, only doing 1 permutation!\n", c + 1);
                Kc[c] = 1;
                continue;
            }
            if (designs[c] != BPG_TWOSAMPLE && designs[c] != BPG_CORRELATION) continue;
            const double mx = bpg_max_permutations(ftest ? 1 : 0, designs[c], (int)N, n2[c]);
            if ((double)K > mx) {
                printf("Warning: Number of possible permutations for your design is %g for contrast %zu, but %g "
                       "permutations were requested. Lowering number of permutations to number of possible "
                       "permutations. \n", mx, c + 1, (double)K);
                Kc[c] = (size_t)mx;
            } else {
                printf("Max number of permutations for contrast %zu is %g \n", c + 1, mx);
            }
        }
    } else {
        const double mx = bpg_max_permutations(2, 0, (int)N, 0);
        if ((double)K > mx) {
            printf("Warning: Number of possible sign flips for group mean is %g, but %g permutations were requested. "
                   "Lowering number of permutations to number of possible sign flips. \n", mx, (double)K);
            Kc[0] = (size_t)mx;
        } else {
            printf("Max number of sign flips is %g \n", mx);
        }
    }
    const bool do_all = [&] {
        for (size_t c = 0; c < nmaps; c++) {
            const double mx = group_mean ? bpg_max_permutations(2, 0, (int)N, 0)
                                         : bpg_max_permutations(ftest ? 1 : 0, designs[c], (int)N, n2[c]);
            if ((double)K >= mx) return true;
        }
        return false;
    }();

    // FP64 model precompute cast to float (RGL:1137-1194)
    std::vector<float> xtxxt(N * p), ctxtxc(ftest ? nc * nc : nc);
    if (bpg_model_precompute((int)N, (int)p, (int)nc, ftest ? 1 : 0, X.data(), C.data(), xtxxt.data(),
                             ctxtxc.data()) != BPG_OK) {
        printf("The design matrix is rank deficient, aborting!\n");
        return EXIT_FAILURE;
    }
    if (verbose) {
        printf("Design matrix is \n");
        for (size_t s = 0; s < N; s++) {
            for (size_t r = 0; r < p; r++) printf(" %f ", X[s + r * N]);
            printf("\n");
        }
        for (size_t c = 0; c < nc; c++) {
            printf("Contrast is ");
            for (size_t r = 0; r < p; r++) printf(" %f ", C[r + c * p]);
            printf("\n");
        }
    }

    // permutation / sign matrices (RGL:1214-1304)
    std::vector<std::vector<unsigned short>> pm(nmaps);
    std::vector<unsigned short*> pmp(nmaps);
    std::vector<std::vector<float>> dist(nmaps);
    std::vector<float*> distp(nmaps);
    for (size_t c = 0; c < nmaps; c++) {
        pm[c].assign(Kc[c] * N, 0);
        pmp[c] = pm[c].data();
        dist[c].assign(Kc[c], 0.0f);
        distp[c] = dist[c].data();
    }
    std::vector<float> sign(Kc[0] * N, 1.0f);
    if (perm_file) {
        std::ifstream pf(perm_file);
        if (!pf.good()) {
            printf("Could not open %s file %s, aborting! \n", group_mean ? "sign" : "permutation", perm_file);
            return EXIT_FAILURE;
        }
        for (size_t k = 0; k < Kc[0]; k++)
            for (size_t s = 0; s < N; s++) {
                float v;
                if (!(pf >> v)) {
                    printf("Could not read all values of the permutation file %s, aborting! \n", perm_file);
                    return EXIT_FAILURE;
                }
                if (group_mean) sign[s + k * N] = v;
                else pm[0][s + k * N] = (unsigned short)((unsigned short)v - 1);
            }
    }

    // volumes as float, mask (RGL:1306-1402)
    std::vector<float> mask(G, 1.0f);
    if (mask_name) mask = maskv.data;

    t0 = wall();
    BROCCOLI_LIB B((cl_uint)platform, (cl_uint)device, 2, verbose);
    if (verbose) printf("It took %f seconds to initiate BROCCOLI\n", (float)(wall() - t0));
    if (const char* dir = getenv("BROCCOLI_DIR")) {
        const std::vector<std::string> info = B.GetOpenCLBuildInfo();
        const std::vector<std::string> names = B.GetKernelFileNames();
        for (int k = 0; k < B.GetNumberOfKernelFiles(); k++) {
            std::string name = names[k];
            name = name.substr(0, name.size() - 4).substr(6);
            const std::string path = std::string(dir) + "compiled/Kernels/buildInfo_" + B.GetOpenCLPlatformName() +
                                     "_" + B.GetOpenCLDeviceName() + "_" + name + ".txt";
            FILE* f = fopen(path.c_str(), "w");
            if (!f) {
                printf("Could not open %s for writing ! \n", path.c_str());
                continue;
            }
            if (fputs(info[k].c_str(), f) == EOF) printf("Could not write to %s ! \n", path.c_str());
            fclose(f);
        }
    }
    if (!B.GetOpenCLInitiated()) {
        printf("Initialization error is \"%s\" \n", B.GetOpenCLInitializationError().c_str());
        printf("OpenCL error is \"%s\" \n", B.GetOpenCLError());
        printf("OpenCL initialization failed, aborting! \n");
        return EXIT_FAILURE;
    }
    std::vector<float> maps(nmaps * G, 0.0f), pv(nmaps * G, 0.0f);
    B.SetInputFirstLevelResults(data.data.data());
    B.SetInputMNIBrainMask(mask.data());
    B.SetMNIWidth(W);
    B.SetMNIHeight(H);
    B.SetMNIDepth(D);
    B.SetAllocatedHostMemory(0);
    B.SetInferenceMode(inference);
    B.SetClusterDefiningThreshold(cdt);
    B.SetSignificanceLevel(alpha);
    B.SetNumberOfSubjects(N);
    B.SetNumberOfSubjectsGroup1(n1.data());
    B.SetNumberOfSubjectsGroup2(n2.data());
    B.SetNumberOfPermutations((size_t)K);
    B.SetNumberOfGroupPermutations(Kc.data());
    B.SetNumberOfGLMRegressors(p);
    B.SetNumberOfContrasts(nc);
    B.SetDesignMatrix(X.data(), xtxxt.data());
    B.SetContrasts(C.data());
    B.SetGLMScalars(ctxtxc.data());
    B.SetOutputStatisticalMapsMNI(maps.data());
    B.SetOutputPermutationDistributions(distp.data());
    B.SetOutputPValuesMNI(pv.data());
    B.SetDoAllPermutations(do_all);
    B.SetPermutationFileUsage(perm_file != nullptr);
    B.SetPrint(print);
    B.SetGroupDesigns(designs.data());
    B.SetPermutationPrecision(precision);
    B.SetClusterConnectivity(connectivity);
    t0 = wall();
    if (group_mean) {
        B.SetSignMatrix(sign.data());
        B.SetStatisticalTest(2);
        B.PerformMeanSecondLevelPermutationWrapper();
    } else if (!ftest) {
        B.SetPermutationMatrices(pmp.data());
        B.SetStatisticalTest(0);
        B.PerformGLMTTestSecondLevelPermutationWrapper();
    } else {
        B.SetPermutationMatrices(pmp.data());
        B.SetStatisticalTest(1);
        B.PerformGLMFTestSecondLevelPermutationWrapper();
    }
    if (verbose) {
        printf("\nIt took %f seconds to run the permutation test\n", (float)(wall() - t0));
        long pties = 0, cties = 0, r32 = 0, r64 = 0;
        B.GetPermutationTies(&pties, &cties);
        B.GetPermutationRepeats(&r32, &r64);
        printf("Permutation p-value comparisons inside the 1e-4 error band (ties): %ld\n", pties);
        if (inference != 0 && inference != 3)
            printf("Permutations with a statistic within the error band of the cluster defining threshold: %ld\n",
                   cties);
        printf("Permutations repeated for ill-conditioned samples: %ld in FP32, %ld in FP64\n", r32, r64);
    }
    bool failed = false;
    {
        const int* e = B.GetOpenCLCreateBufferErrors();
        for (int i = 0; i < B.GetNumberOfOpenCLKernels(); i++)
            if (e[i] != 0) printf("Create buffer error %i is %s \n", i, B.GetOpenCLErrorMessage(e[i])), failed = true;
        e = B.GetOpenCLCreateKernelErrors();
        for (int i = 0; i < B.GetNumberOfOpenCLKernels(); i++)
            if (e[i] != 0)
                printf("Create kernel error for kernel '%s' is '%s' \n", B.GetOpenCLKernelName(i),
                       B.GetOpenCLErrorMessage(e[i])), failed = true;
        e = B.GetOpenCLRunKernelErrors();
        for (int i = 0; i < B.GetNumberOfOpenCLKernels(); i++)
            if (e[i] != 0)
                printf("Run kernel error for kernel '%s' is '%s' \n", B.GetOpenCLKernelName(i),
                       B.GetOpenCLErrorMessage(e[i])), failed = true;
    }

    // text outputs (RGL:1588-1648)
    if (values_file) {
        for (size_t c = 0; c < nmaps; c++) {
            std::string name = values_file;
            const size_t dot = bpg_cli::first_dot(name);
            name.insert(dot == std::string::npos ? name.size() : dot, std::to_string(c + 1));
            std::ofstream f(name);
            if (!f.good()) {
                printf("Could not open %s for writing permutation values!\n", name.c_str());
                continue;
            }
            for (size_t k = 0; k < Kc[c]; k++) f << std::setprecision(6) << std::fixed << (double)dist[c][k] << " " << std::endl;
        }
    }
    if (vectors_file) {
        std::ofstream f(vectors_file);
        if (!f.good()) {
            printf("Could not open %s for writing permutation vectors!\n", vectors_file);
        } else {
            for (size_t k = 0; k < Kc[0]; k++) {
                for (size_t s = 0; s < N; s++) {
                    const double v = group_mean ? (double)sign[s + k * N] : (double)(pm[0][s + k * N] + 1);
                    f << std::setprecision(6) << std::fixed << v << " ";
                }
                f << std::endl;
            }
        }
    }

    // NIfTI outputs (RGL:1650-1703): nc volumes for t, one for F and the mean
    std::string base = output_name ? output_name : argv[1];
    if (output_name) {
        std::string ext;
        if (!bpg_cli::extension_ok(base, ext)) base += ".nii";
    }
    t0 = wall();
    const std::string stat_suffix = ftest ? "_perm_fvalues" : "_perm_tvalues";
    if (!bpg_cli::write_volume(bpg_cli::insert_suffix(base, stat_suffix), data, nmaps, maps.data(), err) ||
        !bpg_cli::write_volume(bpg_cli::insert_suffix(base, "_perm_pvalues"), data, nmaps, pv.data(), err)) {
        printf("%s\n", err.c_str());
        return EXIT_FAILURE;
    }
    if (verbose) printf("It took %f seconds to write the nifti file(s)\n", (float)(wall() - t0));
    return failed ? EXIT_FAILURE : EXIT_SUCCESS;
}
