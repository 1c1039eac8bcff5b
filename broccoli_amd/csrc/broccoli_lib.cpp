// broccoli_lib.cpp -- the BROCCOLI_LIB drop-in for the second-level
// permutation test (include/broccoli_lib.h), on the bpg C ABI.
#include "broccoli_lib.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "bpg.h"

namespace {

// Error-array slots of the reference kernels this path replaces
// (code/BROCCOLI_LIB/broccoli_lib.cpp:2319-2420).
constexpr int kSlotMaxAtomic = 26;
constexpr int kSlotSetStartClusterIndices = 63;
constexpr int kSlotPValuesVoxel = 70;
constexpr int kSlotPValuesExtent = 71;
constexpr int kSlotPValuesMass = 72;
constexpr int kSlotBetaWeightsGLM = 73;
constexpr int kSlotStatMapsGLMTTest = 85;
constexpr int kSlotStatMapsGLMFTest = 86;
constexpr int kSlotTTestPermutation = 89;
constexpr int kSlotFTestPermutation = 90;
constexpr int kSlotMeanPermutation = 91;
constexpr int kSlotTransformData = 92;

double ms_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

BROCCOLI_LIB::BROCCOLI_LIB() { Init(0, 0); }
BROCCOLI_LIB::BROCCOLI_LIB(cl_uint platform, cl_uint device) { Init(platform, device); }
BROCCOLI_LIB::BROCCOLI_LIB(cl_uint platform, cl_uint device, const char*) { Init(platform, device); }
BROCCOLI_LIB::BROCCOLI_LIB(cl_uint platform, cl_uint device, int wrapper, bool verbos) {
    wrapper_ = wrapper;
    verbose_ = verbos;
    Init(platform, device);
}

BROCCOLI_LIB::~BROCCOLI_LIB() {
    if (ctx_) bpg_destroy(ctx_);
}

void BROCCOLI_LIB::Init(cl_uint platform, cl_uint device) {
    std::fill(create_kernel_errors_, create_kernel_errors_ + NUMBER_OF_OPENCL_KERNELS, 0);
    std::fill(create_buffer_errors_, create_buffer_errors_ + NUMBER_OF_OPENCL_KERNELS, 0);
    std::fill(run_kernel_errors_, run_kernel_errors_ + NUMBER_OF_OPENCL_KERNELS, 0);
    // one platform (HIP); an invalid index fails like LIB:1198-1202
    if (platform != 0) {
        init_error_ = "Invalid platform index, the only platform is AMD HIP (index 0)";
        return;
    }
    const int rc = bpg_create((int)device, &ctx_);
    if (rc != BPG_OK) {
        init_error_ = rc == BPG_ERR_NODEVICE ? "Invalid device index, or no HIP device present"
                      : rc == BPG_ERR_UNSUPPORTED ? "The HIP device is not an MI355X (gfx950)"
                                                  : "HIP initialisation failed";
        ctx_ = nullptr;
        return;
    }
    device_name_ = bpg_device_name(ctx_);
    initiated_ = true;
}

void BROCCOLI_LIB::SetPrint(bool print) { print_ = print; }
void BROCCOLI_LIB::SetVerbose(bool verbos) { verbose_ = verbos; }
void BROCCOLI_LIB::SetWrapper(int wrapper) { wrapper_ = wrapper; }
void BROCCOLI_LIB::SetAllocatedHostMemory(size_t allocated) { allocated_ = allocated; }
void BROCCOLI_LIB::SetStatisticalTest(int test) { test_ = test; }
void BROCCOLI_LIB::SetGroupDesigns(int* designs) { designs_ = designs; }
void BROCCOLI_LIB::SetInferenceMode(int mode) { inference_ = mode; }
void BROCCOLI_LIB::SetClusterDefiningThreshold(float threshold) { cdt_ = threshold; }
void BROCCOLI_LIB::SetPermutationMatrices(unsigned short int** matrices) { perms_ = matrices; }
void BROCCOLI_LIB::SetSignMatrix(float* signs) { signs_ = signs; }
void BROCCOLI_LIB::SetPermutationFileUsage(bool use) { use_file_ = use; }
void BROCCOLI_LIB::SetDoAllPermutations(bool all) { do_all_ = all; }
void BROCCOLI_LIB::SetSignificanceLevel(float value) { alpha_ = value; }
void BROCCOLI_LIB::SetNumberOfGLMRegressors(size_t NR) { p_ = NR; }
void BROCCOLI_LIB::SetNumberOfContrasts(size_t NC) { nc_ = NC; }
void BROCCOLI_LIB::SetDesignMatrix(float* X_GLM, float* xtxxt_GLM) {
    X_ = X_GLM;
    xtxxt_ = xtxxt_GLM;
}
void BROCCOLI_LIB::SetContrasts(float* contrasts) { C_ = contrasts; }
void BROCCOLI_LIB::SetGLMScalars(float* ctxtxc) { ctxtxc_ = ctxtxc; }
void BROCCOLI_LIB::SetNumberOfPermutations(size_t K) { K_all_ = K; }
void BROCCOLI_LIB::SetNumberOfGroupPermutations(size_t* K) { K_ = K; }
void BROCCOLI_LIB::SetMNIWidth(size_t w) { W_ = w; }
void BROCCOLI_LIB::SetMNIHeight(size_t h) { H_ = h; }
void BROCCOLI_LIB::SetMNIDepth(size_t d) { D_ = d; }
void BROCCOLI_LIB::SetInputMNIBrainMask(float* input) { h_mask_ = input; }
void BROCCOLI_LIB::SetInputFirstLevelResults(float* input) { h_Y_ = input; }
void BROCCOLI_LIB::SetNumberOfSubjects(size_t N) { N_ = N; }
void BROCCOLI_LIB::SetNumberOfSubjectsGroup1(int* N) { n1_ = N; }
void BROCCOLI_LIB::SetNumberOfSubjectsGroup2(int* N) { n2_ = N; }
void BROCCOLI_LIB::SetOutputStatisticalMapsMNI(float* output) { out_maps_ = output; }
void BROCCOLI_LIB::SetOutputPValuesMNI(float* output) { out_p_ = output; }
void BROCCOLI_LIB::SetOutputPermutationDistributions(float** output) { out_dist_ = output; }
void BROCCOLI_LIB::SetPermutationPrecision(int precision) { precision_ = precision; }
void BROCCOLI_LIB::SetClusterConnectivity(int connectivity) { connectivity_ = connectivity; }

void BROCCOLI_LIB::GetPermutationPhaseTimes(double* permutation_ms, double* observed_ms, double* pvalues_ms) {
    if (permutation_ms) *permutation_ms = t_perm_;
    if (observed_ms) *observed_ms = t_obs_;
    if (pvalues_ms) *pvalues_ms = t_pval_;
}

void BROCCOLI_LIB::GetPermutationTies(long* pvalue_ties, long* cluster_tie_maps) {
    if (pvalue_ties) *pvalue_ties = pvalue_ties_;
    if (cluster_tie_maps) *cluster_tie_maps = cdt_tie_maps_;
}

void BROCCOLI_LIB::GetPermutationRepeats(long* fp32_permutations, long* fp64_permutations) {
    if (fp32_permutations) *fp32_permutations = fp32_perms_;
    if (fp64_permutations) *fp64_permutations = fp64_perms_;
}

void BROCCOLI_LIB::RecordError(int slot, int status, bool buffer) {
    if (status == BPG_OK) return;
    if (buffer || status == BPG_ERR_NOMEM) create_buffer_errors_[slot] = status;
    else run_kernel_errors_[slot] = status;
    if (ctx_) last_error_ = bpg_last_error(ctx_);
}

void BROCCOLI_LIB::SetOutputBetaVolumesMNI(float* output) { out_betas_ = output; }
void BROCCOLI_LIB::SetOutputResidualsMNI(float* output) { out_residuals_ = output; }
void BROCCOLI_LIB::SetOutputResidualVariances(float* output) { out_resvar_ = output; }
void BROCCOLI_LIB::SetSaveResidualsMNI(bool value) { save_residuals_ = value; }
void BROCCOLI_LIB::SetSaveResidualVariances(bool value) { save_resvar_ = value; }
void BROCCOLI_LIB::SetBetasOnly(bool value) { betas_only_ = value; }
void BROCCOLI_LIB::SetContrastsOnly(bool value) { contrasts_only_ = value; }
void BROCCOLI_LIB::SetBetasAndContrastsOnly(bool value) { betas_and_contrasts_only_ = value; }

void BROCCOLI_LIB::PerformGLMTTestSecondLevelWrapper() { PerformGLM(BPG_TTEST); }
void BROCCOLI_LIB::PerformGLMFTestSecondLevelWrapper() { PerformGLM(BPG_FTEST); }
void BROCCOLI_LIB::PerformMeanSecondLevelPermutationWrapper() { Perform(BPG_GROUP_MEAN); }
void BROCCOLI_LIB::PerformGLMTTestSecondLevelPermutationWrapper() { Perform(BPG_TTEST); }
void BROCCOLI_LIB::PerformGLMFTestSecondLevelPermutationWrapper() { Perform(BPG_FTEST); }

// Perform*SecondLevelPermutationWrapper (LIB:13144-13376): for each
// statistical map, the permutation (or sign) matrix unless one was given
// (LIB:16603-16647), the permutation phase, the alpha threshold print
// (LIB:16765-16786); then the observed maps (LIB:15858-15950) and the
// permutation p-values (LIB:16794-16873).
void BROCCOLI_LIB::Perform(int test) {
    const int slot_perm = test == BPG_TTEST ? kSlotTTestPermutation
                          : test == BPG_FTEST ? kSlotFTestPermutation
                                              : kSlotMeanPermutation;
    pvalue_ties_ = cdt_tie_maps_ = fp32_perms_ = fp64_perms_ = 0;
    if (!initiated_ || !ctx_) {
        run_kernel_errors_[slot_perm] = BPG_ERR_NODEVICE;
        return;
    }
    test_ = test == BPG_GROUP_MEAN ? 2 : test;
    const int p = test == BPG_GROUP_MEAN ? 1 : (int)p_;
    const int nc = test == BPG_GROUP_MEAN ? 1 : (int)nc_;
    const int nmaps = test == BPG_TTEST ? nc : 1;
    const int N = (int)N_;
    float one = 1.0f;

    bpg_study_desc d{};
    d.test = test;
    d.inference = inference_;
    d.precision = precision_;
    d.connectivity = connectivity_;
    d.W = (int)W_;
    d.H = (int)H_;
    d.D = (int)D_;
    d.N = N;
    d.p = p;
    d.nc = nc;
    d.Y = h_Y_;
    d.mask = h_mask_;
    d.X = X_;
    d.C = test == BPG_GROUP_MEAN ? (C_ ? C_ : &one) : C_;
    d.cdt = cdt_;

    const auto t0 = std::chrono::steady_clock::now();
    bpg_plan* plan = nullptr;
    int rc = bpg_plan_create(ctx_, &d, &plan);
    if (rc != BPG_OK) {
        RecordError(slot_perm, rc, rc == BPG_ERR_NOMEM);
        return;
    }
    std::vector<long> Ks(nmaps);
    std::vector<std::vector<float>> dists(nmaps);
    for (int c = 0; c < nmaps; c++) {
        const size_t K = K_ ? K_[c] : K_all_;
        Ks[c] = (long)K;
        if (test == BPG_GROUP_MEAN) {
            if (!use_file_ && signs_) {
                rc = bpg_gen_sign_matrix(N, (long)K, 1, 0, signs_);
                if (rc != BPG_OK) { RecordError(slot_perm, rc, false); bpg_plan_destroy(plan); return; }
            }
        } else if (!use_file_ && perms_ && perms_[c]) {
            const int design = designs_ ? designs_[c] : BPG_CORRELATION;
            if (design == BPG_TWOSAMPLE)
                rc = bpg_gen_perm_two_sample(N, n1_ ? n1_[c] : 0, n2_ ? n2_[c] : N, (long)K, 1, 0, perms_[c]);
            else if (design == BPG_CORRELATION)
                rc = bpg_gen_perm_correlation(N, (long)K, 1, 0, perms_[c]);
            else
                for (int s = 0; s < N; s++) perms_[c][s] = (unsigned short)s;
            if (rc != BPG_OK) { RecordError(slot_perm, rc, false); bpg_plan_destroy(plan); return; }
        }
        dists[c].resize(K);
        float* dist = dists[c].data();
        rc = bpg_plan_permutation_phase(plan, c, test == BPG_GROUP_MEAN ? nullptr : perms_[c],
                                        test == BPG_GROUP_MEAN ? signs_ : nullptr, (long)K, dist, nullptr);
        if (rc != BPG_OK) {
            RecordError(slot_perm, rc, rc == BPG_ERR_NOMEM);
            if (inference_ == 0) RecordError(kSlotMaxAtomic, rc, false);
            else RecordError(kSlotSetStartClusterIndices, rc, false);
            bpg_plan_destroy(plan);
            return;
        }
        if (out_dist_ && out_dist_[c]) std::copy(dist, dist + K, out_dist_[c]);
        {
            bpg_timing tm{};
            if (bpg_plan_last_timing(plan, &tm) == BPG_OK) {
                cdt_tie_maps_ += tm.cdt_tie_maps;
                fp32_perms_ += tm.precision_used == BPG_PREC_FP32 ? 0 : tm.fp32_permutations;
                fp64_perms_ += tm.fp64_permutations;
            }
        }
        if (wrapper_ == 2) {
            std::vector<float> srt(dist, dist + K);
            std::sort(srt.begin(), srt.end());
            const int idx = (int)std::ceil((1.0f - alpha_) * (float)K) - 1;
            const float thr = srt[std::min<int>(std::max(idx, 0), (int)K - 1)];
            if (test == BPG_TTEST)
                printf("Permutation threshold for contrast %d for a significance level of %f is %f \n", c + 1,
                       alpha_, thr);
            else if (test == BPG_FTEST)
                printf("Permutation threshold for F-test for a significance level of %f is %f \n", alpha_, thr);
            else
                printf("Permutation threshold for group mean for a significance level of %f is %f \n", alpha_, thr);
        }
    }
    t_perm_ = ms_since(t0);
    (void)kSlotTransformData;
    const auto t1 = std::chrono::steady_clock::now();
    std::vector<float> maps;
    float* mout = out_maps_;
    if (!mout) {
        maps.resize((size_t)nmaps * W_ * H_ * D_);
        mout = maps.data();
    }
    rc = bpg_plan_observed_maps(plan, mout);
    if (rc != BPG_OK) {
        RecordError(kSlotBetaWeightsGLM, rc, rc == BPG_ERR_NOMEM);
        RecordError(test == BPG_FTEST ? kSlotStatMapsGLMFTest : kSlotStatMapsGLMTTest, rc, false);
        bpg_plan_destroy(plan);
        return;
    }
    t_obs_ = ms_since(t1);
    const auto t2 = std::chrono::steady_clock::now();
    if (out_p_) {
        std::vector<const float*> md(nmaps);
        for (int c = 0; c < nmaps; c++) md[c] = dists[c].data();
        rc = bpg_plan_pvalues(plan, Ks.data(), md.data(), out_p_, nullptr);
        bpg_timing tm{};
        if (rc == BPG_OK && bpg_plan_last_timing(plan, &tm) == BPG_OK) pvalue_ties_ = tm.pvalue_ties;
        if (rc != BPG_OK) {
            const int slot = inference_ == 0 ? kSlotPValuesVoxel
                             : inference_ == 1 ? kSlotPValuesExtent
                                               : kSlotPValuesMass;
            RecordError(slot, rc, rc == BPG_ERR_NOMEM);
        }
    }
    t_pval_ = ms_since(t2);
    bpg_plan_destroy(plan);
}

// Perform{GLMTTest,GLMFTest}SecondLevelWrapper (LIB:12872-13086): betas
// always, statistical maps unless one of the *Only flags is set, residual
// variances and residuals when asked for (LIB:12950-12962).
void BROCCOLI_LIB::PerformGLM(int test) {
    const int slot = test == BPG_FTEST ? kSlotStatMapsGLMFTest : kSlotStatMapsGLMTTest;
    if (!initiated_ || !ctx_) {
        run_kernel_errors_[slot] = BPG_ERR_NODEVICE;
        return;
    }
    test_ = test;
    float one = 1.0f;
    bpg_study_desc d{};
    d.test = test;
    d.inference = 0;
    d.precision = precision_;
    d.connectivity = connectivity_;
    d.W = (int)W_;
    d.H = (int)H_;
    d.D = (int)D_;
    d.N = (int)N_;
    d.p = (int)p_;
    d.nc = (int)nc_;
    d.Y = h_Y_;
    d.mask = h_mask_;
    d.X = X_;
    d.C = C_ ? C_ : &one;
    d.cdt = cdt_;
    bpg_plan* plan = nullptr;
    int rc = bpg_plan_create(ctx_, &d, &plan);
    if (rc != BPG_OK) {
        RecordError(kSlotBetaWeightsGLM, rc, rc == BPG_ERR_NOMEM);
        return;
    }
    const bool maps = !betas_only_ && !contrasts_only_ && !betas_and_contrasts_only_;
    rc = bpg_plan_glm(plan, out_betas_, maps ? out_maps_ : nullptr, save_residuals_ ? out_residuals_ : nullptr,
                      save_resvar_ ? out_resvar_ : nullptr);
    if (rc != BPG_OK) {
        RecordError(kSlotBetaWeightsGLM, rc, rc == BPG_ERR_NOMEM);
        RecordError(slot, rc, false);
    }
    bpg_plan_destroy(plan);
}

bool BROCCOLI_LIB::GetOpenCLInitiated() { return initiated_; }
int BROCCOLI_LIB::GetNumberOfOpenCLKernels() { return NUMBER_OF_OPENCL_KERNELS; }

std::vector<std::string> BROCCOLI_LIB::GetKernelFileNames() {
    // the CLI strips "kernel" and ".cpp" from each name (RGL:1431-1443)
    return {"kernelPermutationGLM.cpp"};
}
int BROCCOLI_LIB::GetNumberOfKernelFiles() { return 1; }

std::vector<std::string> BROCCOLI_LIB::GetOpenCLBuildInfo() {
    return {"Ahead-of-time build for gfx950 (hipcc); no runtime compilation."};
}

const char* BROCCOLI_LIB::GetOpenCLErrorMessage(int error) {
    switch (error) {
        case BPG_OK: return "CL_SUCCESS";
        case BPG_ERR_ARG: return "BPG_ERR_ARG (invalid argument or shape)";
        case BPG_ERR_HIP: return "BPG_ERR_HIP (HIP runtime failure)";
        case BPG_ERR_NOMEM: return "BPG_ERR_NOMEM (allocation failed)";
        case BPG_ERR_DESIGN: return "BPG_ERR_DESIGN (rank-deficient design or contrast)";
        case BPG_ERR_UNSUPPORTED: return "BPG_ERR_UNSUPPORTED (outside the built kernels)";
        case BPG_ERR_NODEVICE: return "BPG_ERR_NODEVICE (no usable HIP device)";
    }
    return "Unrecognized error code";
}

const char* BROCCOLI_LIB::GetOpenCLKernelName(int kernel) {
    switch (kernel) {
        case kSlotMaxAtomic: return "CalculateMaxAtomic";
        case kSlotSetStartClusterIndices: return "SetStartClusterIndices";
        case 64: return "ClusterizeScan";
        case 65: return "ClusterizeRelabel";
        case 66: return "CalculateClusterSizes";
        case 67: return "CalculateClusterMasses";
        case 68: return "CalculateLargestCluster";
        case kSlotPValuesVoxel: return "CalculatePermutationPValuesVoxelLevelInference";
        case kSlotPValuesExtent: return "CalculatePermutationPValuesClusterExtentInference";
        case kSlotPValuesMass: return "CalculatePermutationPValuesClusterMassInference";
        case kSlotBetaWeightsGLM: return "CalculateBetaWeightsGLM";
        case kSlotStatMapsGLMTTest: return "CalculateStatisticalMapsGLMTTest";
        case kSlotStatMapsGLMFTest: return "CalculateStatisticalMapsGLMFTest";
        case kSlotTTestPermutation: return "CalculateStatisticalMapsGLMTTestSecondLevelPermutation";
        case kSlotFTestPermutation: return "CalculateStatisticalMapsGLMFTestSecondLevelPermutation";
        case kSlotMeanPermutation: return "CalculateStatisticalMapsMeanSecondLevelPermutation";
        case kSlotTransformData: return "TransformData";
    }
    return "Kernel outside the second-level permutation path";
}

const char* BROCCOLI_LIB::GetOpenCLPlatformName() { return "AMD HIP (ROCm)"; }
const char* BROCCOLI_LIB::GetOpenCLDeviceName() { return device_name_.c_str(); }
int* BROCCOLI_LIB::GetOpenCLCreateKernelErrors() { return create_kernel_errors_; }
int* BROCCOLI_LIB::GetOpenCLCreateBufferErrors() { return create_buffer_errors_; }
int* BROCCOLI_LIB::GetOpenCLRunKernelErrors() { return run_kernel_errors_; }
std::string BROCCOLI_LIB::GetOpenCLInitializationError() { return init_error_; }
const char* BROCCOLI_LIB::GetOpenCLError() { return last_error_.c_str(); }
