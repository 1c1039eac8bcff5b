/*
 * broccoli_lib.h -- drop-in BROCCOLI_LIB for the second-level permutation
 * test, running on one MI355X through the bpg C ABI (include/bpg.h).
 *
 * The class keeps the names, signatures and call sequence of the reference's
 * class for this path (code/BROCCOLI_LIB/broccoli_lib.h, "HDR" below), so a
 * caller such as code/Bash_Wrapper/RandomiseGroupLevel.cpp:1414-1583
 * compiles and links against it unchanged:
 *     construct -> GetOpenCLInitiated() -> setters -> one
 *     Perform*SecondLevelPermutationWrapper() -> error getters -> destruct.
 * Setters store the caller's pointers (no copies, HDR:119-333); outputs are
 * written in place.  Statuses of the C ABI land in the reference's
 * 102-entry error arrays at the slots of the kernels they replace (89/90/91
 * permuted statistics, 26 max, 92 transform, 73/85/86 observed maps, 63-68
 * clustering, 70-72 p-values).
 *
 * Differences from the reference (by design):
 *   - the model is rebuilt in FP64 from X and C; the caller's FP32 xtxxt and
 *     ctxtxc are accepted and not used;
 *   - no runtime kernel compilation; GetOpenCLBuildInfo reports the
 *     ahead-of-time gfx950 build;
 *   - a mean design without a permutation file gets the identity
 *     permutation (the reference reads the caller's uninitialised matrix);
 *   - extensions: SetPermutationPrecision (BPG_PREC_*; default split-BF16),
 *     SetClusterConnectivity (26 default, or 6), GetPermutationPhaseTimes,
 *     GetPermutationTies, GetPermutationRepeats.
 */
#ifndef BROCCOLI_LIB_H
#define BROCCOLI_LIB_H

#include <cstddef>
#include <string>
#include <vector>

typedef unsigned int cl_uint;

struct bpg_ctx;

class BROCCOLI_LIB {
public:
    // HDR:114-117
    BROCCOLI_LIB();
    BROCCOLI_LIB(cl_uint platform, cl_uint device);
    BROCCOLI_LIB(cl_uint platform, cl_uint device, const char* location);
    BROCCOLI_LIB(cl_uint platform, cl_uint device, int wrapper, bool verbos);
    ~BROCCOLI_LIB();

    // control (HDR:122-125, 157-165)
    void SetPrint(bool print);
    void SetVerbose(bool verbos);
    void SetWrapper(int wrapper);
    void SetAllocatedHostMemory(size_t allocated);
    void SetStatisticalTest(int test);
    void SetGroupDesigns(int* designs);
    void SetInferenceMode(int mode);
    void SetClusterDefiningThreshold(float threshold);
    void SetPermutationMatrices(unsigned short int** matrices);
    void SetSignMatrix(float* signs);
    void SetPermutationFileUsage(bool use);
    void SetDoAllPermutations(bool all);
    void SetSignificanceLevel(float value);

    // model (HDR:145-154)
    void SetNumberOfGLMRegressors(size_t NR);
    void SetNumberOfContrasts(size_t NC);
    void SetDesignMatrix(float* X_GLM, float* xtxxt_GLM);
    void SetContrasts(float* contrasts);
    void SetGLMScalars(float* ctxtxc);
    void SetNumberOfPermutations(size_t K);
    void SetNumberOfGroupPermutations(size_t* K);

    // data (HDR:243-258)
    void SetMNIWidth(size_t w);
    void SetMNIHeight(size_t h);
    void SetMNIDepth(size_t d);
    void SetInputMNIBrainMask(float* input);
    void SetInputFirstLevelResults(float* input);
    void SetNumberOfSubjects(size_t N);
    void SetNumberOfSubjectsGroup1(int* N);
    void SetNumberOfSubjectsGroup2(int* N);

    // outputs (HDR:275, 285, 329)
    void SetOutputStatisticalMapsMNI(float* output);
    void SetOutputPValuesMNI(float* output);
    void SetOutputPermutationDistributions(float** output);
    // outputs of the GLM without permutations (HDR:263, 280, 282) and what
    // it copies out (HDR:137-139; SetSaveResiduals*, LIB:4122-4130)
    void SetOutputBetaVolumesMNI(float* output);
    void SetOutputResidualsMNI(float* output);
    void SetOutputResidualVariances(float* output);
    void SetSaveResidualsMNI(bool value);
    void SetSaveResidualVariances(bool value);
    void SetBetasOnly(bool value);
    void SetContrastsOnly(bool value);
    void SetBetasAndContrastsOnly(bool value);

    // run (HDR:447-454)
    void PerformGLMTTestSecondLevelWrapper();
    void PerformGLMFTestSecondLevelWrapper();
    void PerformMeanSecondLevelPermutationWrapper();
    void PerformGLMTTestSecondLevelPermutationWrapper();
    void PerformGLMFTestSecondLevelPermutationWrapper();

    // diagnostics (HDR:359-413)
    bool GetOpenCLInitiated();
    int GetNumberOfOpenCLKernels();
    std::vector<std::string> GetKernelFileNames();
    int GetNumberOfKernelFiles();
    const char* GetOpenCLErrorMessage(int error);
    const char* GetOpenCLKernelName(int kernel);
    const char* GetOpenCLPlatformName();
    const char* GetOpenCLDeviceName();
    std::vector<std::string> GetOpenCLBuildInfo();
    int* GetOpenCLCreateKernelErrors();
    int* GetOpenCLCreateBufferErrors();
    int* GetOpenCLRunKernelErrors();
    std::string GetOpenCLInitializationError();
    const char* GetOpenCLError();

    // extensions
    void SetPermutationPrecision(int precision);
    void SetClusterConnectivity(int connectivity);
    // milliseconds of the last Perform call: permutation phase (all
    // statistical maps), observed maps, p-values
    void GetPermutationPhaseTimes(double* permutation_ms, double* observed_ms, double* pvalues_ms);
    // ties of the last Perform call (bpg_plan_pvalues_ties): p-value
    // comparisons inside the error band, summed over the statistical maps,
    // and (cluster inference) permutations whose map had a statistic within
    // the error band of the cluster-defining threshold
    void GetPermutationTies(long* pvalue_ties, long* cluster_tie_maps);
    // permutations of the last Perform call whose results come from the
    // FP32 and FP64 forms (repeats of ill-conditioned permutations)
    void GetPermutationRepeats(long* fp32_permutations, long* fp64_permutations);

    static const int NUMBER_OF_OPENCL_KERNELS = 102;

private:
    void Init(cl_uint platform, cl_uint device);
    void Perform(int test);
    void PerformGLM(int test);
    void RecordError(int slot, int status, bool buffer);

    bpg_ctx* ctx_ = nullptr;
    bool initiated_ = false;
    std::string init_error_;
    std::string last_error_;
    std::string device_name_;
    int wrapper_ = 0;
    bool verbose_ = false;
    bool print_ = false;

    float* h_Y_ = nullptr;
    float* h_mask_ = nullptr;
    size_t W_ = 0, H_ = 0, D_ = 0, N_ = 0;
    int* n1_ = nullptr;
    int* n2_ = nullptr;
    int test_ = 0, inference_ = 1;
    float cdt_ = 2.5f, alpha_ = 0.05f;
    size_t K_all_ = 0;
    size_t* K_ = nullptr;
    size_t p_ = 0, nc_ = 0;
    float* X_ = nullptr;
    float* xtxxt_ = nullptr;
    float* C_ = nullptr;
    float* ctxtxc_ = nullptr;
    int* designs_ = nullptr;
    unsigned short int** perms_ = nullptr;
    float* signs_ = nullptr;
    bool use_file_ = false, do_all_ = false;
    size_t allocated_ = 0;
    float* out_maps_ = nullptr;
    float* out_p_ = nullptr;
    float** out_dist_ = nullptr;
    float* out_betas_ = nullptr;
    float* out_residuals_ = nullptr;
    float* out_resvar_ = nullptr;
    bool save_residuals_ = false, save_resvar_ = false;
    bool betas_only_ = false, contrasts_only_ = false, betas_and_contrasts_only_ = false;
    int precision_ = 1;
    int connectivity_ = 26;
    double t_perm_ = 0.0, t_obs_ = 0.0, t_pval_ = 0.0;
    long pvalue_ties_ = 0, cdt_tie_maps_ = 0, fp32_perms_ = 0, fp64_perms_ = 0;

    int create_kernel_errors_[NUMBER_OF_OPENCL_KERNELS];
    int create_buffer_errors_[NUMBER_OF_OPENCL_KERNELS];
    int run_kernel_errors_[NUMBER_OF_OPENCL_KERNELS];
};

#endif
