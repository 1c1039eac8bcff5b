/*
 * bpg.h -- C ABI of the MI355X permutation-test GLM: the second-level
 * (group) test and the first-level (single-subject) test.
 *
 * This is the thin layer between the BROCCOLI_LIB drop-in class
 * (include/broccoli_lib.h) and the CDNA4 HIP kernels.  Plain pointers and
 * sizes only; every function returns BPG_OK (0) or a negative status and
 * never throws.  All host buffers are caller-owned; device memory is owned by
 * the context / plan objects.
 *
 * Reference paths (BROCCOLI repository):
 *   LIB = code/BROCCOLI_LIB/broccoli_lib.cpp, HDR = .../broccoli_lib.h
 *   RGL = code/Bash_Wrapper/RandomiseGroupLevel.cpp
 *   CONST = code/BROCCOLI_LIB/broccoli_constants.h
 *
 * Layouts are the reference's (HDR setters, RGL:1148, RGL:1184-1192):
 *   Y[x + y*W + z*W*H + s*W*H*D]   mask[x + y*W + z*W*H] (in-mask iff 1.0f)
 *   X[s + r*N]                     C[r + c*p]
 *   perm[s + k*N] (0-based uint16) sign[s + k*N] (+-1.0f)
 *   maps / p-values [x + y*W + z*W*H + c*W*H*D]
 */
#ifndef BPG_H
#define BPG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define BPG_OK 0
#define BPG_ERR_ARG (-1)         /* invalid argument / shape                 */
#define BPG_ERR_HIP (-2)         /* HIP runtime failure (see bpg_last_error) */
#define BPG_ERR_NOMEM (-3)       /* host or device allocation failed         */
#define BPG_ERR_DESIGN (-4)      /* rank-deficient design or contrast        */
#define BPG_ERR_UNSUPPORTED (-5) /* valid input outside the built kernels    */
#define BPG_ERR_NODEVICE (-6)    /* no HIP device with that index            */

/* statistical tests, CONST:78-80 */
#define BPG_TTEST 0
#define BPG_FTEST 1
#define BPG_GROUP_MEAN 2
/* inference modes, CONST:51-53 */
#define BPG_VOXEL 0
#define BPG_CLUSTER_EXTENT 1
#define BPG_CLUSTER_MASS 2
#define BPG_TFCE 3  /* threshold-free cluster enhancement (LIB:18464-18508) */
/* group designs, CONST:82-83 (+ the CLI's mean design, RGL:972-976) */
#define BPG_TWOSAMPLE 0
#define BPG_CORRELATION 1
#define BPG_MEAN_DESIGN 2
/* arithmetic of the permutation contraction */
#define BPG_PREC_FP32 0       /* FP32 contraction (FP32 MFMA)                  */
#define BPG_PREC_SPLIT_BF16 1 /* hi/lo BF16 MFMA with FP32 accumulation        */
#define BPG_PREC_FP64 2       /* FP64 VALU, residual formed directly (also the
                                 repeat of ill-conditioned permutations)      */
#define BPG_PREC_SPLIT_F16 3  /* hi/lo FP16 MFMA, first-level plans only       */
#define BPG_PREC_SPLIT_FP8 5  /* second level: e4m3 terms with an E8M0 scale per
                                 32 subjects on the block-scaled FP8 matrix
                                 cores (N <= 256, 16 fit columns)              */
#define BPG_PREC_FP32_VALU 4  /* FP32 on the vector ALUs (second level; plain
                                 BPG_PREC_FP32 runs on the FP32 matrix cores
                                 up to 16 fit columns, here above)            */

/* limits of the reference (RGL:630-636 p <= 25, RGL:726-732 F rows <= 10) */
#define BPG_MAX_REGRESSORS 25
#define BPG_MAX_FROWS 10
#define BPG_MAX_SUBJECTS 256

typedef struct bpg_ctx bpg_ctx;
typedef struct bpg_plan bpg_plan;

typedef struct {
    int test;         /* BPG_TTEST / BPG_FTEST / BPG_GROUP_MEAN               */
    int inference;    /* BPG_VOXEL / BPG_CLUSTER_EXTENT / BPG_CLUSTER_MASS /
                         BPG_TFCE                                             */
    int precision;    /* BPG_PREC_* (split-BF16 plans of designs with more
                         than 16 fit columns run the FP32 form)              */
    int connectivity; /* 26 (the reference's) or 6                            */
    int W, H, D;      /* grid                                                 */
    int N;            /* subjects                                             */
    int p;            /* regressors (group mean: 1)                           */
    int nc;           /* t: contrasts; F: rows of the F contrast; mean: 1     */
    const float* Y;    /* host, N*W*H*D                                       */
    const float* mask; /* host, W*H*D                                         */
    const float* X;    /* host, N*p  (group mean: the all-ones column)        */
    const float* C;    /* host, p*nc (group mean: {1})                        */
    float cdt;         /* cluster-defining threshold                          */
} bpg_study_desc;

/* Per-phase timings of the last bpg_plan_permutation_phase call, from HIP
 * events on the plan's stream. */
typedef struct {
    double phase_ms;      /* first permutation launch -> max-T on host        */
    double transform_ms;  /* nuisance residualisation + operand split         */
    double operand_ms;    /* permutation operand build                        */
    double contract_ms;   /* fused contraction + statistic + max-T kernel     */
    double cluster_ms;    /* cluster labelling (cluster inference only)       */
    int contract_launches;
    long perm_voxels;     /* K * V_mask of the phase                          */
    int fit_columns;      /* columns of the permuted fit: p, or p - 1 when the
                             constant vector lies in the nuisance space (it is
                             permutation-invariant and carries no signal)     */
    int slots_per_perm;   /* split-BF16 accumulator rows per permutation      */
    int precision_used;   /* BPG_PREC_* of the results                        */
    int unsafe_waves;     /* waves of the first run that met a sample with
                             SSE below 1/8 (split-BF16) or 1/256 (FP32, and
                             split-BF16 sign flips at N <= 64) of |y*|^2.  The
                             permutations with such samples are repeated in
                             the next form (split-BF16 -> FP32 -> FP64) until
                             none is flagged; FP64 forms the residual itself
                             and is never flagged                            */
    long fp32_permutations; /* permutations whose results come from FP32      */
    long fp64_permutations; /* permutations whose results come from FP64      */
    long cdt_tie_maps;    /* cluster inference: permutations whose map has an
                             in-mask statistic within 1e-4 (1 + |cdt|) of the
                             threshold (not counted for FP64 results): their
                             largest cluster may differ from an exact
                             computation's                                  */
    long pvalue_ties;     /* set by bpg_plan_pvalues: the (voxel, permutation)
                             comparisons inside the error band (the total of
                             its `ties` output)                              */
    long fp64_samples;    /* split-BF16 max-T of a t test with few residual
                             degrees of freedom (up to 64 subjects): the
                             ill-conditioned samples left out of the maxima
                             and recomputed one by one in FP64 (a list sized
                             from the expected count; samples past it send
                             their permutations to the whole repeat above;
                             BPG_SAMPLE_LIST=off|on and BPG_SAMPLE_CAP, the
                             entries per each of 256 segments, override the
                             choice and the size)                            */
} bpg_timing;

/* ---- context: replaces BROCCOLI_LIB(platform, device, ...) +
 *      OpenCLInitiate (LIB:151-158, LIB:1181-1420) and ~BROCCOLI_LIB
 *      (LIB:160-163).  device = HIP ordinal. */
int bpg_create(int device, bpg_ctx** out);
void bpg_destroy(bpg_ctx* ctx);
/* GetOpenCLError / GetOpenCLInitializationError (HDR:412-413) */
const char* bpg_last_error(const bpg_ctx* ctx);
/* GetOpenCLDeviceName / GetOpenCLPlatformName (HDR:391-394) */
const char* bpg_device_name(const bpg_ctx* ctx);

/* ---- plan: device-resident study.  Replaces the buffer creation and
 *      uploads of Perform*SecondLevelPermutationWrapper (LIB:13148-13188,
 *      LIB:13228-13261) and SetupPermutationTestSecondLevel (LIB:16128).
 *      Only in-mask voxels are uploaded (compacted, subject-major). */
int bpg_plan_create(bpg_ctx* ctx, const bpg_study_desc* desc, bpg_plan** out);
/* ---- the checks of bpg_plan_create on the host alone, no device needed (no
 *      counterpart in the reference, which checks sizes in the CLI,
 *      RGL:599-733): argument ranges, a design and contrasts of full rank,
 *      and finite data inside the mask (a NaN or infinite value there is
 *      refused: the reference's (int)(T * 1e4) of the resulting NaN, KM:299,
 *      depends on the device).  Returns the status bpg_plan_create would
 *      return for them; err (optional) receives the message. */
int bpg_validate_study(const bpg_study_desc* desc, char* err, int err_len);
void bpg_plan_destroy(bpg_plan* plan);
long bpg_plan_num_voxels(const bpg_plan* plan); /* V_mask */

/* ---- the permutation phase of one statistical map: nuisance
 *      residualisation (LIB:16654-16727 + TransformData K1:1684), then all K
 *      permutations (LIB:16733-16763: the per-permutation statistic kernels
 *      K2:1458 / K2:1562 / K4:1943 and CalculateMaxAtomic KM:282, or
 *      ClusterizeOpenCLPermutation LIB:18404) in fused launches.
 *      map: contrast index (t) or 0.  perm (t, F) or sign (group mean): host,
 *      [s + k*N].  maxdist: host, K floats, the null distribution as the
 *      reference stores it (quantised max-T, cluster size, or mass / 1e4).
 *      stats (optional, may be NULL): host, K*V_mask floats, the permuted
 *      statistic of every in-mask voxel (compact order), for parity checks. */
int bpg_plan_permutation_phase(bpg_plan* plan, int map, const unsigned short* perm,
                               const float* sign, long K, float* maxdist, float* stats);
/* ---- the permutation phases of n studies at once (no counterpart in the
 *      reference, which analyses one study per RandomiseGroupLevel run,
 *      RGL:44-1715): every plan's phase is enqueued on its own stream before
 *      any is waited for, so small studies overlap on the GPU.  Per study i:
 *      plans[i] (distinct plans), maps[i], perms[i] or signs[i] (either
 *      array may be NULL when no study needs it), K[i], maxdist[i] as for
 *      bpg_plan_permutation_phase.  Returns the first error; the studies
 *      enqueued before it are still completed. */
int bpg_plans_permutation_phase(bpg_plan* const* plans, int n, const int* maps,
                                const unsigned short* const* perms, const float* const* signs,
                                const long* K, float* const* maxdist);
int bpg_plan_last_timing(const bpg_plan* plan, bpg_timing* out);
/* ---- a batch of studies as hipGraphs (no counterpart in the reference,
 *      which analyses one study per RandomiseGroupLevel run, RGL:44-1715;
 *      the per-study work it replaces is LIB:16589-16790).
 *      bpg_batch_create takes n distinct plans of one context (split-BF16 t
 *      tests or group means, voxel inference, up to 128 subjects), their
 *      maps and permutation counts, and captures two graphs.  The studies
 *      are split in two stages: stage 0 holds those likely to meet
 *      ill-conditioned samples (a Gaussian-noise estimate from their
 *      degrees of freedom, fit columns and voxel count, which orders the
 *      work and never changes a result), stage 1 the rest.  The first graph
 *      uploads all permutations / signs and clears the work queues,
 *      ill-conditioned counts and flags with one fill kernel (not a
 *      memset); each graph then runs its stage: per class of studies that
 *      share a contraction instantiation (k-steps, accumulator slots,
 *      operand kind and the autotuner's subtile count), largest work first,
 *      its members' nuisance transforms and operand builds, then one
 *      persistent contraction whose blocks take (study, permutation group,
 *      voxel chunk) work items from per-XCD device queues until all are
 *      empty; one launch of the stage's max reductions; the copies of its
 *      maxima, counts and flags to the host.  The studies' device data and
 *      models are the plans' (uploaded at their creation); every per-run
 *      buffer is the batch's own.  bpg_batch_run replays both graphs for new
 *      permutations (perms[i] or signs[i], [s + k*N], K[i] rows as at
 *      creation) and returns each study's null in maxdist[i] exactly as
 *      bpg_plan_permutation_phase does: the flagged (ill-conditioned)
 *      permutations of each study are repeated through its plan in FP32 /
 *      FP64, stage 0's while the second graph runs.  A plan may be used on
 *      its own between runs. */
typedef struct bpg_batch bpg_batch;
int bpg_batch_create(bpg_plan* const* plans, int n, const int* maps, const long* K, bpg_batch** out);
int bpg_batch_run(bpg_batch* batch, const unsigned short* const* perms, const float* const* signs,
                  float* const* maxdist);
/* host wall time of the last run (graph launch to results) and of its graph
 * launches (launch to their completion), the studies whose flagged permutations were
 * repeated, the number of persistent launches */
int bpg_batch_last_timing(const bpg_batch* batch, double* run_ms, double* graph_ms, int* repeated_studies,
                          int* classes);
void bpg_batch_destroy(bpg_batch* batch);
/* ---- the permutation phase without any device allocation (no counterpart
 *      in the reference): for callers that cannot allocate device memory
 *      while they run (tight serving loops, code that must not call
 *      hipMalloc / hipFree).  The same kernels as
 *      bpg_plan_permutation_phase on the plan's device-resident study, for
 *      voxel inference (max-T) in the plan's split-BF16 or FP32 form.
 *      d_perm (t, F) or d_sign (group mean) and d_maxdist are DEVICE
 *      buffers of the caller: [s + k*N] uint16 / float, and K floats that
 *      receive the null distribution exactly as bpg_plan_permutation_phase
 *      returns it.  The per-call scratch (operand fragments, chunk maxima,
 *      flags, counters) lives in one static device array of 128 MiB per GPU
 *      (allocated with the library's code object, held until the process
 *      ends); a K whose scratch would exceed it is refused with
 *      BPG_ERR_UNSUPPORTED.  Synchronous on the plan's stream.  Calls that
 *      overlap on one GPU share that array and race.  Ill-conditioned
 *      samples, which bpg_plan_permutation_phase repeats in a more accurate
 *      form, are refused (BPG_ERR_UNSUPPORTED) and so are permutation
 *      indices >= N (BPG_ERR_ARG); maxdist is then not valid. */
int bpg_plan_permutation_phase_noalloc(bpg_plan* plan, int map, const unsigned short* d_perm,
                                       const float* d_sign, long K, float* d_maxdist);

/* ---- observed statistical maps on the original data and design
 *      (CalculateStatisticalMapsGLM{T,F}TestSecondLevel, LIB:15858-15950;
 *      K1:51, K1:1456, K1:1560).  maps: host grid, nmaps*W*H*D. */
int bpg_plan_observed_maps(bpg_plan* plan, float* maps);

/* ---- the second-level GLM without permutations
 *      (PerformGLMTTestSecondLevelWrapper / PerformGLMFTestSecondLevelWrapper,
 *      LIB:12872-13086; CalculateBetaWeightsGLM K1:51-128,
 *      CalculateStatisticalMapsGLM{T,F}Test K1:1456-1673), in FP64 on the
 *      original data and design.  Each output is optional (NULL):
 *      betas p*W*H*D (group mean: 1), maps nmaps*W*H*D as
 *      bpg_plan_observed_maps, residuals N*W*H*D, resvar W*H*D (t: sum e^2 /
 *      (N - p); F: sum (e - mean e)^2 / (N - p - 1)).  Out-of-mask voxels
 *      are 0.  Also records the observed maps for bpg_plan_pvalues. */
int bpg_plan_glm(bpg_plan* plan, float* betas, float* maps, float* residuals, float* resvar);

/* ---- permutation p-values (CalculatePermutationPValues, LIB:16794-16873,
 *      KC:516-652), on the observed maps of the last bpg_plan_observed_maps
 *      call.  K[c], maxdist[c][k] per map.  pvalues: host grid,
 *      nmaps*W*H*D.  labels (optional): host grid, uint32 cluster labels of
 *      the observed maps (min linear index of the component, 3*W*H*D
 *      outside clusters), nmaps*W*H*D. */
int bpg_plan_pvalues(bpg_plan* plan, const long* K, const float* const* maxdist,
                     float* pvalues, unsigned* labels);
/* ---- bpg_plan_pvalues with the ties of its counts (no counterpart in the
 *      reference, whose FP32 counts report none).  ties (optional): host
 *      grid, nmaps*W*H*D, per voxel the comparisons T > m_k that lie inside
 *      the error band, i.e. where an exact computation could count
 *      otherwise: voxel inference, the null values within 1e-4 (2 + |T|) of
 *      the observed T (the statistic bound 1e-4 (1 + |t|) plus one
 *      quantum of the max-T quantisation); cluster inference, for voxels
 *      above the threshold, the permutations of the map's last phase whose
 *      map had a statistic within 1e-4 (1 + |cdt|) of the threshold
 *      (cdt_tie_maps; 0 when the null passed is not that phase's); TFCE: 0.
 *      The total lands in bpg_timing.pvalue_ties. */
int bpg_plan_pvalues_ties(bpg_plan* plan, const long* K, const float* const* maxdist,
                          float* pvalues, unsigned* labels, unsigned* ties);
/* ---- the per-permutation tie flags of the last permutation phase (no
 *      counterpart in the reference): flags[k] = 1 when permutation k's map
 *      had an in-mask statistic within 1e-4 (1 + |cdt|) of the
 *      cluster-defining threshold (cluster inference; 0 otherwise and for
 *      results of the FP64 form).  flags: host, K bytes. */
int bpg_plan_last_tie_flags(const bpg_plan* plan, unsigned char* flags);

/* ---- autotuner (no counterpart in the reference, whose launch shapes are
 *      fixed: LIB:2949-2973, LIB:3117-3180).  bpg_plan_tune times the
 *      split-BF16 contraction of the plan's map `map` with K permutations
 *      (seeded, of the phase's shape) in each candidate shape -- operand
 *      subtiles per wave (the default or, for the max-T kernels up to 128
 *      subjects, the alternative) x voxel-chunk length (the planner's,
 *      halved, doubled) -- on the device, HIP-event kernel time, best of 2
 *      after a warm-up, and keeps the fastest for the bucket (k-steps of 16
 *      subjects, accumulator slots per permutation, sign operand, F test,
 *      statistic output, log2 of the padded voxel count) in the context:
 *      every later phase in that bucket uses it.  bpg_tune_save writes the
 *      context's entries to a text file (entries of other devices already
 *      there are kept), bpg_tune_load reads this device's (returns their
 *      number); bpg_create loads the file named by BPG_TUNE_CACHE. */
typedef struct {
    int variant;        /* 0: the default subtiles per wave, 1: the alternative */
    int rows_per_wave;  /* subtiles per wave of the choice                      */
    int chunk_pow;      /* voxel-chunk length x 2^chunk_pow (-1, 0, 1)          */
    int vt_per_chunk;   /* 32-voxel tiles per chunk of the choice               */
    double default_ms;  /* contraction time of the planner's default shape      */
    double best_ms;     /* of the choice                                        */
    int candidates;
} bpg_tune_result;
int bpg_plan_tune(bpg_plan* plan, int map, long K, bpg_tune_result* out);
int bpg_tune_save(const bpg_ctx* ctx, const char* path);
int bpg_tune_load(bpg_ctx* ctx, const char* path);
int bpg_tune_entries(const bpg_ctx* ctx);

/* ---- host-side pieces of the path -------------------------------------- */

/* Generators (LIB:17069-17122, LIB:16925-17011, LIB:17014-17066).
 * compat != 0: the reference's stream -- glibc's rand() algorithm restated
 *   on a library-private state, libstdc++'s std::random_shuffle order,
 *   uniqueness rejection; row 0 is the identity / all +1.  seed != 0
 *   reseeds the stream first (as srand(seed)); seed == 0 continues it.  It
 *   starts as rand() does without srand (seed 1), which is the reference's
 *   stream: it never seeds rand().
 * compat == 0: a seeded SplitMix64 stream with Fisher-Yates shuffles and the
 *   same uniqueness rule.
 * Returns BPG_ERR_ARG when K exceeds the number of distinct vectors. */
int bpg_gen_sign_matrix(int N, long K, int compat, unsigned long long seed, float* sign);
int bpg_gen_perm_two_sample(int N, int n1, int n2, long K, int compat,
                            unsigned long long seed, unsigned short* perm);
int bpg_gen_perm_correlation(int N, long K, int compat, unsigned long long seed,
                             unsigned short* perm);

/* Design-type detection (RGL:908-1044): designs[c] in BPG_TWOSAMPLE /
 * BPG_CORRELATION / BPG_MEAN_DESIGN, n1/n2 group sizes for two-sample
 * designs.  F test: one entry from X*C'. */
int bpg_detect_designs(int N, int p, int nc, const float* X, const float* C,
                       int ftest, int* designs, int* n1, int* n2);
/* Largest permutation count (RGL:38-42, RGL:1048-1134). */
double bpg_max_permutations(int test, int design, int N, int n2);
/* FP64 model precompute of the CLI (RGL:1137-1194), cast to float:
 * xtxxt[s + r*N] = pinv(X)(r,s); ctxtxc[c] (t) or [c + cc*nc] (F). */
int bpg_model_precompute(int N, int p, int nc, int ftest, const float* X,
                         const float* C, float* xtxxt, float* ctxtxc);

/* ======================================================================
 * First-level (single-subject) permutation test:
 * BROCCOLI_LIB::ApplyPermutationTestFirstLevel (LIB:16454-16558), a private
 * member that PerformFirstLevelAnalysisWrapper runs when
 * SetPermuteFirstLevel(true) (HDR:143, LIB:8996-9075).  The volumes are the
 * subject's preprocessed fMRI data; the design is the full first-level
 * design (c_X_GLM, LIB:8795-8837: task regressors, detrending and motion
 * regressors), T time points.
 *   volumes[x + y*W + z*W*H + t*W*H*D]   mask[x + y*W + z*W*H] (== 1.0f)
 *   X[t + r*T]   C[r + c*p]   perm[t + k*T] (0-based uint16)
 *   AR estimates ar[k*W*H*D + i], k = 0..3 for AR1..AR4
 * ====================================================================== */
typedef struct bpg_fl_plan bpg_fl_plan;

typedef struct {
    int test;               /* BPG_TTEST / BPG_FTEST                          */
    int inference;          /* BPG_VOXEL / BPG_CLUSTER_EXTENT /
                               BPG_CLUSTER_MASS (the reference leaves the
                               TFCE distribution unset, LIB:16524-16530)      */
    int precision;          /* the recursion in FP32 (the reference's
                               arithmetic) with the projections on the matrix
                               cores, ill-conditioned permutations repeated in
                               FP64 (p <= 16, T <= 560; other studies run
                               FP64): BPG_PREC_SPLIT_F16 (FP16 hi/lo splits of
                               power-of-two scaled series, samples with
                               SSE < |x|^2/32 repeated; p <= 16 with
                               T <= 480, p <= 25 with T <= 384, the
                               split-BF16 form for p <= 16 above that),
                               BPG_PREC_SPLIT_BF16 (BF16 hi/lo splits,
                               SSE < |x|^2/8) or BPG_PREC_FP32 (FP32 MFMA,
                               SSE < |x|^2/64); BPG_PREC_FP64: everything in
                               FP64                                          */
    int connectivity;       /* 26 (the reference's) or 6                      */
    int W, H, D, T;         /* EPI grid and time points (10 <= T <= 65535)    */
    int p;                  /* regressors, <= BPG_MAX_REGRESSORS              */
    int nc;                 /* t: contrasts; F: contrast rows (<= 10)         */
    const float* volumes;   /* host, T*W*H*D                                  */
    const float* mask;      /* host, W*H*D                                    */
    const float* X;         /* host, T*p                                      */
    const float* C;         /* host, p*nc                                     */
    float voxel_size[3];    /* EPI voxel size (mm): SetEPIVoxelSizeX/Y/Z,
                               HDR:214-216                                    */
    float ar_smoothing_fwhm;/* mm, SetARSmoothingAmount (HDR:177); the
                               reference's default is 8 (LIB:294)            */
    float cdt;              /* cluster-defining threshold                     */
} bpg_fl_desc;

/* Time-point permutations (GeneratePermutationMatrixFirstLevel,
 * LIB:16877-16922): K rows, each a cumulative random_shuffle of 0..T-1
 * distinct from the identity and every earlier row (the identity is never
 * a row).  compat / seed as the generators above. */
int bpg_gen_perm_first_level(int T, long K, int compat, unsigned long long seed, unsigned short* perm);

/* Uploads the in-mask voxels and runs, on the device in FP64, the
 * regression of the full design (PerformRegression, LIB:12144-12194) and the
 * whitening prior to the permutations (PerformWhiteningPriorPermutations,
 * LIB:16363-16452: three rounds of AR(4) estimation, normalised smoothing of
 * the estimates, accumulation and whitening).  Synchronous. */
int bpg_fl_plan_create(bpg_ctx* ctx, const bpg_fl_desc* desc, bpg_fl_plan** out);
void bpg_fl_plan_destroy(bpg_fl_plan* plan);
long bpg_fl_plan_num_voxels(const bpg_fl_plan* plan);
/* The total AR(4) estimates (d_AR1..4_Estimates after LIB:16436-16445):
 * host grid, 4*W*H*D, 0 outside the mask. */
int bpg_fl_plan_ar_estimates(bpg_fl_plan* plan, float* ar);
/* The whitened volumes the permutations draw from: host, T*W*H*D, 0
 * outside the mask. */
int bpg_fl_plan_whitened(bpg_fl_plan* plan, float* volumes);

/* ---- the permutation loop of one statistical map (LIB:16486-16540): for
 *      each of the K permutations, inverse whitening of the permuted
 *      whitened volumes (GeneratePermutedVolumesFirstLevel,
 *      kernelWhitening.cpp:480-535), the first-level statistic
 *      (kernelStatistics3.cpp:2107-2190 / kernelStatistics5.cpp:1943-2060)
 *      and the max statistic (CalculateMaxAtomic) or largest cluster
 *      (ClusterizeOpenCLPermutation).  map: contrast (t) or 0 (F).
 *      maxdist: host, K.  stats (optional): host, K*V_mask compact
 *      statistic values. */
int bpg_fl_plan_permutation_phase(bpg_fl_plan* plan, int map, const unsigned short* perm, long K,
                                  float* maxdist, float* stats);
/* setup_ms = transform_ms (regression + whitening), contract_ms = the
 * permutation kernel, cluster_ms, phase_ms (host wall time);
 * precision_used = the form that ran; with a fast form, unsafe_waves counts
 * the permutations with an ill-conditioned sample, which were repeated in
 * FP64, and fp32_permutations the others. */
int bpg_fl_plan_last_timing(const bpg_fl_plan* plan, bpg_timing* out);

#ifdef __cplusplus
}
#endif
#endif
