/*
 * shorefit — MI355X (gfx950) 3D-SHORE fit behind a C ABI.
 *
 * Drop-in for the volume-level SHORE fit of PennLINC/cs_dmri 0.3.1:
 *   fit_shore(&DwiData, &ShoreFitSpec, on_voxel) -> Result<ShoreFitOutput>
 *       src/fit.rs:687-805 (called by src/bin/cs-fit.rs:287 and the pyo3
 *       binding python/src/models.rs:160)
 *   _cs_dmri.shore_fit(data, bvals, bvecs, mask, **kw) -> dict
 *       python/src/models.rs:70-198
 * plus the per-problem solver seams the reference's tests drive:
 *   FistaSolver::new / fit           src/solver/fista.rs:38-150
 *   TikhonovSolver::new / fit        src/solver/tikhonov.rs:26-76
 *   ShoreBasis::design_matrix        src/basis/shore.rs:122-178
 *
 * Conventions: plain pointers and sizes only. Volumes are float32 (X,Y,Z,N)
 * C-order, masks uint8 (X,Y,Z); coefficients come back float32 (X,Y,Z,K)
 * with zeros outside the mask (src/fit.rs:89,139-142). Outputs are caller
 * allocated (sizes from shorefit_n_coeffs). Every entry point returns a
 * SHOREFIT_* status; the message of the last failure on the calling thread
 * is shorefit_last_error(). Calls are re-entrant: each call owns its device
 * buffers and runs on the stream it is given (NULL = the default stream).
 */
#ifndef SHOREFIT_H
#define SHOREFIT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The reference's Result<_, CsDmriError> (src/lib.rs:63-83)
 * maps to SHOREFIT_E_INVALID / SHOREFIT_E_FIT; its panics (non-PD Cholesky,
 * src/solver/tikhonov.rs:44) map to SHOREFIT_E_NOT_PD. */
#define SHOREFIT_OK 0
#define SHOREFIT_E_INVALID 1     /* bad argument (CsDmriError::Other / Dimension) */
#define SHOREFIT_E_FIT 2         /* fit refused (CsDmriError::Fit: no DWI, single shell) */
#define SHOREFIT_E_NOT_PD 3      /* Tikhonov normal matrix not positive definite */
#define SHOREFIT_E_UNSUPPORTED 4 /* option this library does not implement */
#define SHOREFIT_E_DEVICE 5      /* HIP runtime failure */

/* ShoreRegularization (src/fit.rs:597-612) */
#define SHOREFIT_REG_L1 1
#define SHOREFIT_REG_L2 2
#define SHOREFIT_REG_NONNEG 3 /* AmpNonNeg (ICLS): not implemented -> SHOREFIT_E_UNSUPPORTED */

/* AlphaMode (src/fit.rs:545-552) */
#define SHOREFIT_ALPHA_FIXED 0
#define SHOREFIT_ALPHA_RATIO 1
#define SHOREFIT_ALPHA_PATH_BIC 2
#define SHOREFIT_ALPHA_L2_ANCHORED 3

/* Arithmetic of the FISTA contraction (device knob; no reference counterpart). */
#define SHOREFIT_PREC_FP32 0 /* FP32 MFMA every iteration */
#define SHOREFIT_PREC_FP16 1 /* FP16 MFMA increments + FP32 refresh every refresh_every iterations */
#define SHOREFIT_PREC_BF16 2 /* BF16 MFMA increments + FP32 refresh */

/* Mirrors ShoreFitSpec + build_alpha_strategy's inputs (src/fit.rs:555-669)
 * and the keyword set of _cs_dmri.shore_fit (python/src/models.rs:71-75). */
typedef struct ShorefitSpec {
    uint32_t radial_order;   /* even, > 0 (python/src/models.rs:145-147) */
    double zeta;             /* 700 */
    int32_t regularization;  /* SHOREFIT_REG_* */
    int32_t alpha_mode;      /* SHOREFIT_ALPHA_* */
    double alpha;            /* fixed alpha, also the seed alpha (max(alpha,1e-12)) */
    double alpha_ratio;      /* in (0,1) */
    uint32_t path_n_alphas;  /* >= 2 */
    double path_eps;         /* in (0,1); <= 0 selects the mode default (1e-3 / 1e-4) */
    double slack;            /* >= 0 */
    uint32_t max_iter;       /* FISTA cap */
    double tol;              /* relative-change tolerance */
    int32_t non_negative;
    double lambda_n;         /* Tikhonov radial weight */
    double lambda_l;         /* Tikhonov angular weight */
    int32_t compute_diagnostics;
    int32_t allow_single_shell;
    /* device knobs */
    int32_t precision;       /* SHOREFIT_PREC_* */
    int32_t refresh_every;   /* exact-refresh period of the mixed-precision kernel; <= 0 = the default
                                for the arithmetic (FP16: 40 at order <= 8, else 10; BF16: 10) */
    int32_t reserved[6];
} ShorefitSpec;

/* Gradient table (src/qspace.rs:74-120). NaN deltas mean "absent": both are
 * then estimated with TORTOISE's heuristic at Gmax = 80 mT/m (:213-227). */
typedef struct ShorefitGradients {
    const double* bvals;  /* N, s/mm^2 */
    const double* bvecs;  /* N x 3 row-major */
    int64_t n_grads;
    double big_delta;     /* seconds, NaN = absent */
    double small_delta;   /* seconds, NaN = absent */
    double b0_threshold;  /* 50 */
} ShorefitGradients;

/* DwiData's volume part (src/io/dwi.rs:22-40). */
typedef struct ShorefitVolume {
    const float* data;    /* (X,Y,Z,N) float32, C-order */
    const uint8_t* mask;  /* (X,Y,Z) uint8, NULL = every voxel */
    int64_t nx, ny, nz;
    int32_t on_device;    /* 1: data, mask and every output pointer are device pointers */
    int32_t reserved;
} ShorefitVolume;

/* FitResult + VolumeDiagnostics + the sidecar facts of ShoreFitOutput
 * (src/fit.rs:27-48, :672-683; python/src/models.rs:162-197). */
typedef struct ShorefitOutput {
    float* coefficients;          /* (X,Y,Z,K) float32, required */
    float* r2;                    /* (X,Y,Z) optional (NULL skips) */
    float* residual_l2;           /* (X,Y,Z) optional */
    uint32_t* iterations;         /* (X,Y,Z) optional */
    uint8_t* regularization_kind; /* (X,Y,Z) optional: 1 = L1, 2 = L2 */
    float* alpha;                 /* (X,Y,Z) optional, L1 only */
    float* bic;                   /* (X,Y,Z) optional, L1 only */
    float* rss_l2;                /* (X,Y,Z) optional, L2-anchored only */
    /* filled in by the library */
    int32_t n_coeffs;
    int32_t has_alpha_distribution;
    double alpha_distribution[3]; /* (median, p10, p90) for per-voxel alpha strategies */
    double big_delta, small_delta, tau;
    int32_t delta_source;         /* 0 user supplied, 1 TORTOISE estimate */
    int32_t n_shells;
    double dwi_shells[32];        /* mean b of the non-zero shells, ascending */
    int32_t n_warnings;
    char warnings[2048];          /* n_warnings messages, each NUL-terminated, back to back */
    int64_t n_voxels;             /* masked voxels fitted */
    double lipschitz;             /* sigma_max(A)^2 / N used by FISTA */
    double fit_ms;                /* device time of the fit kernels */
    double graddev_ms[3];         /* gradient-deviation fits: device time of the per-voxel Gram
                                     kernels, of the FISTA kernels (with the residual), 0 (reserved) */
} ShorefitOutput;

/* Fill *spec with the cs-fit / ShoreModel defaults (src/fit.rs:647-669):
 * order 6, zeta 700, L1 + L2-anchored path (20 alphas, eps 1e-4, slack 0.05),
 * max_iter 1000, tol 1e-6, lambda_n = lambda_l = 1e-8, FP32, refresh_every 0
 * (the arithmetic's default). */
void shorefit_spec_default(ShorefitSpec* spec);

/* Whole-volume fit: replaces fit_shore (src/fit.rs:687-805). stream is a
 * hipStream_t (NULL = default stream). */
int shorefit_fit(const ShorefitGradients* grads, const ShorefitVolume* vol, const ShorefitSpec* spec,
                 ShorefitOutput* out, void* stream);

/* Progress of whole-volume fits on the calling thread: the role of fit_shore's
 * on_voxel callback behind cs-fit's heartbeat (src/fit.rs:69-73,
 * src/bin/cs-fit.rs:286-287, src/progress.rs:17-80). The callback receives
 * (voxels finished, masked voxels, user): once with 0 when the solve starts,
 * every interval_seconds while it runs (counts the solve kernel publishes
 * into host memory, polled by this thread), and once with the total after
 * the fit. It is always called on the thread that called shorefit_fit, never
 * from device code. fn = NULL switches reporting off. */
typedef void (*shorefit_progress_fn)(int64_t voxels_done, int64_t voxels_total, void* user);
int shorefit_set_progress(shorefit_progress_fn fn, void* user, double interval_seconds);

/* ------------------------------------------------ gradient deviation ----
 * A gradient-deviation ("graddev") field gives, per voxel, a row-major 3x3 T
 * with g_eff = T^T g in the field's voxel axes (src/graddev.rs:1-24). Each
 * voxel whose T differs from the identity by IDENTITY_TOL = 1e-6 or more
 * (src/graddev.rs:35-36, :89-93) is fitted with its own effective table
 * (src/graddev.rs:160-185) and design matrix, built on the device and never
 * stored; the others use the nominal design, exactly as without a field. The
 * reference applies the field to its DTI / CSD fits through this per-voxel
 * pattern (src/dti.rs:393-407); its SHORE fit takes no field, so this entry
 * point extends fit_shore (src/fit.rs:687-805) with one. */
typedef struct ShorefitGradDev {
    const float* field;         /* (X,Y,Z,9) float32 row-major T, identity included;
                                   device memory when vol->on_device */
    const double* bvecs_vox;    /* N x 3 voxel-axis bvecs (DwiData::bvecs_vox, src/io/dwi.rs:33),
                                   host memory; NULL = grads->bvecs */
    double frame_rotation[9];   /* row-major R, grads->bvecs[i] = R bvecs_vox[i]
                                   (DwiData::frame_rotation, src/io/dwi.rs:34-37; identity for image-axis) */
    int32_t variant;            /* tile-shape variant for this radial order, or -1: the tuned one
                                   (shorefit_graddev_tune) else the default */
    int32_t reserved;
} ShorefitGradDev;

/* fit_shore with a gradient-deviation field. L1 with alpha_mode fixed or
 * alpha-ratio, radial order 2..10 (else SHOREFIT_E_UNSUPPORTED). Outputs as
 * shorefit_fit; out->lipschitz is the nominal design's. */
int shorefit_fit_graddev(const ShorefitGradients* grads, const ShorefitVolume* vol, const ShorefitGradDev* gd,
                         const ShorefitSpec* spec, ShorefitOutput* out, void* stream);

/* Number of tile-shape variants of the gradient-deviation kernels at
 * radial_order, and (when the pointers are not NULL) variant i's padded
 * coefficient count, Gram rows per chunk, Gram waves and FISTA waves. */
int shorefit_graddev_variants(uint32_t radial_order, int32_t i, int32_t* kp, int32_t* rows, int32_t* gram_waves,
                              int32_t* fista_waves);

/* Autotuner: fits the first min(max_voxels, deviated voxels) voxels with
 * every variant of spec->radial_order, stores the fastest for (radial order,
 * N) in this process (used by shorefit_fit_graddev with variant -1) and
 * writes each variant's device time in ms to ms_out (n_variants entries, may
 * be NULL; -1 for a variant whose Gram kernel cannot hold the table's
 * gradients in LDS). Returns the chosen variant's index, or a negative
 * status. */
int shorefit_graddev_tune(const ShorefitGradients* grads, const ShorefitVolume* vol, const ShorefitGradDev* gd,
                          const ShorefitSpec* spec, int64_t max_voxels, double* ms_out, void* stream);

/* Per-voxel set-up of the gradient-deviation fit (FistaSolver::new and the
 * alpha_max right-hand side for A_v, src/solver/fista.rs:38-62,
 * src/solver/alpha.rs:47-53): for each listed voxel (linear index into the
 * field and the volume, which must be host memory), G = A_v^T A_v (K x K,
 * float32 of an FP64 accumulation), b = A_v^T s (K, float64) and the
 * Lipschitz constant L = max(sigma_max(A_v)^2 / N, eps) the device solver
 * uses. Host outputs; any may be NULL. The field is used as given (no
 * identity test). */
int shorefit_graddev_gram(const ShorefitGradients* grads, const ShorefitVolume* vol, const ShorefitGradDev* gd,
                          uint32_t radial_order, double zeta, const int64_t* voxels, int64_t n_voxels,
                          float* gram_out, double* rhs_out, double* lipschitz_out);

/* ------------------------------------------------------ differentiation ----
 * The reference has no derivative of the fit; these entry points are new.
 * A design kept on the device between calls (its FP64 copy, A^T A, the
 * Lipschitz constant and the FP32 kernel operands): the set-up of
 * FistaSolver::new (src/solver/fista.rs:38-62) done once. */
typedef struct ShorefitDesign ShorefitDesign;

/* A: host, n_rows x n_cols row-major float64. */
int shorefit_design_create(const double* A, int32_t n_rows, int32_t n_cols, ShorefitDesign** out);
void shorefit_design_destroy(ShorefitDesign* design);
int shorefit_design_info(const ShorefitDesign* design, int32_t* n_rows, int32_t* n_cols, double* lipschitz);

/* FistaSolver fits (src/solver/fista.rs:70-143) of V signals on the device
 * with a per-signal alpha (alpha: V float32, or NULL for alpha_scalar) and
 * an optional warm start (warm: V x K float32, x0 = y0 = warm, t = 1, as at
 * src/solver/fista.rs:76-80; may alias coef). FP32 kernel. Device pointers:
 * signals V x N, coef V x K out, iterations V (may be NULL). */
int shorefit_lasso_forward(const ShorefitDesign* design, const float* signals, int64_t n_signals, const float* alpha,
                           double alpha_scalar, const float* warm, uint32_t max_iter, double tol,
                           int32_t non_negative, float* coef, uint32_t* iterations, void* stream);

/* shorefit_lasso_forward without device allocation. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as shorefit_lasso_forward (bit for bit); the
 * per-call buffers (right-hand sides, signal means, momentum table, work
 * counter) live in a 64 MiB device array the library holds on each GPU until
 * the process ends, so a call needs 4 (Kp + 1) bytes per signal plus
 * 4 max_iter + 512 bytes (Kp = 160 at order 8: about 100k signals per call);
 * larger calls return SHOREFIT_E_INVALID and name the limit. The call only
 * queues work on `stream` (it returns before the results are ready). Every
 * call on one GPU shares that array: calls that overlap on one GPU race, so
 * callers serialise them. */
int shorefit_lasso_forward_noalloc(const ShorefitDesign* design, const float* signals, int64_t n_signals,
                                   const float* alpha, double alpha_scalar, const float* warm, uint32_t max_iter,
                                   double tol, int32_t non_negative, float* coef, uint32_t* iterations,
                                   void* stream);

/* Vector-Jacobian product of the lasso solution by implicit differentiation
 * on each signal's active set S = {k : coef_k != 0}, z = sign(coef_S): with
 * G = A^T A, w = G_SS^-1 grad_coef_S (FP64 Cholesky), grad_signal = A_S w
 * (V x N) and grad_alpha = -N z^T w (V). Device pointers; grad_signal and
 * grad_alpha may be NULL; status (V, may be NULL) is 0, or 1 when G_SS is not
 * positive definite (the signal's gradients are then NaN). Active sets of at
 * most 192 coefficients. */
int shorefit_lasso_adjoint(const ShorefitDesign* design, const float* coef, const float* grad_coef,
                           int64_t n_signals, float* grad_signal, float* grad_alpha, uint8_t* status, void* stream);

/* Number of SHORE coefficients at radial_order (ShoreBasis::n_coeffs,
 * src/basis/shore.rs:29-44, :118-120): 72 / 145 / 256 at order 6 / 8 / 10.
 * Returns -1 for an odd order. */
int shorefit_n_coeffs(uint32_t radial_order);

/* ShoreBasis::design_matrix (src/basis/shore.rs:122-178), evaluated on the
 * device in FP64. A_out: host, N x K row-major float64. */
int shorefit_design_matrix(const ShorefitGradients* grads, uint32_t radial_order, double zeta, double* A_out);

/* FistaSolver fits of V signals against one caller design (the Solver seam,
 * src/solver/fista.rs:38-150). Host pointers: A (N x K row-major f64),
 * signals (V x N f32). Outputs: coef (V x K f32), iterations (V), converged
 * (V), residual_l2 (V f32); the last three may be NULL. lipschitz_out (may
 * be NULL) receives L = max(sigma_max^2/N, eps). */
int shorefit_fista_batch(const double* A, int32_t n_rows, int32_t n_cols, const float* signals, int64_t n_signals,
                         double alpha, uint32_t max_iter, double tol, int32_t non_negative, int32_t precision,
                         int32_t refresh_every, float* coef, uint32_t* iterations, uint8_t* converged,
                         float* residual_l2, double* lipschitz_out);

/* TikhonovSolver fits of V signals (src/solver/tikhonov.rs:26-76). d_n and
 * d_l are the K regularisation diagonals (either may be NULL = zeros).
 * Returns SHOREFIT_E_NOT_PD when A^T A + lambda_n D_n + lambda_l D_l is not
 * positive definite. */
int shorefit_tikhonov_batch(const double* A, int32_t n_rows, int32_t n_cols, const double* d_n, const double* d_l,
                            double lambda_n, double lambda_l, const float* signals, int64_t n_signals, float* coef,
                            float* residual_l2);

/* ---------------------------------------------------------- post-fit maps ----
 * Coefficient volumes are (X,Y,Z,K) float32 in C order, as shorefit_fit
 * writes them. Output volumes are zeroed outside the mask. With on_device the
 * pointers are device memory; otherwise host memory (copied in and out). */

#define SHOREFIT_UNITS_UM 0 /* q in 1/um: RTOP /um^3, RTAP /um^2, RTPP /um, MSD um^2, QIV um^5 */
#define SHOREFIT_UNITS_MM 1 /* q in 1/mm (the fit's own units) */

/* SHORE -> Tournier (MRtrix) SH projection, row-major (lmax+1)(lmax+2)/2 x K
 * (FP64, host memory); lmax even and <= radial_order. Replaces
 * cs_dmri::odf::shore_to_tournier_sh_matrix (src/odf.rs:79-102). */
int shorefit_tournier_matrix(uint32_t radial_order, double zeta, uint32_t lmax, double* out);

/* ODF in Tournier SH per voxel: out (X,Y,Z,(lmax+1)(lmax+2)/2) float32 =
 * projection x coefficients (FP64 sums), zeros outside mask (mask may be
 * NULL = every voxel). Replaces _cs_dmri.shore_odf_sh
 * (python/src/models.rs:203-248). */
int shorefit_odf_sh(const float* coefficients, const uint8_t* mask, int64_t nx, int64_t ny, int64_t nz,
                    uint32_t radial_order, double zeta, uint32_t lmax, float* out, int32_t on_device, void* stream);

/* Signal synthesized for a target gradient table: out (X,Y,Z,N') float32 =
 * design(target) x coefficients (FP64 sums); voxels whose coefficients are
 * all zero give zeros. Replaces cs_dmri::synthesize_volume
 * (src/synth.rs:15-84) behind _cs_dmri.shore_predict
 * (python/src/models.rs:255-278). */
int shorefit_predict(const float* coefficients, int64_t nx, int64_t ny, int64_t nz, uint32_t radial_order,
                     double zeta, const ShorefitGradients* target, float* out, int32_t on_device, void* stream);

/* RTOP, RTAP, RTPP, MSD, QIV and NG maps over the mask (required), in the
 * given units; directions (X,Y,Z,3) float32 (may be NULL) are the peak
 * directions for RTAP / RTPP (zero or non-finite = no peak -> NaN). When
 * outlier_factor > 0, each map's finite values above outlier_factor x its
 * 99th percentile become NaN. Any output pointer may be NULL. Replaces
 * _cs_dmri.shore_microstructure (python/src/models.rs:286-331) /
 * cs_dmri::io::microstructure::compute_microstructure
 * (src/io/microstructure.rs:160-300). */
int shorefit_microstructure(const float* coefficients, const uint8_t* mask, int64_t nx, int64_t ny, int64_t nz,
                            uint32_t radial_order, double zeta, const float* directions, int32_t units,
                            float outlier_factor, float* rtop, float* rtap, float* rtpp, float* msd, float* qiv,
                            float* ng, int32_t on_device, void* stream);

/* NIfTI storage order -> voxel-major f32. raw holds the elements of a 4D
 * series as the file stores them (volume after volume, x fastest; native byte
 * order; NIfTI-1 datatype code 2, 4, 8, 16, 64, 256, 512 or 768); out
 * receives (nx, ny, nz, nt) float32 in C order, each voxel's samples
 * contiguous. With scl_slope != 0 every value is f32(raw) * scl_slope +
 * scl_inter in f32. raw and out are both host or both device pointers
 * (on_device). Replaces cs_dmri::io::dwi::read_dwi_volume / write_volume_block
 * (src/io/dwi.rs:219-313). */
int shorefit_volume_to_voxel_major(const void* raw, int32_t datatype, int64_t nx, int64_t ny, int64_t nz, int64_t nt,
                                   float scl_slope, float scl_inter, float* out, int32_t on_device, void* stream);

/* Voxel-major (nx, ny, nz, nt) float32 -> NIfTI storage order (volume after
 * volume, x fastest), float32. Replaces the layout step of the coefficient
 * NIfTI writer, CoefficientsFile::write (src/io/coeffs.rs:112-160). */
int shorefit_volume_to_file_order(const float* in, int64_t nx, int64_t ny, int64_t nz, int64_t nt, float* out,
                                  int32_t on_device, void* stream);

/* Message of the last failure on this thread ("" if none). */
const char* shorefit_last_error(void);

/* Library version string. */
const char* shorefit_version(void);

/* Build facts for provenance records, as a JSON object: version, git SHA of
 * the sources (with "-dirty" for uncommitted changes), UTC build time and
 * offload architecture. Replaces the compile-time VERSION / GIT_SHA /
 * BUILD_TIMESTAMP the reference bakes in (src/lib.rs, build.rs:1-37). */
const char* shorefit_build_info(void);

/* Returns the calling thread's cached device blocks to the driver (after a
 * device synchronization). Entry points keep the blocks they release for
 * later calls on the same thread, up to 1/16 of the device's memory (at most
 * 16 GiB); a process that shares the device with another allocator (torch)
 * calls this when that allocator runs short or at teardown. */
int shorefit_release_cache(void);

#ifdef __cplusplus
}
#endif

#endif /* SHOREFIT_H */
