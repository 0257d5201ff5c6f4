// Internal declarations shared by the shorefit translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "shorefit.h"

namespace shorefit {

// Thread-local last-error string behind shorefit_last_error().
void set_error(const std::string& msg);

struct Error {
    int code;
    std::string msg;
};

#define SF_HIP(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            throw ::shorefit::Error{SHOREFIT_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)}; \
    } while (0)

// Device memory of the entry points. While a StreamScope is active (one per
// entry-point call, on the calling thread), DevBuf blocks come from a
// per-thread, per-device cache of earlier calls' blocks; a block released
// during the call is only parked, and becomes reusable when the scope closes,
// after the call's stream has been synchronized, so no block is handed out
// while queued work may still use it. Repeated fits therefore allocate no
// device memory after the first call. Outside a StreamScope, DevBuf uses
// hipMalloc / hipFree directly.
void* cache_alloc(size_t bytes);
void cache_release(void* p);
bool& stream_scope_active();

struct StreamScope {
    hipStream_t s;
    bool prev_active;
    explicit StreamScope(hipStream_t stream);
    ~StreamScope();
};

// Owning device buffer (freed on scope exit; errors on free are ignored).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    bool cached = false;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    void alloc(size_t count) {
        release();
        n = count;
        if (!count) return;
        if (stream_scope_active()) {
            p = static_cast<T*>(cache_alloc(count * sizeof(T)));
            cached = true;
        } else {
            SF_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        }
    }
    void release() {
        if (p) {
            if (cached)
                cache_release(p);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        n = 0;
        cached = false;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
};

// ---------------------------------------------------------------- basis ----
int n_coeffs(uint32_t radial_order);

// Device FP64 design matrix (N x K row-major) from q-magnitudes and bvecs.
void launch_design_matrix(const double* d_q, const double* d_bvecs, int N, uint32_t radial_order, double zeta,
                          double* d_A, hipStream_t s);

// Masked voxels (mask != 0, linear order) of a device mask of n voxels into
// vox (allocated here, n entries); returns their count.
int64_t compact_mask(const uint8_t* d_mask, int64_t n, DevBuf<int64_t>& vox, hipStream_t s);
// dst[i] = src[idx[i]] (src[i] when idx is null), i < n.
void launch_gather_f32(const float* d_src, const int64_t* d_idx, int64_t n, float* d_dst, hipStream_t s);

// ------------------------------------------------------------ FP64 setup ----
// G = A^T A (K x K), A row-major N x K.
void launch_gram(const double* d_A, int N, int K, double* d_G, hipStream_t s);
// Largest eigenvalue of the symmetric K x K matrix G by power iteration
// (Rayleigh quotient), written to d_lambda[0].
void launch_power_iteration(const double* d_G, int K, double* d_lambda, hipStream_t s);
// out = G + lam_n diag(dn) + lam_l diag(dl) (K x K; dn / dl may be null),
// the two diagonal terms added in that order.
void launch_add_diagonals(const double* d_G, int K, const double* d_dn, double lam_n, const double* d_dl,
                          double lam_l, double* d_out, hipStream_t s);
// From the largest eigenvalue lam of A^T A: out[0] = L = max(lam / N, eps)
// (src/solver/fista.rs:51,58), out[1] = c = (1 / L) / N, on the device.
void launch_fit_scalars(const double* d_lam, int N, double* d_out, hipStream_t s);
// Synchronize s unless an entry point's StreamScope is active (whose buffers
// are only recycled after the call's final synchronization).
void sync_if_unscoped(hipStream_t s);
// In-place lower Cholesky of Greg (K x K, upper triangle zeroed); d_status[0]
// = 0 on success, j+1 when pivot j is not positive.
void launch_cholesky(double* d_G, int K, int* d_status, hipStream_t s);

// ------------------------------------------------------------- packing ----
// Fragment-major FP32 operands of the 16x16x4 f32 MFMA (layout described in
// fista.hip). Padded rows/cols are zero.
//   M  = -c * G               (Kp x Kp), c = step / N (identity applied in-kernel)
//   AT = c * A^T              (Kp x Np)
//   A                          (Np x Kp)
void launch_pack_M(const double* d_G, int K, int Kp, double c, float* d_frag, hipStream_t s);
void launch_pack_rowblock(const double* d_src, int rows, int cols, int src_ld, bool transpose, double scale,
                          int rows_p, int cols_p, float* d_frag, hipStream_t s);
// -c G in FP16 / BF16 16x16x32 fragment order (mixed kernel increments);
// part 1 packs the FP16 rounding remainder scaled by 2^12 (split refresh),
// part 2 the remainder unscaled in the same type (split increments).
void launch_pack_M_lowp(const double* d_G, int K, int Kp, double c, bool bf16, uint16_t* d_frag, hipStream_t s,
                        int part = 0);
// Split-FP16 operands of the Tikhonov apply (P: K x N, A: N x K, FP64, row
// major): per-row inverse scales (16 MB rows of P, then 32 G2 rows of A) and
// G1e + G2 + 2 chunks of 4 ceil(MB / 2) x 64 FP16 fragments (8 halves a
// lane; the last two are zeros). With P == nullptr and G1e = 0, A's chunks
// alone (the residual pass of the FISTA epilogue).
void launch_pack_tik16(const double* d_P, const double* d_A, int K, int N, int MB, int G1e, int G2, float* d_inv,
                       uint16_t* d_frag, hipStream_t s);
// The split-FP16 apply serves padded coefficient counts up to 16 kTik16MaxMB
// and up to 16 kTik16MaxBlocks measurements (its signals stay in registers).
constexpr int kTik16MaxMB = 10;
constexpr int kTik16MaxBlocks = 18;
// dst (rows x cols_p, FP64) = src (rows x cols) zero-padded on the right.
void launch_pad_cols(const double* d_src, int rows, int cols, int cols_p, double* d_dst, hipStream_t s);

// ---------------------------------------------------------------- fits ----
// Largest padded coefficient count whose mixed-precision increments apply
// both halves of the split operator (hi and lo, 2 x Kp^2 x 2 bytes of LDS).
constexpr int kMixedSplitMaxKp = 160;

struct FistaLaunch {
    // signals: row v of the fit is data[vox[v] * sig_stride ...] (N floats);
    // vox == nullptr means row v itself.
    const float* data;
    const int64_t* vox;
    int64_t n_vox;
    int N;
    int64_t sig_stride;
    int K, Kp, Np;
    const float* Mfrag;
    const float* ATfrag;
    const float* Afrag;
    const uint16_t* AH16;   // A's split-FP16 chunks for the epilogue (null: the FP32 epilogue)
    const float* Ainv16;    // their row scales' inverses
    const uint16_t* Mlow;   // FP16/BF16 increment operand (mixed precision only)
    const uint16_t* Mlow2;  // FP16 remainder of -cG x 2^12 (split refresh, FP16 only)
    const uint16_t* Mlow3;  // remainder -cG - lowp(-cG), unscaled (split increments, padded K <= 160)
    const double* A64p;     // FP64 design zero-padded to N x Kp (right-hand sides)
    const double* c_dev;    // step / N on the device (read by the right-hand-side kernel)
    float* Bbuf;            // scratch [n_vox][Kp]: b = c A^T s
    float* meanbuf;         // scratch [n_vox]: signal means
    float thr;              // step * alpha (fixed alpha)
    float step;             // 1 / L
    double step64;          // 1 / L in FP64
    const float* alpha_vox; // optional per-voxel alpha [n_vox] (FP32 kernel, fixed strategy)
    const float* warm;      // optional warm start, indexed like coef (FP32 kernel)
    float ratio;            // alpha-ratio strategy: alpha = ratio * alpha_max (0 = fixed)
    double alpha;           // fixed alpha (for maps / BIC)
    uint32_t max_iter;
    float tol;
    int nonneg;
    int precision;
    int refresh;
    // outputs, indexed by vox[v] (or v when vox == nullptr)
    float* coef;            // [.. x K]
    float* resid;           // optional
    float* r2;              // optional
    uint32_t* iters;        // optional
    uint8_t* converged;     // optional
    uint8_t* kind;          // optional
    float* alpha_map;       // optional
    float* bic;             // optional
    // mixed precision (persistent kernel) only
    const float* mom;       // momentum table [max_iter]
    unsigned int* counter;  // work counter
    float* Xbuf;            // [n_vox][Kp]
    uint32_t* itsbuf;       // [n_vox]
    uint8_t* statusbuf;     // [n_vox]
    float* alphabuf;        // [n_vox]
    void* event_mid;        // optional hipEvent_t recorded right after the solve kernel
    // FP32 only: persistent form (slots refilled from `counter`, momentum from
    // `mom`; with Xbuf / itsbuf / statusbuf / alphabuf the outputs go through
    // the epilogue kernel, without them the coefficients are written directly)
    bool persist;
    bool counter_zeroed;    // the caller has cleared `counter` (no memset here)
    // alpha path strategies (path_n > 0): point i of voxel v is the FISTA
    // solve at alpha_i = alpha_max(v) path_f[i], warm-started from point i-1;
    // a selection pass then writes the chosen point's outputs.
    int path_n;
    const float* path_f;    // [path_n] eps^(i/(path_n-1)), device
    const double* path_f64; // [path_n] same in FP64, device
    int path_rule;          // 0: BIC argmin, 1: L2-anchored
    double slack;
    const float* rss_l2_ref;  // L2-anchored: Tikhonov RSS, indexed like the outputs
    float* Xpath;           // [n_vox][path_n][Kp]
    uint32_t* itspath;      // [n_vox][path_n]
    uint8_t* convpath;      // [n_vox][path_n]
    // progress of the persistent solve kernels (optional): host-pinned,
    // device-visible, one count of finished voxels per wave of the grid
    unsigned int* progress;
    int progress_slots;
};
// b = c A^T s and signal means into L.Bbuf / L.meanbuf (FP64 accumulation).
void launch_fista_rhs(const FistaLaunch& L, hipStream_t s);
// The FISTA solve proper (reads L.Bbuf / L.meanbuf).
void launch_fista(const FistaLaunch& L, hipStream_t s);

// The allocation-free lasso forward's device scratch: kNoallocScratchBytes of
// a file-scope device array, one per device, shared by every call on it.
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;
// Its address on the current device (hipGetSymbolAddress on first use).
void* noalloc_scratch();
// Clears the work counter and writes the momentum table, both with kernels.
void launch_noalloc_prepare(float* mom, uint32_t n_mom, unsigned int* counter, hipStream_t s);

struct TikhonovLaunch {
    const float* data;
    const int64_t* vox;
    int64_t n_vox;
    int N;
    int64_t sig_stride;
    int K, Kp, Np;
    const float* Pfrag;  // (G + Lambda)^-1 A^T, Kp x Np fragment-major
    const float* Afrag;
    const uint16_t* H16;  // split-FP16 chunks of P and A (null: the FP32 apply)
    const float* inv16;   // their row scales' inverses
    float* coef;
    float* resid;
    float* r2;
    uint32_t* iters;
    uint8_t* kind;
    float* rss_l2;        // optional: rss of this fit (L2-anchored reference)
};
void launch_tikhonov(const TikhonovLaunch& L, hipStream_t s);

// ------------------------------------------------ gradient-deviation fits ----
// SHORE basis tables for per-voxel designs (radial order <= 10, K <= 256):
// column -> (n, l) pair and angular factor; pair -> (n, l, kappa, first
// column); angular factor -> (l, |m|, kind, normalisation). Angular factors
// are numbered l (l - 1) / 2 + slot over even l, slots in BrainSuite order.
struct GdBasis {
    int R, K, P, Q;
    double zeta;
    uint8_t col_pair[256], col_ang[256];
    uint8_t pair_n[40], pair_l[40];
    int16_t pair_col0[40];
    double kappa[40];
    uint8_t ang_l[72], ang_k[72], ang_kind[72];  // kind 0: l = 0; 1: cos(k theta); 2: m = 0; 3: sin(k theta)
    double ang_c[72];                           // normalisation (with sqrt 2 and sign)
};
GdBasis gd_basis_host(uint32_t radial_order, double zeta);

// Nominal table on the device and the frame of the effective directions.
struct GdTable {
    const double* bvals;  // [N] nominal b
    const double* gvox;   // [N][3] voxel-axis bvecs
    const double* gnom;   // [N][3] bvecs in the fit frame (kept for b0 / zero rows)
    double R[9];          // frame rotation, row-major
    double b0_thr;
    double qdenom;        // 4 pi^2 tau
    int N;
};

// One chunk of deviated voxels (slot i of the chunk is voxel vox[i]).
struct GdChunk {
    GdBasis B;
    GdTable tab;
    const float* data;
    int64_t sig_stride;
    const float* field;   // (X,Y,Z,9) f32 row-major T
    const int64_t* vox;
    int64_t n;
    float* G;             // [n][Kp][Kp]
    double* b;            // [n][Kp]
    double* sstat;        // [n][2]: s's and the sum of squares about the mean
};

// Per-slot results of the FISTA kernel.
struct GdSlotOut {
    double rss;           // |s - A x|^2 (Gram identity, FP64)
    float alpha;
    uint32_t its;
    uint16_t nnz;
    uint16_t flags;       // bit 0 converged, bit 1 all-zero short circuit
};

struct GdFistaArgs {
    const float* G;
    const double* b;
    const double* sstat;
    const int64_t* vox;
    int64_t n;
    int N, K;
    double alpha;         // fixed alpha
    double ratio;         // > 0: alpha = ratio * alpha_max (alpha-ratio strategy)
    uint32_t max_iter;
    double tol;
    int nonneg;
    const float* mom;     // [max_iter] (t_k - 1) / t_{k+1}
    unsigned* counter;
    float* coef;          // volume [.. x K], indexed by vox[slot]
    GdSlotOut* slot_out;  // [n]
    double* lip;          // [n] optional
};

// Output maps (indexed by voxel), any may be null.
struct GdMaps {
    float* resid;
    float* r2;
    uint32_t* iters;
    uint8_t* kind;
    float* alpha_map;
    float* bic;
};

// Tile shapes: padded coefficient count, Gram rows per LDS chunk, Gram
// waves, FISTA waves (RB = kp / (4 fista_waves) rows per lane group).
struct GdVariant {
    int order, kp, rc, gram_waves, fista_waves;
};
int gd_variants(uint32_t order, const GdVariant** first);
// Whether variant v's Gram kernel fits its LDS for N gradients.
bool gd_gram_fits(const GdVariant& v, int N);
void gd_launch_gram(const GdVariant& v, const GdChunk& c, hipStream_t s);
void gd_launch_fista(const GdVariant& v, const GdFistaArgs& a, int max_blocks, hipStream_t s);
void gd_launch_finish(const GdSlotOut* so, const double* sstat, const int64_t* vox, int64_t n, int N,
                      const GdMaps& m, hipStream_t s);
void gd_launch_identity(const float* field, const int64_t* vox, int64_t n, uint8_t* ident, uint8_t* devi,
                        hipStream_t s);
void launch_gather_i64(const int64_t* src, const int64_t* pos, int64_t n, int64_t* dst, hipStream_t s);

// -------------------------------------------------------- differentiation ----
size_t lasso_adjoint_lds(int K);
void launch_lasso_adjoint(const double* G, const double* A, int N, int K, const float* coef, const float* gcoef,
                          int64_t V, float* gsig, float* galpha, uint8_t* status, hipStream_t s);

// ------------------------------------------------------------- post-fit ----
// Per-mode weights of the microstructure scalars for one (radial order,
// zeta); blocks are the (n, l) pairs in coefficient order.
struct MicroTables {
    int R;
    int n00[16];                      // column of (n, 0, 0)
    double e0w[16], rtopw[16], msdw[16], qivw[16];
    int nblocks;
    int blk_l[64], blk_start[64];
    double scale_ap[64], scale_pp[64];  // RTAP / RTPP kernel over alpha(n, l)
    double qiv_pre, pref_ap, pref_pp;
    float unit[6];                    // rtop, rtap, rtpp, msd, qiv, ng
};
struct MicroThresholds {
    int active[6];
    float thr[6];
};
int n_tournier(unsigned lmax);
void tournier_matrix_host(unsigned radial_order, double zeta, unsigned lmax, double* out);
void micro_tables_host(unsigned radial_order, double zeta, int units, MicroTables& T);
// out[vox[v] * R + r] = sum_k M[r][k] coef[vox[v] * K + k] (FP64); with
// skip_zero an all-zero coefficient vector gives zeros.
void launch_apply_rows(const float* coef, const int64_t* vox, int64_t V, int K, const double* M, int R, bool skip_zero,
                       float* out, hipStream_t s);
void launch_micro(const float* coef, const int64_t* vox, int64_t V, int K, const MicroTables& T, const float* dirs,
                  float* raw, hipStream_t s);
// Fit-failure thresholds of the six maps into d_th (device), all on the device.
void launch_micro_thresholds(const float* raw, int64_t V, float factor, MicroThresholds* d_th, hipStream_t s);
void launch_micro_scatter(const float* raw, const int64_t* vox, int64_t V, const MicroThresholds* d_th,
                          float* const* maps, hipStream_t s);

// NIfTI storage order <-> voxel-major (volume.hip). raw holds the file's
// elements (NIfTI datatype code, native byte order).
void launch_to_voxel_major(const void* raw, int datatype, int64_t nx, int64_t ny, int64_t nz, int64_t nt, float slope,
                           float inter, float* out, hipStream_t s);
void launch_to_file_order(const float* in, int64_t nx, int64_t ny, int64_t nz, int64_t nt, float* out,
                          hipStream_t s);

}  // namespace shorefit
