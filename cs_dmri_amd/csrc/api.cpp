// C ABI of the shorefit library: validation and dispatch of the SHORE fit
// (src/fit.rs:545-805), gradient-table handling (src/qspace.rs:90-227), and
// the device set-up / fit launches.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <thread>
#include <string>
#include <vector>

#include "common.h"

namespace shorefit {

namespace {
thread_local std::string g_last_error;
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double kTortoiseGmax = 80e-3;      // src/qspace.rs:15
constexpr double kGyro = 267.51532e6;        // src/qspace.rs:18
constexpr double kShellTolerance = 50.0;     // src/fit.rs:618
constexpr size_t kMinShoreShells = 2;        // src/fit.rs:614
}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }

bool& stream_scope_active() {
    thread_local bool a = false;
    return a;
}

namespace {
struct BlockCache {
    std::map<int, std::multimap<size_t, void*>> free;  // device -> size -> block
    std::map<void*, std::pair<int, size_t>> size_of;   // block -> (device, bytes)
    std::vector<void*> parked;                         // released during the current call
    size_t cached_bytes = 0;
    ~BlockCache() {
        for (auto& kv : size_of) (void)hipFree(kv.first);
    }
};
BlockCache& block_cache() {
    thread_local BlockCache c;
    return c;
}
// Blocks are rounded up to 2 MiB (small ones to 4 KiB) so that sizes recur.
size_t round_block(size_t bytes) {
    const size_t g = bytes >= (size_t(1) << 20) ? (size_t(2) << 20) : 4096;
    return (bytes + g - 1) / g * g;
}
// Cached (unused) bytes kept per thread beyond which free blocks go back to
// the driver: 1/16 of the device's memory, and never more than 16 GiB, so
// that a process sharing the device (torch's allocator) keeps its room.
size_t cache_keep_bytes(int dev) {
    static std::mutex mu;
    static std::map<int, size_t> keep;
    std::lock_guard<std::mutex> lock(mu);
    auto it = keep.find(dev);
    if (it != keep.end()) return it->second;
    size_t free_b = 0, total_b = 0;
    size_t k = size_t(1) << 30;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0)
        k = std::min(total_b / 16, size_t(16) << 30);
    else
        (void)hipGetLastError();
    keep[dev] = k;
    return k;
}

// Frees every cached block of the calling thread (all devices).
void drop_cached_blocks(BlockCache& c) {
    for (auto& dv : c.free) {
        for (auto& kv : dv.second) {
            c.size_of.erase(kv.second);
            (void)hipFree(kv.second);
        }
        dv.second.clear();
    }
    c.cached_bytes = 0;
}
}  // namespace

void* cache_alloc(size_t bytes) {
    BlockCache& c = block_cache();
    int dev = 0;
    SF_HIP(hipGetDevice(&dev));
    const size_t need = round_block(bytes);
    auto& fl = c.free[dev];
    auto it = fl.lower_bound(need);
    if (it != fl.end() && it->first <= 2 * need) {
        void* p = it->second;
        c.cached_bytes -= it->first;
        fl.erase(it);
        return p;
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, need);
    if (e != hipSuccess) {
        // give the cached blocks of this device back and retry once
        (void)hipGetLastError();
        SF_HIP(hipDeviceSynchronize());
        for (auto& kv : fl) {
            c.cached_bytes -= kv.first;
            c.size_of.erase(kv.second);
            (void)hipFree(kv.second);
        }
        fl.clear();
        SF_HIP(hipMalloc(&p, need));
    }
    c.size_of[p] = {dev, need};
    return p;
}

void cache_release(void* p) { block_cache().parked.push_back(p); }

StreamScope::StreamScope(hipStream_t stream) : s(stream), prev_active(stream_scope_active()) {
    stream_scope_active() = true;
}

StreamScope::~StreamScope() {
    stream_scope_active() = prev_active;
    if (prev_active) return;  // an enclosing scope parks and recycles
    BlockCache& c = block_cache();
    if (c.parked.empty()) return;
    // The call's queued work may still read parked blocks (also on an error
    // path): wait for it before they become reusable.
    (void)hipStreamSynchronize(s);
    for (void* p : c.parked) {
        auto it = c.size_of.find(p);
        if (it == c.size_of.end()) continue;
        if (c.cached_bytes + it->second.second > cache_keep_bytes(it->second.first)) {
            (void)hipFree(p);
            c.size_of.erase(it);
            continue;
        }
        c.free[it->second.first].emplace(it->second.second, p);
        c.cached_bytes += it->second.second;
    }
    c.parked.clear();
}

namespace {

// Padded coefficient count: the smallest instantiated multiple of 16 (of 32
// for the mixed-precision kernel, whose FP16/BF16 MFMA reduces 32 at a time).
int pad_coeffs(int K, bool mixed = false) {
    static const int mbs[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 16};
    static const int mbs_mixed[] = {2, 4, 6, 8, 10, 16};
    if (mixed) {
        for (int mb : mbs_mixed)
            if (16 * mb >= K) return 16 * mb;
        return -1;
    }
    for (int mb : mbs)
        if (16 * mb >= K) return 16 * mb;
    return -1;
}
int pad16(int n) { return (n + 15) / 16 * 16; }

// src/qspace.rs:213-227 — TORTOISE's (Delta, delta) heuristic.
void tortoise_deltas(const double* bvals, int64_t n, double* big, double* small) {
    double mb = 0.0;
    for (int64_t i = 0; i < n; ++i) mb = std::max(mb, bvals[i]);
    if (mb <= 0.0) {
        *big = 3e-3;
        *small = 1e-3;
        return;
    }
    const double temp = mb / kGyro / kGyro / kTortoiseGmax / kTortoiseGmax / 2.0 * 1e6;
    const double small_ms = std::pow(temp, 1.0 / 3.0) * 1000.0;
    *big = small_ms * 3.0 * 1e-3;
    *small = small_ms * 1e-3;
}

// src/qspace.rs:171-197 + src/fit.rs:615-617 — greedy shell clustering.
std::vector<double> dwi_shells(const double* bvals, int64_t n, double b0_thr) {
    std::vector<double> b;
    std::vector<double> cnt;
    for (int64_t i = 0; i < n; ++i) {
        if (bvals[i] <= b0_thr) continue;
        bool placed = false;
        for (size_t s = 0; s < b.size(); ++s)
            if (std::fabs(b[s] - bvals[i]) <= kShellTolerance) {
                b[s] = (b[s] * cnt[s] + bvals[i]) / (cnt[s] + 1.0);
                cnt[s] += 1.0;
                placed = true;
                break;
            }
        if (!placed) {
            b.push_back(bvals[i]);
            cnt.push_back(1.0);
        }
    }
    std::stable_sort(b.begin(), b.end());
    return b;
}

std::string fmt_b(double b) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.0f", b);
    return buf;
}

struct Warnings {
    std::vector<std::string> msgs;
    void to(ShorefitOutput* out) const {
        out->n_warnings = 0;
        size_t pos = 0;
        std::memset(out->warnings, 0, sizeof(out->warnings));
        for (const auto& m : msgs) {
            if (pos + m.size() + 1 > sizeof(out->warnings)) break;
            std::memcpy(out->warnings + pos, m.c_str(), m.size() + 1);
            pos += m.size() + 1;
            out->n_warnings++;
        }
    }
};

// Device-side design and its derived operands.
struct Design {
    int N = 0, K = 0, Np = 0, Kp = 0;
    DevBuf<double> A;   // N x K row-major, FP64
    DevBuf<double> G;   // K x K
    DevBuf<double> qv;  // q and the gradient directions the design was built from (4N)
    double L = 0.0;     // Lipschitz constant sigma_max^2 / N (floored at eps)
};

// Per-thread pinned host staging for the small uploads of a call (q and the
// gradient directions, the momentum table), so that they are true
// asynchronous copies: a copy from pageable memory would wait for the work
// already queued on the stream, which leaves the device idle while the host
// catches up. Before a buffer is rewritten the host waits for the event
// recorded after its previous copy (normally long complete).
template <int TAG>
struct PinnedStage {
    void* p = nullptr;
    size_t bytes = 0;
    hipEvent_t done = nullptr;
    ~PinnedStage() {
        if (done) (void)hipEventSynchronize(done), (void)hipEventDestroy(done);
        if (p) (void)hipHostFree(p);
    }
    // host memory of at least n bytes, free for writing
    void* get(size_t n) {
        if (done) SF_HIP(hipEventSynchronize(done));
        if (n > bytes) {
            if (p) SF_HIP(hipHostFree(p));
            p = nullptr;
            bytes = 0;
            SF_HIP(hipHostMalloc(&p, n));
            bytes = n;
        }
        return p;
    }
    // after the copy out of the buffer is queued on s
    void queued(hipStream_t s) {
        if (!done) SF_HIP(hipEventCreateWithFlags(&done, hipEventDisableTiming));
        SF_HIP(hipEventRecord(done, s));
    }
};
template <int TAG>
PinnedStage<TAG>& pinned_stage() {
    thread_local PinnedStage<TAG> st;
    return st;
}

void upload_design(Design& d, const double* A_host, int N, int K, hipStream_t s, bool mixed = false) {
    d.N = N;
    d.K = K;
    d.Np = pad16(N);
    d.Kp = pad_coeffs(K, mixed);
    if (d.Kp < 0)
        throw Error{SHOREFIT_E_UNSUPPORTED, "at most 256 coefficients (radial order 10) are supported, got " +
                                                std::to_string(K)};
    d.A.alloc(size_t(N) * K);
    SF_HIP(hipMemcpyAsync(d.A.p, A_host, sizeof(double) * N * K, hipMemcpyHostToDevice, s));
}

void build_design(Design& d, const ShorefitGradients& g, double big, double small, uint32_t order, double zeta,
                  hipStream_t s, bool mixed = false) {
    const int N = int(g.n_grads);
    const int K = n_coeffs(order);
    d.N = N;
    d.K = K;
    d.Np = pad16(N);
    d.Kp = pad_coeffs(K, mixed);
    if (d.Kp < 0)
        throw Error{SHOREFIT_E_UNSUPPORTED,
                    "radial order " + std::to_string(order) + " (" + std::to_string(K) +
                        " coefficients) exceeds the supported maximum of 10 (256 coefficients)"};
    // src/qspace.rs:130-149 — q = sqrt(b / (4 pi^2 tau)), 0 at or below the b0 threshold.
    auto& stage = pinned_stage<0>();
    double* q = static_cast<double*>(stage.get(sizeof(double) * 4 * N));  // q, then the directions
    const double tau = big - small / 3.0;
    const double denom = 4.0 * kPi * kPi * tau;
    for (int i = 0; i < N; ++i) q[i] = (g.bvals[i] <= g.b0_threshold) ? 0.0 : std::sqrt(g.bvals[i] / denom);
    std::memcpy(q + N, g.bvecs, sizeof(double) * 3 * N);
    d.qv.alloc(size_t(4) * N);  // kept with the design: no wait for the kernel here
    SF_HIP(hipMemcpyAsync(d.qv.p, q, sizeof(double) * 4 * N, hipMemcpyHostToDevice, s));
    stage.queued(s);
    d.A.alloc(size_t(N) * K);
    launch_design_matrix(d.qv.p, d.qv.p + N, N, order, zeta, d.A.p, s);
}

struct FistaOps {
    DevBuf<float> M, A;
    DevBuf<double> Ap;  // FP64 design zero-padded to N x Kp
    DevBuf<uint16_t> Mlow, Mlow2, Mlow3;
    DevBuf<uint16_t> AH16;  // A's split-FP16 chunks (epilogue residual)
    DevBuf<float> Ainv16;
    double step = 0.0;
};

// Device buffers of one FISTA solve: right-hand sides, the momentum table,
// the Lipschitz scalars, and for the mixed-precision persistent kernel its
// work counter and work-order results (and the alpha-path points).
struct FistaWork {
    DevBuf<float> B, mean, mom, X, alpha, Xpath;
    DevBuf<double> lam, sc;
    DevBuf<unsigned int> counter;
    DevBuf<uint32_t> its, itspath;
    DevBuf<uint8_t> status, convpath;
};

// Progress reporting (the reference's on_voxel callback behind cs-fit's
// heartbeat, src/fit.rs:69-73, src/progress.rs:17-80): a callback registered
// per thread, polled from the calling thread while the persistent solve
// kernel runs. The kernel publishes each wave's finished-voxel count into a
// host-pinned array; nothing is called from the device.
struct ProgressReg {
    shorefit_progress_fn fn = nullptr;
    void* user = nullptr;
    double interval = 1.0;
};
ProgressReg& progress_reg() {
    thread_local ProgressReg r;
    return r;
}
constexpr int kProgressSlots = 1 << 16;  // waves of the largest solve grid
unsigned int* progress_slots_pinned() {
    struct Pinned {
        unsigned int* p = nullptr;
        ~Pinned() {
            if (p) (void)hipHostFree(p);
        }
    };
    thread_local Pinned h;
    if (!h.p)
        SF_HIP(hipHostMalloc(reinterpret_cast<void**>(&h.p), sizeof(unsigned int) * kProgressSlots,
                             hipHostMallocCoherent));
    return h.p;
}
// Calls the registered callback until `solve_done` has completed, every
// interval seconds, with the sum of the published counts.
void poll_progress(const unsigned int* slots, int64_t total, hipEvent_t solve_done) {
    const ProgressReg& r = progress_reg();
    auto last = std::chrono::steady_clock::now();
    const auto step = std::chrono::duration<double>(std::max(r.interval, 0.001));
    for (;;) {
        const hipError_t q = hipEventQuery(solve_done);
        if (q != hipErrorNotReady) {
            if (q != hipSuccess) SF_HIP(q);
            break;
        }
        std::this_thread::sleep_for(std::min(std::chrono::duration<double>(0.005), step));
        const auto now = std::chrono::steady_clock::now();
        if (now - last < step) continue;
        last = now;
        int64_t done = 0;
        for (int i = 0; i < kProgressSlots; ++i)
            done += reinterpret_cast<const volatile unsigned int*>(slots)[i];
        r.fn(std::min(done, total), total, r.user);
    }
}

// A small pinned host buffer per thread for device scalars read mid-call.
double* pinned_scalars() {
    struct Pinned {
        double* p = nullptr;
        ~Pinned() {
            if (p) (void)hipHostFree(p);
        }
    };
    thread_local Pinned h;
    if (!h.p) SF_HIP(hipHostMalloc(reinterpret_cast<void**>(&h.p), 4 * sizeof(double)));
    return h.p;
}

// Refresh period of the mixed-precision kernel: the caller's, or (<= 0) the
// default for the arithmetic. Measured on the whole-brain fit, FP16 with split
// increments keeps its stopping accuracy up to a period of 40 (it drops at 80);
// BF16 rounds dy to 8 bits and is best refreshed every 10.
int refresh_period(int requested, int precision, int Kp) {
    if (requested >= 1) return requested;
    return (precision == SHOREFIT_PREC_FP16 && Kp <= kMixedSplitMaxKp) ? 40 : 10;
}

// Momentum table (t_k - 1) / t_{k+1}, t_0 = 1 (src/solver/fista.rs:111-112),
// in FP64, rounded to FP32; max(max_iter, 1) entries, on the device.
void upload_momentum(DevBuf<float>& mom, uint32_t max_iter, hipStream_t s) {
    const size_t n = std::max<uint32_t>(max_iter, 1);
    auto& stage = pinned_stage<1>();
    float* host = static_cast<float*>(stage.get(sizeof(float) * n));
    double t = 1.0;
    for (size_t k = 0; k < n; ++k) {
        const double tn = 0.5 * (1.0 + std::sqrt(1.0 + 4.0 * t * t));
        host[k] = float((t - 1.0) / tn);
        t = tn;
    }
    mom.alloc(n);
    SF_HIP(hipMemcpyAsync(mom.p, host, sizeof(float) * n, hipMemcpyHostToDevice, s));
    stage.queued(s);
}

// FISTA set-up (src/solver/fista.rs:38-62) and right-hand sides. The caller
// fills L's signal fields (data, vox, n_vox, N, sig_stride, K, Kp, Np),
// precision, max_iter and path_n. In device order: G = A^T A, its largest
// eigenvalue, L and c = 1 / (L N) on the device, the FP64 right-hand sides
// b = c A^T s, then the operands that need c on the host. The host waits for
// L on an event while the right-hand-side kernel runs.
void setup_fista(Design& d, FistaOps& ops, FistaWork& w, FistaLaunch& L, hipStream_t s) {
    const int64_t V = L.n_vox;
    const bool table = L.precision != SHOREFIT_PREC_FP32 || L.path_n > 0 || L.persist;
    if (table) {
        // uploaded first, while the stream holds little work
        upload_momentum(w.mom, L.max_iter, s);
        L.mom = w.mom.p;
    }
    d.G.alloc(size_t(d.K) * d.K);
    launch_gram(d.A.p, d.N, d.K, d.G.p, s);
    w.lam.alloc(1);
    w.sc.alloc(2);
    launch_power_iteration(d.G.p, d.K, w.lam.p, s);
    launch_fit_scalars(w.lam.p, d.N, w.sc.p, s);
    double* hs = pinned_scalars();
    SF_HIP(hipMemcpyAsync(hs, w.sc.p, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
    hipEvent_t ev;
    SF_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    struct EventGuard {
        hipEvent_t e;
        ~EventGuard() { (void)hipEventDestroy(e); }
    } guard{ev};
    SF_HIP(hipEventRecord(ev, s));
    ops.Ap.alloc(size_t(d.N) * d.Kp);
    launch_pad_cols(d.A.p, d.N, d.K, d.Kp, ops.Ap.p, s);
    w.B.alloc(size_t(V) * d.Kp);
    w.mean.alloc(V);
    L.A64p = ops.Ap.p;
    L.c_dev = w.sc.p + 1;
    L.Bbuf = w.B.p;
    L.meanbuf = w.mean.p;
    launch_fista_rhs(L, s);
    SF_HIP(hipEventSynchronize(ev));
    d.L = hs[0];
    // operands that take c on the host
    ops.step = 1.0 / d.L;
    const double c = ops.step / double(d.N);
    ops.M.alloc(size_t(d.Kp) * d.Kp);
    ops.A.alloc(size_t(d.Np) * d.Kp);
    launch_pack_M(d.G.p, d.K, d.Kp, c, ops.M.p, s);
    launch_pack_rowblock(d.A.p, d.N, d.K, d.K, /*transpose=*/false, 1.0, d.Np, d.Kp, ops.A.p, s);
    if (d.Kp / 16 <= kTik16MaxMB && d.Np / 16 <= kTik16MaxBlocks) {
        // A's chunks of the split-FP16 residual pass (fixed count, two zero chunks after them)
        const int G2 = kTik16MaxBlocks / 2;
        ops.AH16.alloc(size_t(G2 + 2) * 4 * ((d.Kp / 16 + 1) / 2) * 64 * 8);
        ops.Ainv16.alloc(size_t(d.Kp) + 32 * G2);
        launch_pack_tik16(nullptr, d.A.p, d.K, d.N, d.Kp / 16, 0, G2, ops.Ainv16.p, ops.AH16.p, s);
    }
    if (L.precision != SHOREFIT_PREC_FP32) {
        ops.Mlow.alloc(size_t(d.Kp) * d.Kp);
        launch_pack_M_lowp(d.G.p, d.K, d.Kp, c, L.precision == SHOREFIT_PREC_BF16, ops.Mlow.p, s);
        if (L.precision == SHOREFIT_PREC_FP16) {
            ops.Mlow2.alloc(size_t(d.Kp) * d.Kp);
            launch_pack_M_lowp(d.G.p, d.K, d.Kp, c, false, ops.Mlow2.p, s, 1);
        }
        if (d.Kp <= kMixedSplitMaxKp) {
            ops.Mlow3.alloc(size_t(d.Kp) * d.Kp);
            launch_pack_M_lowp(d.G.p, d.K, d.Kp, c, L.precision == SHOREFIT_PREC_BF16, ops.Mlow3.p, s, 2);
        }
    }
    L.Mfrag = ops.M.p;
    L.Afrag = ops.A.p;
    L.AH16 = ops.AH16.p;
    L.Ainv16 = ops.Ainv16.p;
    L.Mlow = ops.Mlow.p;
    L.Mlow2 = ops.Mlow2.p;
    L.Mlow3 = ops.Mlow3.p;
    L.step = float(ops.step);
    L.step64 = ops.step;
    if (L.path_n > 0) {
        w.Xpath.alloc(size_t(V) * L.path_n * d.Kp);
        w.itspath.alloc(size_t(V) * L.path_n);
        w.convpath.alloc(size_t(V) * L.path_n);
        L.Xpath = w.Xpath.p;
        L.itspath = w.itspath.p;
        L.convpath = w.convpath.p;
    }
    if (L.precision == SHOREFIT_PREC_FP32 && !L.persist) return;
    w.counter.alloc(1);
    L.counter = w.counter.p;
    if (L.precision == SHOREFIT_PREC_FP32 && L.path_n > 0) return;  // the path points hold the results
    w.X.alloc(size_t(V) * d.Kp);
    w.its.alloc(V);
    w.status.alloc(V);
    w.alpha.alloc(V);
    L.counter = w.counter.p;
    L.Xbuf = w.X.p;
    L.itsbuf = w.its.p;
    L.statusbuf = w.status.p;
    L.alphabuf = w.alpha.p;
}

// P = (A^T A + lambda_n D_n + lambda_l D_l)^-1 A^T in FP64 (Cholesky on the
// device), packed as the FP32 prologue operand and, where the split-FP16
// apply takes the shape, as its hi / lo chunks.
struct TikOps {
    DevBuf<float> P, A;
    DevBuf<uint16_t> H16;
    DevBuf<float> inv16;
};

void set_tik_ops(TikhonovLaunch& L, const TikOps& ops) {
    L.Pfrag = ops.P.p;
    L.Afrag = ops.A.p;
    L.H16 = ops.H16.p;
    L.inv16 = ops.inv16.p;
}

void pack_tikhonov(const Design& d, const double* dn, const double* dl, double lam_n, double lam_l, TikOps& ops,
                   hipStream_t s);

hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

template <typename F>
int guarded(F&& f) {
    try {
        set_error("");
        f();
        return SHOREFIT_OK;
    } catch (const Error& e) {
        set_error(e.msg);
        return e.code;
    } catch (const std::exception& e) {
        set_error(e.what());
        return SHOREFIT_E_DEVICE;
    }
}

}  // namespace

// FP64 host-side triangular solves are avoided: the columns of A^T are solved
// against the Cholesky factor on the device.
void launch_chol_solve_columns(const double* d_L, int K, const double* d_A, int N, double* d_P, hipStream_t s);

namespace {
void pack_tikhonov(const Design& d, const double* dn, const double* dl, double lam_n, double lam_l, TikOps& ops,
                   hipStream_t s) {
    // src/solver/tikhonov.rs:33-44: G + lambda_N D_N + lambda_L D_L, on the device
    std::vector<double> diag(size_t(2) * d.K, 0.0);
    for (int i = 0; i < d.K; ++i) {
        if (dn) diag[i] = dn[i];
        if (dl) diag[d.K + i] = dl[i];
    }
    DevBuf<double> ddiag(diag.size());
    SF_HIP(hipMemcpyAsync(ddiag.p, diag.data(), sizeof(double) * diag.size(), hipMemcpyHostToDevice, s));
    DevBuf<double> Lf(size_t(d.K) * d.K);
    DevBuf<int> status(1);
    launch_add_diagonals(d.G.p, d.K, dn ? ddiag.p : nullptr, lam_n, dl ? ddiag.p + d.K : nullptr, lam_l, Lf.p, s);
    launch_cholesky(Lf.p, d.K, status.p, s);
    int st = 0;
    SF_HIP(hipMemcpyAsync(&st, status.p, sizeof(int), hipMemcpyDeviceToHost, s));
    SF_HIP(hipStreamSynchronize(s));
    if (st != 0)
        throw Error{SHOREFIT_E_NOT_PD,
                    "normal-equations matrix should be PD (Cholesky pivot " + std::to_string(st - 1) +
                        " is not positive)"};
    DevBuf<double> P(size_t(d.K) * d.N);  // K x N row-major
    launch_chol_solve_columns(Lf.p, d.K, d.A.p, d.N, P.p, s);
    ops.P.alloc(size_t(d.Kp) * d.Np);
    ops.A.alloc(size_t(d.Np) * d.Kp);
    launch_pack_rowblock(P.p, d.K, d.N, d.N, /*transpose=*/false, 1.0, d.Kp, d.Np, ops.P.p, s);
    launch_pack_rowblock(d.A.p, d.N, d.K, d.K, /*transpose=*/false, 1.0, d.Np, d.Kp, ops.A.p, s);
    const int MB = d.Kp / 16, NC = d.Np / 16;
    if (MB <= kTik16MaxMB && NC <= kTik16MaxBlocks) {
        // the apply kernel's fixed chunk counts: G2 chunks of two sample
        // blocks, G1 (even) chunks of P, two zero chunks after them
        const int G2 = kTik16MaxBlocks / 2, G1 = (G2 + 1) & ~1;
        ops.H16.alloc(size_t(G1 + G2 + 2) * 4 * ((MB + 1) / 2) * 64 * 8);
        ops.inv16.alloc(size_t(d.Kp) + 32 * G2);
        launch_pack_tik16(P.p, d.A.p, d.K, d.N, MB, G1, G2, ops.inv16.p, ops.H16.p, s);
    }
    SF_HIP(hipStreamSynchronize(s));
}

void regularization_diagonals(uint32_t order, std::vector<double>& dn, std::vector<double>& dl) {
    // src/basis/shore.rs:180-195
    dn.clear();
    dl.clear();
    for (uint32_t n = 0; n <= order; ++n)
        for (uint32_t l = 0; l <= n; l += 2)
            for (uint32_t m = 0; m < 2 * l + 1; ++m) {
                const double nn = double(n * (n + 1)), ll = double(l * (l + 1));
                dn.push_back(nn * nn);
                dl.push_back(ll * ll);
            }
}

}  // namespace
}  // namespace shorefit

using namespace shorefit;

namespace {
// Device copy of a host array (or the device pointer itself).
template <typename T>
const T* dev_in(const T* p, size_t count, bool on_device, DevBuf<T>& buf, hipStream_t s) {
    if (!p || on_device) return p;
    buf.alloc(count);
    if (count) SF_HIP(hipMemcpyAsync(buf.p, p, sizeof(T) * count, hipMemcpyHostToDevice, s));
    return buf.p;
}

// Masked-voxel list on the device (null = every voxel); returns the count.
int64_t voxel_list(const uint8_t* d_mask, int64_t n, DevBuf<int64_t>& vox, hipStream_t s) {
    if (!d_mask) return n;
    return compact_mask(d_mask, n, vox, s);
}

void check_dims(int64_t nx, int64_t ny, int64_t nz) {
    if (nx < 0 || ny < 0 || nz < 0) throw Error{SHOREFIT_E_INVALID, "negative volume dimension"};
}
// ------------------------------------------------ gradient-deviation fits --
// Tuned variant per (radial order, N), filled by shorefit_graddev_tune.
std::mutex& tuned_mutex() {
    static std::mutex m;
    return m;
}
std::map<std::pair<uint32_t, int64_t>, int>& tuned_variants() {
    static std::map<std::pair<uint32_t, int64_t>, int> t;
    return t;
}

// The requested variant, else the tuned one for (order, N), else the first
// (the default) of the order whose Gram kernel fits its LDS for N gradients.
const GdVariant& pick_variant(uint32_t order, int64_t N, int32_t requested) {
    const GdVariant* vs = nullptr;
    const int nv = gd_variants(order, &vs);
    if (nv == 0)
        throw Error{SHOREFIT_E_UNSUPPORTED,
                    "gradient-deviation fits support even radial orders 2 to 10, got " + std::to_string(order)};
    if (requested >= 0) {
        if (requested >= nv)
            throw Error{SHOREFIT_E_INVALID, "variant " + std::to_string(requested) + " out of range (radial order " +
                                                std::to_string(order) + " has " + std::to_string(nv) + ")"};
        return vs[requested];
    }
    std::lock_guard<std::mutex> lk(tuned_mutex());
    const auto it = tuned_variants().find({order, N});
    if (it != tuned_variants().end()) return vs[it->second];
    for (int i = 0; i < nv; ++i)
        if (gd_gram_fits(vs[i], int(N))) return vs[i];
    return vs[0];  // none fits: the launch reports it
}

// Identity voxels (max |T - I| < 1e-6) of the voxel list into id, the others
// into dv; both in list order. Returns the identity count.
int64_t gd_partition(const float* field, const int64_t* vox, int64_t V, DevBuf<int64_t>& id, DevBuf<int64_t>& dv,
                     int64_t* n_dv, hipStream_t s) {
    DevBuf<uint8_t> fid(V), fdv(V);
    gd_launch_identity(field, vox, V, fid.p, fdv.p, s);
    DevBuf<int64_t> pid, pdv;
    const int64_t n_id = V ? compact_mask(fid.p, V, pid, s) : 0;
    *n_dv = V ? compact_mask(fdv.p, V, pdv, s) : 0;
    id.alloc(std::max<int64_t>(n_id, 1));
    dv.alloc(std::max<int64_t>(*n_dv, 1));
    launch_gather_i64(vox, pid.p, n_id, id.p, s);
    launch_gather_i64(vox, pdv.p, *n_dv, dv.p, s);
    return n_id;
}

struct GdSetup {
    GdBasis B;
    GdTable tab;
    DevBuf<double> bvals, gvox, gnom;
};

void gd_setup(GdSetup& S, const ShorefitGradients& g, double big, double small, const ShorefitGradDev& gd,
              uint32_t order, double zeta, hipStream_t s) {
    const int N = int(g.n_grads);
    S.B = gd_basis_host(order, zeta);
    S.bvals.alloc(N);
    S.gvox.alloc(size_t(3) * N);
    S.gnom.alloc(size_t(3) * N);
    SF_HIP(hipMemcpyAsync(S.bvals.p, g.bvals, sizeof(double) * N, hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(S.gvox.p, gd.bvecs_vox ? gd.bvecs_vox : g.bvecs, sizeof(double) * 3 * N,
                          hipMemcpyHostToDevice, s));
    SF_HIP(hipMemcpyAsync(S.gnom.p, g.bvecs, sizeof(double) * 3 * N, hipMemcpyHostToDevice, s));
    S.tab.bvals = S.bvals.p;
    S.tab.gvox = S.gvox.p;
    S.tab.gnom = S.gnom.p;
    for (int k = 0; k < 9; ++k) S.tab.R[k] = gd.frame_rotation[k];
    S.tab.b0_thr = g.b0_threshold;
    const double tau = big - small / 3.0;
    S.tab.qdenom = 4.0 * kPi * kPi * tau;  // src/qspace.rs:130-149
    S.tab.N = N;
}

struct GdRunOut {
    float* coef;
    float* resid;
    float* r2;
    uint32_t* iters;
    uint8_t* kind;
    float* alpha_map;
    float* bic;
};

// Workspace per chunk: G_v (FP32 Kp x Kp) and b_v (FP64) of every voxel of
// the chunk; 8 GiB unless SHOREFIT_GRADDEV_WORKSPACE_MB says otherwise.
int64_t gd_chunk_voxels(const GdVariant& v, int64_t V) {
    size_t cap = size_t(8) << 30;
    if (const char* e = std::getenv("SHOREFIT_GRADDEV_WORKSPACE_MB")) {
        const long long mb = std::atoll(e);
        if (mb > 0) cap = size_t(mb) << 20;
    }
    const size_t per = size_t(v.kp) * v.kp * sizeof(float) + size_t(v.kp) * sizeof(double) + 16;
    return std::max<int64_t>(1, std::min<int64_t>(V, int64_t(cap / per)));
}

// The deviated voxels vox[0..V): Gram, FISTA and residual kernels per chunk.
// lip (optional, device, V entries) receives each voxel's Lipschitz constant;
// phase_ms (optional) accumulates the three kernels' device times.
void gd_run(const GdSetup& S, const GdVariant& var, const ShorefitSpec& sp, const float* data, int64_t sig_stride,
            const float* field, const int64_t* vox, int64_t V, const GdRunOut& o, double* lip, hipStream_t s,
            double* phase_ms = nullptr) {
    if (V <= 0) return;
    const int N = S.tab.N, K = S.B.K, KP = var.kp;
    DevBuf<float> d_mom;
    upload_momentum(d_mom, sp.max_iter, s);
    const int64_t Vb = gd_chunk_voxels(var, V);
    DevBuf<float> G(size_t(Vb) * KP * KP);
    DevBuf<double> b(size_t(Vb) * KP), sstat(size_t(Vb) * 2);
    DevBuf<GdSlotOut> so(Vb);
    DevBuf<unsigned> ctr(1);
    hipEvent_t ev[3] = {};
    if (phase_ms)
        for (auto& e : ev) SF_HIP(hipEventCreate(&e));
    for (int64_t c0 = 0; c0 < V; c0 += Vb) {
        const int64_t n = std::min(Vb, V - c0);
        GdChunk c{S.B, S.tab, data, sig_stride, field, vox + c0, n, G.p, b.p, sstat.p};
        if (phase_ms) SF_HIP(hipEventRecord(ev[0], s));
        gd_launch_gram(var, c, s);
        if (phase_ms) SF_HIP(hipEventRecord(ev[1], s));
        SF_HIP(hipMemsetAsync(ctr.p, 0, sizeof(unsigned), s));
        GdFistaArgs fa{};
        fa.G = G.p;
        fa.b = b.p;
        fa.sstat = sstat.p;
        fa.vox = vox + c0;
        fa.n = n;
        fa.N = N;
        fa.K = K;
        fa.alpha = sp.alpha;
        fa.ratio = sp.alpha_mode == SHOREFIT_ALPHA_RATIO ? sp.alpha_ratio : 0.0;
        fa.max_iter = sp.max_iter;
        fa.tol = sp.tol;
        fa.nonneg = sp.non_negative;
        fa.mom = d_mom.p;
        fa.counter = ctr.p;
        fa.coef = o.coef;
        fa.slot_out = so.p;
        fa.lip = lip ? lip + c0 : nullptr;
        gd_launch_fista(var, fa, 0, s);
        const GdMaps maps{o.resid, o.r2, o.iters, o.kind, o.alpha_map, o.bic};
        gd_launch_finish(so.p, sstat.p, vox + c0, n, N, maps, s);
        if (phase_ms) {
            SF_HIP(hipEventRecord(ev[2], s));
            SF_HIP(hipEventSynchronize(ev[2]));
            for (int k = 0; k < 2; ++k) {
                float ms = 0.f;
                SF_HIP(hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
                phase_ms[k] += ms;
            }
        }
    }
    if (phase_ms)
        for (auto& e : ev) (void)hipEventDestroy(e);
    SF_HIP(hipStreamSynchronize(s));  // the chunk buffers go out of scope
}

// The volume fit behind shorefit_fit and shorefit_fit_graddev: validation,
// gradient table, voxel list, the shared-design solve and, with a
// gradient-deviation field, the per-voxel-design solve of the deviated
// voxels.
void fit_impl(const ShorefitGradients* grads, const ShorefitVolume* vol, const ShorefitSpec* spec,
              ShorefitOutput* out, hipStream_t s, const ShorefitGradDev* gd) {
    out->graddev_ms[0] = out->graddev_ms[1] = out->graddev_ms[2] = 0.0;
    const ShorefitSpec& sp = *spec;
    Warnings warn;

    // --- spec validation (build_alpha_strategy, src/fit.rs:555-591; the
    // binding's own checks, python/src/models.rs:117-147) -------------
    if (sp.regularization == SHOREFIT_REG_NONNEG)
        throw Error{SHOREFIT_E_UNSUPPORTED, "regularization 'nonneg' (ICLS) is not implemented on this device"};
    if (sp.regularization != SHOREFIT_REG_L1 && sp.regularization != SHOREFIT_REG_L2)
        throw Error{SHOREFIT_E_INVALID, "regularization must be l1, l2 or nonneg"};
    double path_eps = sp.path_eps;
    if (sp.regularization == SHOREFIT_REG_L1) {
        switch (sp.alpha_mode) {
            case SHOREFIT_ALPHA_FIXED: break;
            case SHOREFIT_ALPHA_RATIO:
                if (!(sp.alpha_ratio > 0.0 && sp.alpha_ratio < 1.0))
                    throw Error{SHOREFIT_E_INVALID,
                                "--alpha-ratio must lie in (0, 1), got " + std::to_string(sp.alpha_ratio)};
                break;
            case SHOREFIT_ALPHA_PATH_BIC:
            case SHOREFIT_ALPHA_L2_ANCHORED: {
                if (sp.path_n_alphas < 2)
                    throw Error{SHOREFIT_E_INVALID,
                                "--path-n-alphas must be ≥ 2, got " + std::to_string(sp.path_n_alphas)};
                if (!(path_eps > 0.0)) path_eps = sp.alpha_mode == SHOREFIT_ALPHA_PATH_BIC ? 1e-3 : 1e-4;
                if (!(path_eps > 0.0 && path_eps < 1.0))
                    throw Error{SHOREFIT_E_INVALID, "--path-eps must lie in (0, 1), got " + std::to_string(path_eps)};
                if (sp.alpha_mode == SHOREFIT_ALPHA_L2_ANCHORED && sp.slack < 0.0)
                    throw Error{SHOREFIT_E_INVALID, "--slack must be ≥ 0, got " + std::to_string(sp.slack)};
                break;
            }
            default:
                throw Error{SHOREFIT_E_INVALID, "alpha_mode must be fixed, alpha-ratio, path-bic or l2-anchored"};
        }
    }
    if (sp.radial_order % 2 != 0 || sp.radial_order == 0)
        throw Error{SHOREFIT_E_INVALID,
                    "radial_order must be a positive even number, got " + std::to_string(sp.radial_order)};
    if (sp.precision != SHOREFIT_PREC_FP32 && sp.precision != SHOREFIT_PREC_FP16 &&
        sp.precision != SHOREFIT_PREC_BF16)
        throw Error{SHOREFIT_E_INVALID, "precision must be fp32, fp16 or bf16"};
    const GdVariant* gd_var = nullptr;
    if (gd) {
        if (!gd->field) throw Error{SHOREFIT_E_INVALID, "the gradient-deviation field is required"};
        if (sp.regularization != SHOREFIT_REG_L1 ||
            (sp.alpha_mode != SHOREFIT_ALPHA_FIXED && sp.alpha_mode != SHOREFIT_ALPHA_RATIO))
            throw Error{SHOREFIT_E_UNSUPPORTED,
                        "gradient-deviation fits support regularization l1 with alpha_mode fixed or alpha-ratio"};
        gd_var = &pick_variant(sp.radial_order, grads->n_grads, gd->variant);
    }

    // --- gradient table (src/qspace.rs:90-120) ------------------------
    const int64_t N = grads->n_grads;
    if (N <= 0 || !grads->bvals || !grads->bvecs) throw Error{SHOREFIT_E_INVALID, "empty gradient table"};
    if (vol->nx < 0 || vol->ny < 0 || vol->nz < 0) throw Error{SHOREFIT_E_INVALID, "negative volume shape"};
    double big = grads->big_delta, small = grads->small_delta;
    out->delta_source = 0;
    if (std::isnan(big) || std::isnan(small)) {
        tortoise_deltas(grads->bvals, N, &big, &small);
        out->delta_source = 1;
    }
    out->big_delta = big;
    out->small_delta = small;
    out->tau = big - small / 3.0;

    // --- fit_shore refusals and warnings (src/fit.rs:691-730) ---------
    const auto shells = dwi_shells(grads->bvals, N, grads->b0_threshold);
    out->n_shells = int32_t(shells.size());
    for (size_t i = 0; i < shells.size() && i < 32; ++i) out->dwi_shells[i] = shells[i];
    if (shells.empty())
        throw Error{SHOREFIT_E_FIT,
                    "3D-SHORE needs diffusion-weighted data; every volume is at or below the b=0 threshold"};
    if (shells.size() < kMinShoreShells) {
        const std::string what = shells.size() == 1 ? "a single shell (b = " + fmt_b(shells[0]) + " s/mm²)"
                                                    : std::to_string(shells.size()) + " shells";
        if (!sp.allow_single_shell)
            throw Error{
                SHOREFIT_E_FIT,
                "3D-SHORE needs diffusion-weighted data on at least 2 b-value shells; this series has " + what +
                    ". With a single shell the radial decay of the signal is not determined by the data: a fit "
                    "would match the measurements, but its propagator-derived scalars (RTOP, RTAP, RTPP, MSD, "
                    "QIV) and the signals it predicts at other b-values would reflect the regularization rather "
                    "than the data. For single-shell data use single-shell three-tissue CSD (cs-ss3t-full, "
                    "cs_dmri.ss3t_pipeline). To fit anyway for orientation information only, allow single-shell "
                    "data (--allow-single-shell, or allow_single_shell=True in Python)."};
        warn.msgs.push_back("the data have " + what +
                            ": only the orientation information of this fit is meaningful, and "
                            "propagator-derived scalars are not computed");
    }
    const int K = n_coeffs(sp.radial_order);
    out->n_coeffs = K;
    if (N < K && sp.regularization != SHOREFIT_REG_L1)
        warn.msgs.push_back(std::to_string(N) + " measurements for " + std::to_string(K) +
                            " coefficients (radial order " + std::to_string(sp.radial_order) +
                            "): without L1 regularization the fit is underdetermined and its fit statistics "
                            "are not meaningful; use a lower radial order or L1");

    // --- device buffers -------------------------------------------------
    // Queued ahead of the mask compaction, whose voxel count the host waits
    // for: the device clears the outputs and builds the design meanwhile.
    const int64_t nvox_total = vol->nx * vol->ny * vol->nz;
    const size_t coef_bytes = sizeof(float) * size_t(nvox_total) * K;
    DevBuf<float> d_data, d_coef, d_r2, d_res, d_alpha, d_bic, d_rssl2;
    DevBuf<uint32_t> d_it;
    DevBuf<uint8_t> d_kind;
    const bool dev = vol->on_device != 0;
    const float* data = vol->data;
    if (!dev) {
        d_data.alloc(size_t(nvox_total) * N);
        if (nvox_total)
            SF_HIP(hipMemcpyAsync(d_data.p, vol->data, sizeof(float) * nvox_total * N, hipMemcpyHostToDevice, s));
        data = d_data.p;
    }
    auto out_buf = [&](auto* user, auto& buf, size_t count) {
        using T = std::remove_pointer_t<decltype(user)>;
        if (!user) return static_cast<T*>(nullptr);
        if (dev) return user;
        buf.alloc(count);
        return buf.p;
    };
    const bool l1 = sp.regularization == SHOREFIT_REG_L1;
    float* coef = out_buf(out->coefficients, d_coef, size_t(nvox_total) * K);
    float* r2 = out_buf(out->r2, d_r2, nvox_total);
    float* res = out_buf(out->residual_l2, d_res, nvox_total);
    uint32_t* its = out_buf(out->iterations, d_it, nvox_total);
    uint8_t* kind = out_buf(out->regularization_kind, d_kind, nvox_total);
    float* amap = l1 ? out_buf(out->alpha, d_alpha, nvox_total) : nullptr;
    float* bic = l1 ? out_buf(out->bic, d_bic, nvox_total) : nullptr;
    const bool path = l1 && (sp.alpha_mode == SHOREFIT_ALPHA_PATH_BIC || sp.alpha_mode == SHOREFIT_ALPHA_L2_ANCHORED);
    const bool anchored = l1 && sp.alpha_mode == SHOREFIT_ALPHA_L2_ANCHORED;
    // L2-anchored: the Tikhonov reference RSS per voxel (also the rss_l2 map)
    float* rssl2 = anchored ? out_buf(out->rss_l2, d_rssl2, nvox_total) : nullptr;
    if (anchored && !rssl2) {
        d_rssl2.alloc(nvox_total);
        rssl2 = d_rssl2.p;
    }
    DevBuf<float> d_alpha_tmp;
    const bool per_voxel_alpha = l1 && sp.alpha_mode != SHOREFIT_ALPHA_FIXED;
    if (per_voxel_alpha && !amap) {
        d_alpha_tmp.alloc(nvox_total);
        amap = d_alpha_tmp.p;
    }
    // zeros outside the mask (src/fit.rs:89-101)
    if (nvox_total) {
        SF_HIP(hipMemsetAsync(coef, 0, coef_bytes, s));
        if (r2) SF_HIP(hipMemsetAsync(r2, 0, sizeof(float) * nvox_total, s));
        if (res) SF_HIP(hipMemsetAsync(res, 0, sizeof(float) * nvox_total, s));
        if (its) SF_HIP(hipMemsetAsync(its, 0, sizeof(uint32_t) * nvox_total, s));
        if (kind) SF_HIP(hipMemsetAsync(kind, 0, nvox_total, s));
        if (amap) SF_HIP(hipMemsetAsync(amap, 0, sizeof(float) * nvox_total, s));
        if (bic) SF_HIP(hipMemsetAsync(bic, 0, sizeof(float) * nvox_total, s));
    }
    if (out->rss_l2 && dev && nvox_total) SF_HIP(hipMemsetAsync(out->rss_l2, 0, sizeof(float) * nvox_total, s));
    if (rssl2 && nvox_total) SF_HIP(hipMemsetAsync(rssl2, 0, sizeof(float) * nvox_total, s));

    // --- design + solver set-up (src/fit.rs:718-759) -------------------
    const int precision = (l1 && sp.max_iter > 0) ? sp.precision : SHOREFIT_PREC_FP32;
    Design d;
    build_design(d, *grads, big, small, sp.radial_order, sp.zeta, s, precision != SHOREFIT_PREC_FP32);
    // --- masked voxels in x -> y -> z order (src/voxel_loop.rs:28-42) --
    // A device mask is compacted on the device; a host mask on the host;
    // without a mask every voxel is fitted in place (no index list).
    DevBuf<int64_t> d_vox;
    int64_t V = 0;
    if (!vol->mask) {
        V = nvox_total;
    } else if (vol->on_device) {
        V = compact_mask(vol->mask, nvox_total, d_vox, s);
    } else {
        std::vector<int64_t> vox;
        vox.reserve(nvox_total);
        for (int64_t i = 0; i < nvox_total; ++i)
            if (vol->mask[i]) vox.push_back(i);
        V = int64_t(vox.size());
        d_vox.alloc(V);
        if (V) SF_HIP(hipMemcpyAsync(d_vox.p, vox.data(), sizeof(int64_t) * V, hipMemcpyHostToDevice, s));
        SF_HIP(hipStreamSynchronize(s));  // vox is released at scope exit
    }
    const int64_t* vox_dev = d_vox.p;  // null: voxel v is row v
    out->n_voxels = V;

    // --- gradient deviation: identity voxels keep the shared design -------
    const int64_t* vox_shared = vox_dev;
    int64_t V_shared = V;
    DevBuf<float> d_field;
    DevBuf<int64_t> d_vox_id, d_vox_dv;
    const float* field = nullptr;
    int64_t V_dv = 0;
    if (gd) {
        field = dev_in(gd->field, size_t(nvox_total) * 9, vol->on_device != 0, d_field, s);
        V_shared = gd_partition(field, vox_dev, V, d_vox_id, d_vox_dv, &V_dv, s);
        vox_shared = d_vox_id.p;
    }

    hipEvent_t e0, e1, e2, e3;
    SF_HIP(hipEventCreate(&e0));
    SF_HIP(hipEventCreate(&e1));
    SF_HIP(hipEventCreate(&e2));
    SF_HIP(hipEventCreate(&e3));
    if (l1) {
        FistaOps ops;
        FistaLaunch L{};
        FistaWork work;
        L.data = data;
        L.vox = vox_shared;
        L.n_vox = V_shared;
        L.N = d.N;
        L.sig_stride = N;
        L.K = d.K;
        L.Kp = d.Kp;
        L.Np = d.Np;
        L.ratio = sp.alpha_mode == SHOREFIT_ALPHA_RATIO ? float(sp.alpha_ratio) : 0.f;
        L.alpha = sp.alpha;
        L.max_iter = sp.max_iter;
        L.tol = float(sp.tol);
        L.nonneg = sp.non_negative;
        L.precision = precision;
        L.persist = precision == SHOREFIT_PREC_FP32;  // FP32: the persistent form (no alpha path)
        L.refresh = refresh_period(sp.refresh_every, precision, d.Kp);
        L.coef = coef;
        L.resid = res;
        L.r2 = r2;
        L.iters = its;
        L.kind = kind;
        L.alpha_map = amap;
        L.bic = bic;
        // alpha path: f_i = exp(t_i ln eps), t_i = i / (n - 1), so that
        // alpha_i = alpha_max f_i (src/solver/alpha.rs:57-69)
        DevBuf<float> d_pf;
        DevBuf<double> d_pf64;
        TikOps tik;
        if (path) {
            double eps = sp.path_eps;
            if (!(eps > 0.0)) eps = sp.alpha_mode == SHOREFIT_ALPHA_PATH_BIC ? 1e-3 : 1e-4;
            const int n = int(sp.path_n_alphas);
            std::vector<double> f64(n);
            std::vector<float> f32(n);
            for (int i = 0; i < n; ++i) {
                f64[i] = std::exp(double(i) / double(n - 1) * std::log(eps));
                f32[i] = float(f64[i]);
            }
            d_pf.alloc(n);
            d_pf64.alloc(n);
            SF_HIP(hipMemcpyAsync(d_pf.p, f32.data(), sizeof(float) * n, hipMemcpyHostToDevice, s));
            SF_HIP(hipMemcpyAsync(d_pf64.p, f64.data(), sizeof(double) * n, hipMemcpyHostToDevice, s));
            SF_HIP(hipStreamSynchronize(s));
            L.path_n = n;
            L.path_f = d_pf.p;
            L.path_f64 = d_pf64.p;
            L.path_rule = anchored ? 1 : 0;
            L.slack = sp.slack;
            L.rss_l2_ref = rssl2;
        }
        setup_fista(d, ops, work, L, s);
        out->lipschitz = d.L;
        L.thr = float(ops.step * sp.alpha);
        if (anchored && V > 0) {
            // the Tikhonov reference fit (src/fit.rs:404-410, 760-761)
            std::vector<double> dn, dl;
            regularization_diagonals(sp.radial_order, dn, dl);
            pack_tikhonov(d, dn.data(), dl.data(), sp.lambda_n, sp.lambda_l, tik, s);
            TikhonovLaunch T{};
            T.data = data;
            T.vox = vox_dev;
            T.n_vox = V;
            T.N = d.N;
            T.sig_stride = N;
            T.K = d.K;
            T.Kp = d.Kp;
            T.Np = d.Np;
            set_tik_ops(T, tik);
            T.rss_l2 = rssl2;
            launch_tikhonov(T, s);
        }
        SF_HIP(hipEventRecord(e0, s));
        L.event_mid = e1;  // recorded right after the solve kernel
        const bool report = progress_reg().fn && V_shared > 0;
        if (report) {
            L.progress = progress_slots_pinned();
            L.progress_slots = kProgressSlots;
            std::memset(L.progress, 0, sizeof(unsigned int) * kProgressSlots);
            progress_reg().fn(0, V, progress_reg().user);
        }
        launch_fista(L, s);
        if (V_shared == 0) SF_HIP(hipEventRecord(e1, s));
        if (report) poll_progress(L.progress, V_shared, e1);
        SF_HIP(hipEventRecord(e2, s));
        if (gd && V_dv > 0) {
            // deviated voxels: per-voxel designs (src/graddev.rs:160-185)
            GdSetup S;
            gd_setup(S, *grads, big, small, *gd, sp.radial_order, sp.zeta, s);
            GdRunOut o{coef, res, r2, its, kind, amap, bic};
            gd_run(S, *gd_var, sp, data, N, field, d_vox_dv.p, V_dv, o, nullptr, s, out->graddev_ms);
        }
        SF_HIP(hipEventRecord(e3, s));
    } else {
        d.G.alloc(size_t(d.K) * d.K);
        launch_gram(d.A.p, d.N, d.K, d.G.p, s);
        std::vector<double> dn, dl;
        regularization_diagonals(sp.radial_order, dn, dl);
        TikOps ops;
        pack_tikhonov(d, dn.data(), dl.data(), sp.lambda_n, sp.lambda_l, ops, s);
        TikhonovLaunch L{};
        L.data = data;
        L.vox = vox_dev;
        L.n_vox = V;
        L.N = d.N;
        L.sig_stride = N;
        L.K = d.K;
        L.Kp = d.Kp;
        L.Np = d.Np;
        set_tik_ops(L, ops);
        L.coef = coef;
        L.resid = res;
        L.r2 = r2;
        L.iters = its;
        L.kind = kind;
        if (progress_reg().fn) progress_reg().fn(0, V, progress_reg().user);
        SF_HIP(hipEventRecord(e0, s));
        launch_tikhonov(L, s);
        SF_HIP(hipEventRecord(e1, s));
        SF_HIP(hipEventRecord(e2, s));
        SF_HIP(hipEventRecord(e3, s));
    }
    SF_HIP(hipStreamSynchronize(s));
    if (progress_reg().fn) progress_reg().fn(V, V, progress_reg().user);
    float ms = 0.f, ms_gd = 0.f;
    SF_HIP(hipEventElapsedTime(&ms, e0, e1));
    SF_HIP(hipEventElapsedTime(&ms_gd, e2, e3));
    out->fit_ms = double(ms) + double(ms_gd);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipEventDestroy(e2);
    (void)hipEventDestroy(e3);

    // --- alpha distribution (src/fit.rs:334-352) ------------------------
    out->has_alpha_distribution = 0;
    if (per_voxel_alpha) {
        // the masked voxels' alphas in work order, then three order
        // statistics (the values a full sort would put at those ranks)
        std::vector<float> a_v(V);
        if (V) {
            DevBuf<float> d_av(V);
            launch_gather_f32(amap, vox_dev, V, d_av.p, s);
            SF_HIP(hipMemcpyAsync(a_v.data(), d_av.p, sizeof(float) * V, hipMemcpyDeviceToHost, s));
            SF_HIP(hipStreamSynchronize(s));
        }
        std::vector<double> seen(a_v.begin(), a_v.end());
        out->has_alpha_distribution = 1;
        if (seen.empty()) {
            out->alpha_distribution[0] = out->alpha_distribution[1] = out->alpha_distribution[2] = 0.0;
        } else {
            auto pct = [&](double q) {
                size_t i = std::min(size_t(std::llround(double(seen.size() - 1) * q)), seen.size() - 1);
                std::nth_element(seen.begin(), seen.begin() + i, seen.end());
                return seen[i];
            };
            out->alpha_distribution[0] = pct(0.5);
            out->alpha_distribution[1] = pct(0.10);
            out->alpha_distribution[2] = pct(0.90);
        }
    }

    // --- results back to the caller -------------------------------------
    if (!dev && nvox_total) {
        SF_HIP(hipMemcpy(out->coefficients, coef, coef_bytes, hipMemcpyDeviceToHost));
        if (out->r2) SF_HIP(hipMemcpy(out->r2, r2, sizeof(float) * nvox_total, hipMemcpyDeviceToHost));
        if (out->residual_l2)
            SF_HIP(hipMemcpy(out->residual_l2, res, sizeof(float) * nvox_total, hipMemcpyDeviceToHost));
        if (out->iterations)
            SF_HIP(hipMemcpy(out->iterations, its, sizeof(uint32_t) * nvox_total, hipMemcpyDeviceToHost));
        if (out->regularization_kind)
            SF_HIP(hipMemcpy(out->regularization_kind, kind, nvox_total, hipMemcpyDeviceToHost));
        if (l1 && out->alpha) SF_HIP(hipMemcpy(out->alpha, amap, sizeof(float) * nvox_total, hipMemcpyDeviceToHost));
        if (l1 && out->bic) SF_HIP(hipMemcpy(out->bic, bic, sizeof(float) * nvox_total, hipMemcpyDeviceToHost));
        if (anchored && out->rss_l2)
            SF_HIP(hipMemcpy(out->rss_l2, rssl2, sizeof(float) * nvox_total, hipMemcpyDeviceToHost));
    }
    warn.to(out);
}

}  // namespace

// A design held on the device between calls (shorefit_design_create).
struct ShorefitDesign {
    shorefit::Design d;
    FistaOps ops;
    FistaWork work;
};

extern "C" {

const char* shorefit_last_error(void) { return g_last_error.c_str(); }

const char* shorefit_version(void) { return "shorefit 0.1.0 (gfx950)"; }

#ifndef SHOREFIT_GIT_SHA
#define SHOREFIT_GIT_SHA "unknown"
#endif
#ifndef SHOREFIT_BUILD_TIME
#define SHOREFIT_BUILD_TIME "unknown"
#endif

const char* shorefit_build_info(void) {
    return "{\"library\": \"shorefit\", \"version\": \"0.1.0\", \"git_sha\": \"" SHOREFIT_GIT_SHA
           "\", \"build_timestamp\": \"" SHOREFIT_BUILD_TIME "\", \"offload_arch\": \"gfx950\"}";
}

int shorefit_set_progress(shorefit_progress_fn fn, void* user, double interval_seconds) {
    return guarded([&] {
        if (fn && !(interval_seconds > 0.0))
            throw Error{SHOREFIT_E_INVALID, "progress interval must be positive"};
        ProgressReg& r = progress_reg();
        r.fn = fn;
        r.user = user;
        r.interval = interval_seconds;
    });
}

int shorefit_release_cache(void) {
    return guarded([&] {
        BlockCache& c = block_cache();
        if (c.size_of.empty()) return;
        SF_HIP(hipDeviceSynchronize());
        drop_cached_blocks(c);
    });
}

int shorefit_n_coeffs(uint32_t radial_order) {
    if (radial_order % 2 != 0) return -1;
    return n_coeffs(radial_order);
}

void shorefit_spec_default(ShorefitSpec* s) {
    std::memset(s, 0, sizeof(*s));
    s->radial_order = 6;
    s->zeta = 700.0;
    s->regularization = SHOREFIT_REG_L1;
    s->alpha_mode = SHOREFIT_ALPHA_L2_ANCHORED;
    s->alpha = 1.0;
    s->alpha_ratio = 1e-3;
    s->path_n_alphas = 20;
    s->path_eps = 0.0;
    s->slack = 0.05;
    s->max_iter = 1000;
    s->tol = 1e-6;
    s->non_negative = 0;
    s->lambda_n = 1e-8;
    s->lambda_l = 1e-8;
    s->compute_diagnostics = 0;
    s->allow_single_shell = 0;
    s->precision = SHOREFIT_PREC_FP32;
    s->refresh_every = 0;
}

int shorefit_design_matrix(const ShorefitGradients* g, uint32_t radial_order, double zeta, double* A_out) {
    return guarded([&] {
        if (!g || !A_out) throw Error{SHOREFIT_E_INVALID, "null argument"};
        double big = g->big_delta, small = g->small_delta;
        if (std::isnan(big) || std::isnan(small)) tortoise_deltas(g->bvals, g->n_grads, &big, &small);
        StreamScope scope(nullptr);
        Design d;
        build_design(d, *g, big, small, radial_order, zeta, nullptr);
        SF_HIP(hipMemcpy(A_out, d.A.p, sizeof(double) * d.N * d.K, hipMemcpyDeviceToHost));
    });
}

int shorefit_fista_batch(const double* A, int32_t n_rows, int32_t n_cols, const float* signals, int64_t n_signals,
                         double alpha, uint32_t max_iter, double tol, int32_t non_negative, int32_t precision,
                         int32_t refresh_every, float* coef, uint32_t* iterations, uint8_t* converged,
                         float* residual_l2, double* lipschitz_out) {
    return guarded([&] {
        if (!A || !signals || !coef || n_rows <= 0 || n_cols <= 0 || n_signals < 0)
            throw Error{SHOREFIT_E_INVALID, "invalid fista_batch arguments"};
        hipStream_t s = nullptr;
        StreamScope scope(s);
        if (precision != SHOREFIT_PREC_FP32 && precision != SHOREFIT_PREC_FP16 && precision != SHOREFIT_PREC_BF16)
            throw Error{SHOREFIT_E_INVALID, "precision must be fp32, fp16 or bf16"};
        if (max_iter == 0) precision = SHOREFIT_PREC_FP32;
        const bool mixed = precision != SHOREFIT_PREC_FP32;
        Design d;
        upload_design(d, A, n_rows, n_cols, s, mixed);
        FistaOps ops;
        const int64_t V = n_signals;
        DevBuf<float> S(size_t(V) * n_rows), C(size_t(V) * n_cols), R(V);
        DevBuf<uint32_t> It(V);
        DevBuf<uint8_t> Cv(V);
        if (V) SF_HIP(hipMemcpyAsync(S.p, signals, sizeof(float) * V * n_rows, hipMemcpyHostToDevice, s));
        FistaLaunch L{};
        FistaWork work;
        L.data = S.p;
        L.vox = nullptr;
        L.n_vox = V;
        L.N = d.N;
        L.sig_stride = d.N;
        L.K = d.K;
        L.Kp = d.Kp;
        L.Np = d.Np;
        L.ratio = 0.f;
        L.alpha = alpha;
        L.max_iter = max_iter;
        L.tol = float(tol);
        L.nonneg = non_negative;
        L.precision = precision;
        L.refresh = refresh_period(refresh_every, precision, d.Kp);
        L.coef = C.p;
        L.resid = R.p;
        L.iters = It.p;
        L.converged = Cv.p;
        setup_fista(d, ops, work, L, s);
        if (lipschitz_out) *lipschitz_out = d.L;
        L.thr = float(ops.step * alpha);
        launch_fista(L, s);
        SF_HIP(hipStreamSynchronize(s));
        if (V) {
            SF_HIP(hipMemcpy(coef, C.p, sizeof(float) * V * n_cols, hipMemcpyDeviceToHost));
            if (iterations) SF_HIP(hipMemcpy(iterations, It.p, sizeof(uint32_t) * V, hipMemcpyDeviceToHost));
            if (converged) SF_HIP(hipMemcpy(converged, Cv.p, V, hipMemcpyDeviceToHost));
            if (residual_l2) SF_HIP(hipMemcpy(residual_l2, R.p, sizeof(float) * V, hipMemcpyDeviceToHost));
        }
    });
}

int shorefit_design_create(const double* A, int32_t n_rows, int32_t n_cols, ShorefitDesign** out) {
    return guarded([&] {
        if (!A || !out || n_rows <= 0 || n_cols <= 0) throw Error{SHOREFIT_E_INVALID, "invalid design arguments"};
        *out = nullptr;
        auto h = std::make_unique<ShorefitDesign>();
        hipStream_t s = nullptr;
        upload_design(h->d, A, n_rows, n_cols, s);
        FistaLaunch L{};
        L.N = h->d.N;
        L.K = h->d.K;
        L.Kp = h->d.Kp;
        L.Np = h->d.Np;
        L.precision = SHOREFIT_PREC_FP32;
        L.max_iter = 1;
        setup_fista(h->d, h->ops, h->work, L, s);
        SF_HIP(hipStreamSynchronize(s));
        *out = h.release();
    });
}

void shorefit_design_destroy(ShorefitDesign* design) { delete design; }

int shorefit_design_info(const ShorefitDesign* h, int32_t* n_rows, int32_t* n_cols, double* lipschitz) {
    return guarded([&] {
        if (!h) throw Error{SHOREFIT_E_INVALID, "null design"};
        if (n_rows) *n_rows = h->d.N;
        if (n_cols) *n_cols = h->d.K;
        if (lipschitz) *lipschitz = h->d.L;
    });
}

}  // extern "C"

namespace {
// The FISTA launch of a lasso forward on a kept design (both entry points);
// the caller supplies the right-hand-side, momentum and counter buffers.
FistaLaunch lasso_launch(const ShorefitDesign* h, const float* signals, int64_t V, const float* alpha,
                         double alpha_scalar, const float* warm, uint32_t max_iter, double tol, int32_t non_negative,
                         float* coef, uint32_t* iterations) {
    const Design& d = h->d;
    FistaLaunch L{};
    L.data = signals;
    L.vox = nullptr;
    L.n_vox = V;
    L.N = d.N;
    L.sig_stride = d.N;
    L.K = d.K;
    L.Kp = d.Kp;
    L.Np = d.Np;
    L.alpha = alpha_scalar;
    L.max_iter = max_iter;
    L.tol = float(tol);
    L.nonneg = non_negative;
    L.precision = SHOREFIT_PREC_FP32;
    L.coef = coef;
    L.iters = iterations;
    L.A64p = h->ops.Ap.p;
    L.c_dev = h->work.sc.p + 1;
    L.Mfrag = h->ops.M.p;
    L.Afrag = h->ops.A.p;
    L.step = float(h->ops.step);
    L.step64 = h->ops.step;
    L.thr = float(h->ops.step * alpha_scalar);
    L.alpha_vox = alpha;
    L.warm = warm;
    // persistent FP32 form: voxels of uneven convergence refill their slots
    L.persist = true;
    return L;
}

void check_lasso_args(const ShorefitDesign* h, const float* signals, int64_t n_signals, float* coef,
                      const char* who) {
    if (!h || (n_signals > 0 && (!signals || !coef)) || n_signals < 0)
        throw Error{SHOREFIT_E_INVALID, std::string("invalid ") + who + " arguments"};
}
}  // namespace

extern "C" {

int shorefit_lasso_forward(const ShorefitDesign* h, const float* signals, int64_t n_signals, const float* alpha,
                           double alpha_scalar, const float* warm, uint32_t max_iter, double tol,
                           int32_t non_negative, float* coef, uint32_t* iterations, void* stream) {
    return guarded([&] {
        check_lasso_args(h, signals, n_signals, coef, "lasso_forward");
        hipStream_t s = as_stream(stream);
        StreamScope scope(s);
        const int64_t V = n_signals;
        FistaLaunch L = lasso_launch(h, signals, V, alpha, alpha_scalar, warm, max_iter, tol, non_negative, coef,
                                     iterations);
        DevBuf<float> B(size_t(V) * h->d.Kp), mean(V);
        L.Bbuf = B.p;
        L.meanbuf = mean.p;
        DevBuf<float> mom;
        upload_momentum(mom, max_iter, s);
        DevBuf<unsigned int> counter(1);
        L.mom = mom.p;
        L.counter = counter.p;
        launch_fista_rhs(L, s);
        launch_fista(L, s);
    });
}

int shorefit_lasso_forward_noalloc(const ShorefitDesign* h, const float* signals, int64_t n_signals,
                                   const float* alpha, double alpha_scalar, const float* warm, uint32_t max_iter,
                                   double tol, int32_t non_negative, float* coef, uint32_t* iterations,
                                   void* stream) {
    return guarded([&] {
        check_lasso_args(h, signals, n_signals, coef, "lasso_forward_noalloc");
        if (n_signals == 0) return;
        const size_t V = size_t(n_signals), Kp = size_t(h->d.Kp);
        const size_t n_mom = std::max<uint32_t>(max_iter, 1);
        auto up = [](size_t bytes) { return (bytes + 255) & ~size_t(255); };
        const size_t off_mom = 256, off_B = off_mom + up(4 * n_mom), off_mean = off_B + up(4 * V * Kp);
        const size_t need = off_mean + up(4 * V);
        if (need > kNoallocScratchBytes) {
            const size_t fixed = off_B, per = 4 * Kp + 4;
            const size_t fit = kNoallocScratchBytes > fixed + 512 ? (kNoallocScratchBytes - fixed - 512) / per : 0;
            throw Error{SHOREFIT_E_INVALID,
                        "lasso_forward_noalloc: " + std::to_string(V) + " signals need " + std::to_string(need) +
                            " bytes of device scratch; the library holds " + std::to_string(kNoallocScratchBytes) +
                            " (at most " + std::to_string(fit) + " signals per call with this design and max_iter)"};
        }
        hipStream_t s = as_stream(stream);
        char* base = static_cast<char*>(noalloc_scratch());
        FistaLaunch L = lasso_launch(h, signals, int64_t(V), alpha, alpha_scalar, warm, max_iter, tol, non_negative,
                                     coef, iterations);
        float* mom = reinterpret_cast<float*>(base + off_mom);
        L.counter = reinterpret_cast<unsigned int*>(base);
        L.counter_zeroed = true;
        L.mom = mom;
        L.Bbuf = reinterpret_cast<float*>(base + off_B);
        L.meanbuf = reinterpret_cast<float*>(base + off_mean);
        launch_noalloc_prepare(mom, uint32_t(n_mom), L.counter, s);
        launch_fista_rhs(L, s);
        launch_fista(L, s);
    });
}

int shorefit_lasso_adjoint(const ShorefitDesign* h, const float* coef, const float* grad_coef, int64_t n_signals,
                           float* grad_signal, float* grad_alpha, uint8_t* status, void* stream) {
    return guarded([&] {
        if (!h || (n_signals > 0 && (!coef || !grad_coef)) || n_signals < 0)
            throw Error{SHOREFIT_E_INVALID, "invalid lasso_adjoint arguments"};
        hipStream_t s = as_stream(stream);
        StreamScope scope(s);
        launch_lasso_adjoint(h->d.G.p, h->d.A.p, h->d.N, h->d.K, coef, grad_coef, n_signals, grad_signal,
                             grad_alpha, status, s);
    });
}

int shorefit_tikhonov_batch(const double* A, int32_t n_rows, int32_t n_cols, const double* d_n, const double* d_l,
                            double lambda_n, double lambda_l, const float* signals, int64_t n_signals, float* coef,
                            float* residual_l2) {
    return guarded([&] {
        if (!A || !signals || !coef || n_rows <= 0 || n_cols <= 0 || n_signals < 0)
            throw Error{SHOREFIT_E_INVALID, "invalid tikhonov_batch arguments"};
        hipStream_t s = nullptr;
        StreamScope scope(s);
        Design d;
        upload_design(d, A, n_rows, n_cols, s);
        d.G.alloc(size_t(d.K) * d.K);
        launch_gram(d.A.p, d.N, d.K, d.G.p, s);
        TikOps ops;
        pack_tikhonov(d, d_n, d_l, lambda_n, lambda_l, ops, s);
        const int64_t V = n_signals;
        DevBuf<float> S(size_t(V) * n_rows), C(size_t(V) * n_cols), R(V);
        if (V) SF_HIP(hipMemcpyAsync(S.p, signals, sizeof(float) * V * n_rows, hipMemcpyHostToDevice, s));
        TikhonovLaunch L{};
        L.data = S.p;
        L.n_vox = V;
        L.N = d.N;
        L.sig_stride = d.N;
        L.K = d.K;
        L.Kp = d.Kp;
        L.Np = d.Np;
        set_tik_ops(L, ops);
        L.coef = C.p;
        L.resid = R.p;
        launch_tikhonov(L, s);
        SF_HIP(hipStreamSynchronize(s));
        if (V) {
            SF_HIP(hipMemcpy(coef, C.p, sizeof(float) * V * n_cols, hipMemcpyDeviceToHost));
            if (residual_l2) SF_HIP(hipMemcpy(residual_l2, R.p, sizeof(float) * V, hipMemcpyDeviceToHost));
        }
    });
}

int shorefit_fit(const ShorefitGradients* grads, const ShorefitVolume* vol, const ShorefitSpec* spec,
                 ShorefitOutput* out, void* stream) {
    return guarded([&] {
        if (!grads || !vol || !spec || !out) throw Error{SHOREFIT_E_INVALID, "null argument"};
        if (!out->coefficients) throw Error{SHOREFIT_E_INVALID, "coefficients output is required"};
        hipStream_t s = as_stream(stream);
        StreamScope scope(s);
        fit_impl(grads, vol, spec, out, s, nullptr);
    });
}

int shorefit_fit_graddev(const ShorefitGradients* grads, const ShorefitVolume* vol, const ShorefitGradDev* gd,
                         const ShorefitSpec* spec, ShorefitOutput* out, void* stream) {
    return guarded([&] {
        if (!grads || !vol || !spec || !out || !gd) throw Error{SHOREFIT_E_INVALID, "null argument"};
        if (!out->coefficients) throw Error{SHOREFIT_E_INVALID, "coefficients output is required"};
        hipStream_t s = as_stream(stream);
        StreamScope scope(s);
        fit_impl(grads, vol, spec, out, s, gd);
    });
}

int shorefit_graddev_variants(uint32_t radial_order, int32_t i, int32_t* kp, int32_t* rows, int32_t* gram_waves,
                              int32_t* fista_waves) {
    const GdVariant* vs = nullptr;
    const int nv = gd_variants(radial_order, &vs);
    if (i >= 0 && i < nv) {
        if (kp) *kp = vs[i].kp;
        if (rows) *rows = vs[i].rc;
        if (gram_waves) *gram_waves = vs[i].gram_waves;
        if (fista_waves) *fista_waves = vs[i].fista_waves;
    }
    return nv;
}

int shorefit_graddev_tune(const ShorefitGradients* grads, const ShorefitVolume* vol, const ShorefitGradDev* gd,
                          const ShorefitSpec* spec, int64_t max_voxels, double* ms_out, void* stream) {
    int best = -1;
    const int rc = guarded([&] {
        if (!grads || !vol || !spec || !gd || !gd->field) throw Error{SHOREFIT_E_INVALID, "null argument"};
        const ShorefitSpec& sp = *spec;
        if (sp.regularization != SHOREFIT_REG_L1 ||
            (sp.alpha_mode != SHOREFIT_ALPHA_FIXED && sp.alpha_mode != SHOREFIT_ALPHA_RATIO))
            throw Error{SHOREFIT_E_UNSUPPORTED,
                        "gradient-deviation fits support regularization l1 with alpha_mode fixed or alpha-ratio"};
        const GdVariant* vs = nullptr;
        const int nv = gd_variants(sp.radial_order, &vs);
        if (nv == 0) throw Error{SHOREFIT_E_UNSUPPORTED, "gradient-deviation fits support even radial orders 2 to 10"};
        check_dims(vol->nx, vol->ny, vol->nz);
        const int64_t N = grads->n_grads;
        if (N <= 0 || !grads->bvals || !grads->bvecs) throw Error{SHOREFIT_E_INVALID, "empty gradient table"};
        double big = grads->big_delta, small = grads->small_delta;
        if (std::isnan(big) || std::isnan(small)) tortoise_deltas(grads->bvals, N, &big, &small);
        hipStream_t s = as_stream(stream);
        StreamScope scope(s);
        const int64_t nvox = vol->nx * vol->ny * vol->nz;
        const bool dev = vol->on_device != 0;
        DevBuf<uint8_t> d_mask;
        DevBuf<float> d_data, d_field;
        const uint8_t* mask = dev_in(vol->mask, size_t(nvox), dev, d_mask, s);
        const float* data = dev_in(vol->data, size_t(nvox) * N, dev, d_data, s);
        const float* field = dev_in(gd->field, size_t(nvox) * 9, dev, d_field, s);
        DevBuf<int64_t> d_vox, d_id, d_dv;
        const int64_t V = voxel_list(mask, nvox, d_vox, s);
        int64_t V_dv = 0;
        gd_partition(field, d_vox.p, V, d_id, d_dv, &V_dv, s);
        const int64_t n = std::min(V_dv, std::max<int64_t>(max_voxels, 1));
        if (n == 0) throw Error{SHOREFIT_E_INVALID, "no deviated voxels to tune on"};
        GdSetup S;
        gd_setup(S, *grads, big, small, *gd, sp.radial_order, sp.zeta, s);
        DevBuf<float> coef(size_t(nvox) * S.B.K);
        GdRunOut o{coef.p, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
        hipEvent_t e0, e1;
        SF_HIP(hipEventCreate(&e0));
        SF_HIP(hipEventCreate(&e1));
        double best_ms = 0.0;
        const GdVariant& dflt = pick_variant(sp.radial_order, N, -1);
        for (int i = -1; i < nv; ++i) {  // i = -1: warm-up with the default
            const GdVariant& v = i < 0 ? dflt : vs[i];
            if (!gd_gram_fits(v, int(N))) {  // too many gradients for this chunk's LDS
                if (ms_out && i >= 0) ms_out[i] = -1.0;
                continue;
            }
            SF_HIP(hipEventRecord(e0, s));
            gd_run(S, v, sp, data, N, field, d_dv.p, n, o, nullptr, s);
            SF_HIP(hipEventRecord(e1, s));
            SF_HIP(hipEventSynchronize(e1));
            float ms = 0.f;
            SF_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (i < 0) continue;
            if (ms_out) ms_out[i] = ms;
            if (best < 0 || ms < best_ms) {
                best = i;
                best_ms = ms;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (best < 0)
            throw Error{SHOREFIT_E_UNSUPPORTED, "gradient-deviation fit: too many gradients for LDS"};
        std::lock_guard<std::mutex> lk(tuned_mutex());
        tuned_variants()[{sp.radial_order, N}] = best;
    });
    return rc == SHOREFIT_OK ? best : -rc;
}

int shorefit_graddev_gram(const ShorefitGradients* grads, const ShorefitVolume* vol, const ShorefitGradDev* gd,
                          uint32_t radial_order, double zeta, const int64_t* voxels, int64_t n_voxels,
                          float* gram_out, double* rhs_out, double* lipschitz_out) {
    return guarded([&] {
        if (!grads || !vol || !gd || !gd->field || (!voxels && n_voxels > 0))
            throw Error{SHOREFIT_E_INVALID, "null argument"};
        if (vol->on_device) throw Error{SHOREFIT_E_INVALID, "shorefit_graddev_gram takes host memory"};
        check_dims(vol->nx, vol->ny, vol->nz);
        const GdVariant& v = pick_variant(radial_order, grads->n_grads, -1);
        const int64_t N = grads->n_grads;
        if (N <= 0 || !grads->bvals || !grads->bvecs) throw Error{SHOREFIT_E_INVALID, "empty gradient table"};
        const int64_t nvox = vol->nx * vol->ny * vol->nz;
        for (int64_t i = 0; i < n_voxels; ++i)
            if (voxels[i] < 0 || voxels[i] >= nvox) throw Error{SHOREFIT_E_INVALID, "voxel index out of range"};
        if (n_voxels == 0) return;
        double big = grads->big_delta, small = grads->small_delta;
        if (std::isnan(big) || std::isnan(small)) tortoise_deltas(grads->bvals, N, &big, &small);
        hipStream_t s = nullptr;
        StreamScope scope(s);
        DevBuf<float> d_data, d_field;
        const float* data = dev_in(vol->data, size_t(nvox) * N, false, d_data, s);
        const float* field = dev_in(gd->field, size_t(nvox) * 9, false, d_field, s);
        DevBuf<int64_t> d_vox(n_voxels);
        SF_HIP(hipMemcpyAsync(d_vox.p, voxels, sizeof(int64_t) * n_voxels, hipMemcpyHostToDevice, s));
        GdSetup S;
        gd_setup(S, *grads, big, small, *gd, radial_order, zeta, s);
        const int KP = v.kp, K = S.B.K;
        DevBuf<float> G(size_t(n_voxels) * KP * KP);
        DevBuf<double> b(size_t(n_voxels) * KP), sstat(size_t(n_voxels) * 2);
        GdChunk c{S.B, S.tab, data, N, field, d_vox.p, n_voxels, G.p, b.p, sstat.p};
        gd_launch_gram(v, c, s);
        std::vector<double> lip(n_voxels);
        if (lipschitz_out) {
            // the FISTA kernel's set-up (Lanczos) with no iterations
            DevBuf<double> d_lip(n_voxels);
            DevBuf<float> coef(size_t(nvox) * K), mom(1);
            DevBuf<GdSlotOut> so(n_voxels);
            DevBuf<unsigned> ctr(1);
            SF_HIP(hipMemsetAsync(ctr.p, 0, sizeof(unsigned), s));
            GdFistaArgs fa{};
            fa.G = G.p;
            fa.b = b.p;
            fa.sstat = sstat.p;
            fa.vox = d_vox.p;
            fa.n = n_voxels;
            fa.N = int(N);
            fa.K = K;
            fa.alpha = 1.0;
            fa.max_iter = 0;
            fa.tol = 1e-6;
            fa.mom = mom.p;
            fa.counter = ctr.p;
            fa.coef = coef.p;
            fa.slot_out = so.p;
            fa.lip = d_lip.p;
            gd_launch_fista(v, fa, 0, s);
            SF_HIP(hipMemcpyAsync(lip.data(), d_lip.p, sizeof(double) * n_voxels, hipMemcpyDeviceToHost, s));
        }
        std::vector<float> Gh(gram_out ? size_t(n_voxels) * KP * KP : 0);
        std::vector<double> bh(rhs_out ? size_t(n_voxels) * KP : 0);
        if (gram_out) SF_HIP(hipMemcpyAsync(Gh.data(), G.p, sizeof(float) * Gh.size(), hipMemcpyDeviceToHost, s));
        if (rhs_out) SF_HIP(hipMemcpyAsync(bh.data(), b.p, sizeof(double) * bh.size(), hipMemcpyDeviceToHost, s));
        SF_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < n_voxels; ++i) {
            if (gram_out)
                for (int r = 0; r < K; ++r)
                    std::memcpy(gram_out + (size_t(i) * K + r) * K, Gh.data() + (size_t(i) * KP + r) * KP,
                                sizeof(float) * K);
            if (rhs_out) std::memcpy(rhs_out + size_t(i) * K, bh.data() + size_t(i) * KP, sizeof(double) * K);
            if (lipschitz_out) lipschitz_out[i] = lip[i];
        }
    });
}

// --------------------------------------------------------- post-fit maps ----

int shorefit_tournier_matrix(uint32_t radial_order, double zeta, uint32_t lmax, double* out) {
    return guarded([&] {
        if (!out) throw Error{SHOREFIT_E_INVALID, "null argument"};
        if (lmax % 2 != 0 || lmax > radial_order)
            throw Error{SHOREFIT_E_INVALID, "lmax must be even and ≤ radial_order (" + std::to_string(radial_order) +
                                                "), got " + std::to_string(lmax)};
        tournier_matrix_host(radial_order, zeta, lmax, out);
    });
}


int shorefit_odf_sh(const float* coefficients, const uint8_t* mask, int64_t nx, int64_t ny, int64_t nz,
                    uint32_t radial_order, double zeta, uint32_t lmax, float* out, int32_t on_device, void* stream) {
    return guarded([&] {
        if (!coefficients || !out) throw Error{SHOREFIT_E_INVALID, "null argument"};
        check_dims(nx, ny, nz);
        if (lmax % 2 != 0 || lmax > radial_order)
            throw Error{SHOREFIT_E_INVALID, "lmax must be even and ≤ radial_order (" + std::to_string(radial_order) +
                                                "), got " + std::to_string(lmax)};
        hipStream_t s = as_stream(stream);
        StreamScope scope(s);
        const int K = n_coeffs(radial_order), R = n_tournier(lmax);
        const int64_t nv = nx * ny * nz;
        std::vector<double> M(size_t(R) * K);
        tournier_matrix_host(radial_order, zeta, lmax, M.data());
        DevBuf<double> dM(M.size());
        SF_HIP(hipMemcpyAsync(dM.p, M.data(), sizeof(double) * M.size(), hipMemcpyHostToDevice, s));
        const bool dev = on_device != 0;
        DevBuf<float> bc, bo;
        DevBuf<uint8_t> bm;
        const float* c = dev_in(coefficients, size_t(nv) * K, dev, bc, s);
        const uint8_t* m = dev_in(mask, size_t(nv), dev, bm, s);
        float* o = out;
        if (!dev) {
            bo.alloc(size_t(nv) * R);
            o = bo.p;
        }
        if (nv) SF_HIP(hipMemsetAsync(o, 0, sizeof(float) * size_t(nv) * R, s));
        DevBuf<int64_t> vox;
        const int64_t V = voxel_list(m, nv, vox, s);
        launch_apply_rows(c, vox.p, V, K, dM.p, R, false, o, s);
        if (!dev && nv) SF_HIP(hipMemcpyAsync(out, o, sizeof(float) * size_t(nv) * R, hipMemcpyDeviceToHost, s));
        SF_HIP(hipStreamSynchronize(s));
    });
}

int shorefit_predict(const float* coefficients, int64_t nx, int64_t ny, int64_t nz, uint32_t radial_order,
                     double zeta, const ShorefitGradients* target, float* out, int32_t on_device, void* stream) {
    return guarded([&] {
        if (!coefficients || !out || !target) throw Error{SHOREFIT_E_INVALID, "null argument"};
        check_dims(nx, ny, nz);
        hipStream_t s = as_stream(stream);
        StreamScope scope(s);
        double big = target->big_delta, small = target->small_delta;
        if (std::isnan(big) || std::isnan(small)) tortoise_deltas(target->bvals, target->n_grads, &big, &small);
        Design d;
        build_design(d, *target, big, small, radial_order, zeta, s);
        const int K = d.K, Nt = d.N;
        const int64_t nv = nx * ny * nz;
        const bool dev = on_device != 0;
        DevBuf<float> bc, bo;
        const float* c = dev_in(coefficients, size_t(nv) * K, dev, bc, s);
        float* o = out;
        if (!dev) {
            bo.alloc(size_t(nv) * Nt);
            o = bo.p;
        }
        launch_apply_rows(c, nullptr, nv, K, d.A.p, Nt, true, o, s);
        if (!dev && nv) SF_HIP(hipMemcpyAsync(out, o, sizeof(float) * size_t(nv) * Nt, hipMemcpyDeviceToHost, s));
        SF_HIP(hipStreamSynchronize(s));
    });
}

namespace {
size_t nifti_element_size(int32_t datatype) {
    switch (datatype) {
        case 2: case 256: return 1;
        case 4: case 512: return 2;
        case 8: case 16: case 768: return 4;
        case 64: return 8;
        default:
            throw Error{SHOREFIT_E_UNSUPPORTED, "NIfTI datatype " + std::to_string(datatype) +
                                                    " is not supported (uint8, int8, int16, uint16, int32, uint32, "
                                                    "float32, float64 are)"};
    }
}
}  // namespace

int shorefit_volume_to_voxel_major(const void* raw, int32_t datatype, int64_t nx, int64_t ny, int64_t nz, int64_t nt,
                                   float scl_slope, float scl_inter, float* out, int32_t on_device, void* stream) {
    return guarded([&] {
        if (!raw || !out) throw Error{SHOREFIT_E_INVALID, "null argument"};
        check_dims(nx, ny, nz);
        if (nt < 0) throw Error{SHOREFIT_E_INVALID, "negative volume dimension"};
        const size_t es = nifti_element_size(datatype);
        const size_t n = size_t(nx) * ny * nz * nt;
        hipStream_t s = as_stream(stream);
        StreamScope scope(s);
        const bool dev = on_device != 0;
        DevBuf<uint8_t> braw;
        DevBuf<float> bout;
        const void* r = raw;
        float* o = out;
        if (!dev) {
            braw.alloc(n * es);
            if (n) SF_HIP(hipMemcpyAsync(braw.p, raw, n * es, hipMemcpyHostToDevice, s));
            r = braw.p;
            bout.alloc(n);
            o = bout.p;
        }
        launch_to_voxel_major(r, datatype, nx, ny, nz, nt, scl_slope, scl_inter, o, s);
        if (!dev && n) SF_HIP(hipMemcpyAsync(out, o, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        SF_HIP(hipStreamSynchronize(s));
    });
}

int shorefit_volume_to_file_order(const float* in, int64_t nx, int64_t ny, int64_t nz, int64_t nt, float* out,
                                  int32_t on_device, void* stream) {
    return guarded([&] {
        if (!in || !out) throw Error{SHOREFIT_E_INVALID, "null argument"};
        check_dims(nx, ny, nz);
        if (nt < 0) throw Error{SHOREFIT_E_INVALID, "negative volume dimension"};
        const size_t n = size_t(nx) * ny * nz * nt;
        hipStream_t s = as_stream(stream);
        StreamScope scope(s);
        const bool dev = on_device != 0;
        DevBuf<float> bin, bout;
        const float* i = dev_in(in, n, dev, bin, s);
        float* o = out;
        if (!dev) {
            bout.alloc(n);
            o = bout.p;
        }
        launch_to_file_order(i, nx, ny, nz, nt, o, s);
        if (!dev && n) SF_HIP(hipMemcpyAsync(out, o, sizeof(float) * n, hipMemcpyDeviceToHost, s));
        SF_HIP(hipStreamSynchronize(s));
    });
}

int shorefit_microstructure(const float* coefficients, const uint8_t* mask, int64_t nx, int64_t ny, int64_t nz,
                            uint32_t radial_order, double zeta, const float* directions, int32_t units,
                            float outlier_factor, float* rtop, float* rtap, float* rtpp, float* msd, float* qiv,
                            float* ng, int32_t on_device, void* stream) {
    return guarded([&] {
        if (!coefficients || !mask) throw Error{SHOREFIT_E_INVALID, "null argument"};
        check_dims(nx, ny, nz);
        if (units != SHOREFIT_UNITS_UM && units != SHOREFIT_UNITS_MM)
            throw Error{SHOREFIT_E_INVALID, "units must be 'um' or 'mm'"};
        if (radial_order > 14) throw Error{SHOREFIT_E_UNSUPPORTED, "radial order above 14"};
        hipStream_t s = as_stream(stream);
        StreamScope scope(s);
        const int K = n_coeffs(radial_order);
        const int64_t nv = nx * ny * nz;
        const bool dev = on_device != 0;
        MicroTables T;
        micro_tables_host(radial_order, zeta, units, T);
        DevBuf<float> bc, bd;
        DevBuf<uint8_t> bm;
        const float* c = dev_in(coefficients, size_t(nv) * K, dev, bc, s);
        const uint8_t* m = dev_in(mask, size_t(nv), dev, bm, s);
        const float* dirs = dev_in(directions, size_t(nv) * 3, dev, bd, s);
        DevBuf<int64_t> vox;
        const int64_t V = voxel_list(m, nv, vox, s);
        DevBuf<float> raw(size_t(6) * std::max<int64_t>(V, 1));
        launch_micro(c, vox.p, V, K, T, dirs, raw.p, s);
        // src/io/microstructure.rs:215-268, 303-326: threshold = factor x p99
        // of the finite values (index round((n - 1) 0.99) of the sorted list),
        // selected on the device
        DevBuf<MicroThresholds> th(1);
        SF_HIP(hipMemsetAsync(th.p, 0, sizeof(MicroThresholds), s));
        if (outlier_factor > 0.f && V > 0) launch_micro_thresholds(raw.p, V, outlier_factor, th.p, s);
        float* user[6] = {rtop, rtap, rtpp, msd, qiv, ng};
        DevBuf<float> bo[6];
        float* maps[6];
        for (int q = 0; q < 6; ++q) {
            maps[q] = user[q];
            if (user[q] && !dev) {
                bo[q].alloc(std::max<int64_t>(nv, 1));
                maps[q] = bo[q].p;
            }
            if (maps[q] && nv) SF_HIP(hipMemsetAsync(maps[q], 0, sizeof(float) * nv, s));
        }
        launch_micro_scatter(raw.p, vox.p, V, th.p, maps, s);
        if (!dev)
            for (int q = 0; q < 6; ++q)
                if (user[q] && nv)
                    SF_HIP(hipMemcpyAsync(user[q], maps[q], sizeof(float) * nv, hipMemcpyDeviceToHost, s));
        SF_HIP(hipStreamSynchronize(s));
    });
}

}  // extern "C"
